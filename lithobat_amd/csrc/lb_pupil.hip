// Transfer tables of the coherent systems, computed in FP64 on the device.
//
// For system j and box position (a, b) the frequency index is
// (box_y + a, box_x + b), evaluated at its canonical FFT bin frequency
// nu * (1 / (N * pixel)) exactly as numpy.fft.fftfreq forms it
// (core/utils.py:64-66).  The in-pupil test repeats the reference's float64
// operation order with no fused multiply-add and correctly rounded sqrt and
// division (expose/aerial_image.py:131-138), so the pupil support is the
// reference's bit for bit.  Inside the pupil:
//
//   W   = defocus OPD (paraxial or exact, expose/pupil.py:155-221)
//       + sum_j 2 pi c_j Z_j(rho, phi)            (expose/pupil.py:81-141)
//   H   = weight * exp(i W)                         scalar model
//   H   = exp(i W) * V_c(rho, phi; Jones)           vector model, component c
//                                                   (expose/pupil.py:292-369)
// and H = 0 outside.  dc[j] receives the open-frame contribution of system j
// (|H(0)|^2 / weight-free, as expose/aerial_image.py:464,482-486 book it).
#include "lb_kernels.h"

namespace lb {

__device__ __forceinline__ int canonical_nu(int nu, int n) {
  int i = nu % n;
  if (i < 0) i += n;
  return i <= (n - 1) / 2 ? i : i - n;
}

__device__ double rho_exact(double FX, double FY, double fsx, double fsy, double k_max) {
#pragma clang fp contract(off)
  const double fx = FX - fsx;
  const double fy = FY - fsy;
  const double r2 = fx * fx + fy * fy;
  return __dsqrt_rn(r2) / k_max;
}

__device__ double defocus_phase(double rho, const PupilParams& pp, double dz, double n_img) {
  if (dz == 0.0) return 0.0;
  if (!pp.exact) {
    return M_PI * (pp.NA * pp.NA) * (rho * rho) * dz / (pp.wavelength * n_img);
  }
  double u = rho * pp.NA / n_img;
  double sin_sq = fmin(fmax(u * u, 0.0), 1.0);
  double cos_t = sqrt(1.0 - sin_sq);
  return (2.0 * M_PI / pp.wavelength) * (n_img * dz * (1.0 - cos_t));
}

__device__ double zernike_sum(double rho, double phi, const PupilParams& pp) {
  double W = 0.0;
  const double rc = fmin(fmax(rho, 0.0), 1.0);
  for (int t = 0; t < pp.n_zern; ++t) {
    const int n = pp.zn[t], m = pp.zm[t];
    const int ma = m < 0 ? -m : m;
    double R = 0.0;
    for (int s = 0; s <= (n - ma) / 2; ++s) {
      double p = 1.0;
      for (int e = 0; e < n - 2 * s; ++e) p *= rc;
      R += pp.zrad[t][s] * p;
    }
    double Z = pp.znorm[t] * R;
    if (m > 0) Z *= cos((double)m * phi);
    if (m < 0) Z *= sin((double)ma * phi);
    W += 2.0 * M_PI * pp.zcoef[t] * Z;
  }
  return W;
}

// Linear interpolation of (xp, fp) at x as numpy.interp defines it: the end
// values outside [xp[0], xp[n-1]], fp[j] at a table point, else
// slope * (x - xp[j]) + fp[j] on the bracketing interval.
__device__ double interp_table(double x, const double* xp, const double* fp, int n) {
  if (x < xp[0]) return fp[0];
  if (x >= xp[n - 1]) return fp[n - 1];
  int lo = 0, hi = n - 1;  // xp[lo] <= x < xp[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xp[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  if (xp[lo] == x) return fp[lo];
  const double slope = (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
  return slope * (x - xp[lo]) + fp[lo];
}

// The blank's amplitude reflection for an order leaving the mask at direction
// sine st (provider.py:252-266): zero for an evanescent order (st > 1), else
// |r| and arg r interpolated at asin(st).
__device__ void blank_reflection(double st, const PupilParams& pp, double* rr, double* ri) {
  *rr = 0.0;
  *ri = 0.0;
  if (!(st <= 1.0)) return;
  const double th = asin(fmin(fmax(st, 0.0), 1.0));
  const double mag = interp_table(th, pp.refl_theta, pp.refl_mag, pp.n_refl);
  const double ph = interp_table(th, pp.refl_theta, pp.refl_phase, pp.n_refl);
  double sp, cp;
  sincos(ph, &sp, &cp);
  *rr = mag * cp;
  *ri = mag * sp;
}

template <typename T>
__global__ void k_pupil(const PupilSys* __restrict__ sys, int n_sys, PupilParams pp,
                        cx<T>* __restrict__ H, double* __restrict__ dc) {
  const int j = blockIdx.y;
  const int WW = pp.W * pp.W;
  const PupilSys s = sys[j];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < WW; e += gridDim.x * blockDim.x) {
    const int a = e / pp.W, b = e - a * pp.W;
    const int nuy = canonical_nu(s.box_y + a, pp.N);
    const int nux = canonical_nu(s.box_x + b, pp.N);
    const double FX = (double)nux * pp.val;
    const double FY = (double)nuy * pp.val;
    const double rho = rho_exact(FX, FY, s.fsx, s.fsy, pp.k_max);
    cx<T> h = cx<T>{0, 0};
    double dcv = 0.0;
    if (rho <= 1.0) {
      const double fx = FX - s.fsx, fy = FY - s.fsy;
      const bool need_phi = s.comp >= 0 || pp.n_zern > 0;
      const double phi = need_phi ? atan2(fy, fx) : 0.0;
      double W = defocus_phase(rho, pp, s.defocus, s.n_image);
      if (pp.n_zern > 0) W += zernike_sum(rho, phi, pp);
      double sw, cw;
      sincos(W, &sw, &cw);
      double hr, hi;
      if (s.comp < 0) {
        hr = s.weight * cw;
        hi = s.weight * sw;
        dcv = 1.0;
      } else {
        const double stm = pp.NA / s.n_image;
        const double u = rho * stm;
        const double sin_sq = fmin(fmax(u * u, 0.0), 1.0);
        const double cos_t = sqrt(1.0 - sin_sq);
        const double sin_t = sqrt(sin_sq);
        double sp, cp;
        sincos(phi, &sp, &cp);
        const double e_s = -s.jx * sp + s.jy * cp;
        const double e_p = s.jx * cp + s.jy * sp;
        double v;
        if (s.comp == 0) v = -e_s * sp + e_p * cos_t * cp;
        else if (s.comp == 1) v = e_s * cp + e_p * cos_t * sp;
        else v = -e_p * sin_t;
        if (pp.obliquity) v *= sqrt(cos_t);
        hr = s.weight * cw * v;
        hi = s.weight * sw * v;
        dcv = v * v;
      }
      if (pp.n_refl > 0) {
        // reflective mask: this order leaves the mask at its own angle
        const double sx = s.sin_x + pp.refl_doff * FX, sy = s.sin_y + pp.refl_doff * FY;
        double rr, ri;
        blank_reflection(sqrt(sx * sx + sy * sy), pp, &rr, &ri);
        const double nr = hr * rr - hi * ri, ni = hr * ri + hi * rr;
        hr = nr;
        hi = ni;
        dcv *= rr * rr + ri * ri;  // the clear field of this source point (specular order)
      }
      h = cx<T>{(T)hr, (T)hi};
    }
    H[(long long)j * WW + e] = h;
    if (nux == 0 && nuy == 0) dc[j] = dcv;
  }
}

template <typename T>
hipError_t pupil_tables(const PupilSys* sys_dev, int n_sys, const PupilParams& pp, cx<T>* H,
                        double* dc, hipStream_t st) {
  if (n_sys <= 0) return hipSuccess;
  const int WW = pp.W * pp.W;
  dim3 g((WW + 255) / 256, n_sys);
  hipLaunchKernelGGL(k_pupil<T>, g, dim3(256), 0, st, sys_dev, n_sys, pp, H, dc);
  return hipGetLastError();
}

template hipError_t pupil_tables<float>(const PupilSys*, int, const PupilParams&, cx<float>*,
                                        double*, hipStream_t);
template hipError_t pupil_tables<double>(const PupilSys*, int, const PupilParams&, cx<double>*,
                                         double*, hipStream_t);

}  // namespace lb
