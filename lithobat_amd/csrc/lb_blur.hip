// Separable Gaussian blur with the semantics of the reference's PEB / ILT
// blur (opc/ilt.py:274-289 -> scipy.ndimage.gaussian_filter, default mode
// "reflect" and truncate 4.0): taps w_k = exp(-k^2 / (2 sigma^2)) / sum for
// |k| <= R = int(4 sigma + 0.5), correlated along each axis, half-sample
// symmetric extension (d c b a | a b c d | d c b a) at the borders.
//
// One output per thread, neighbours read from HBM / L2 (each input is reused
// by the 2R+1 outputs around it, so the traffic stays on chip).  The weights
// are computed on the device, in FP64, so the call allocates nothing and can
// be captured into a graph.
#include "lb_kernels.h"

namespace lb {

// half-sample symmetric index of i on [0, n): period 2n, mirrored halves
__device__ __forceinline__ int reflect_idx(int i, int n) {
  const int p = 2 * n;
  int r = i % p;
  if (r < 0) r += p;
  return r < n ? r : p - 1 - r;
}

// w[0..2R] (FP64), normalised; one block
__global__ void k_gauss_weights(double sigma, int R, double* __restrict__ w) {
  __shared__ double part[256];
  double s = 0;
  for (int k = threadIdx.x; k <= 2 * R; k += blockDim.x) {
    const double x = (double)(k - R);
    const double v = exp(-0.5 / (sigma * sigma) * x * x);
    w[k] = v;
    s += v;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  // deterministic tree sum
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  const double inv = 1.0 / part[0];
  for (int k = threadIdx.x; k <= 2 * R; k += blockDim.x) w[k] *= inv;
}

// AXIS 0: along y (rows of the image index), AXIS 1: along x.
template <typename T, int AXIS>
__global__ void k_gauss_pass(const T* __restrict__ in, T* __restrict__ out, int n, int R,
                             const double* __restrict__ w) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  const long long img = blockIdx.z;
  if (x >= n) return;
  const T* src = in + img * n * n;
  T acc = 0;
  for (int k = -R; k <= R; ++k) {
    const T wk = (T)w[k + R];
    const T v = AXIS == 0 ? src[(long long)reflect_idx(y + k, n) * n + x]
                          : src[(long long)y * n + reflect_idx(x + k, n)];
    acc += wk * v;
  }
  out[img * n * n + (long long)y * n + x] = acc;
}

int gauss_radius(double sigma) { return (int)(4.0 * sigma + 0.5); }

template <typename T>
hipError_t gaussian_blur(const T* in, T* out, int n_img, int n, double sigma, T* tmp, double* w,
                         hipStream_t st) {
  const int R = gauss_radius(sigma);
  hipLaunchKernelGGL(k_gauss_weights, dim3(1), dim3(256), 0, st, sigma, R, w);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const dim3 grid((n + 255) / 256, n, n_img);
  hipLaunchKernelGGL((k_gauss_pass<T, 0>), grid, dim3(256), 0, st, in, tmp, n, R, w);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL((k_gauss_pass<T, 1>), grid, dim3(256), 0, st, tmp, out, n, R, w);
  return hipGetLastError();
}

template hipError_t gaussian_blur<float>(const float*, float*, int, int, double, float*, double*,
                                         hipStream_t);
template hipError_t gaussian_blur<double>(const double*, double*, int, int, double, double*,
                                          double*, hipStream_t);

}  // namespace lb
