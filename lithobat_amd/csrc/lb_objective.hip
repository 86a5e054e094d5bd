// The inverse-lithography objective around the imaging adjoint, fused into
// single passes over the images (opc/ilt.py:311-337, 368-389):
//
//   relaxed resist   s = sign (f - level), f = scale * I (or its PEB blur),
//                    z = sigmoid(clip(a s, +-60)), a = steepness / |level|,
//                    cost_b = sum_pixels (z - t)^2 / npix,
//                    d f   = sign (2 / npix) (z - t) a z (1 - z),
//                    mismatch_b = #{(s > 0) != (t > 0.5)}
//   mask             m = sigmoid(clip(beta x, +-60)),  d x = d m beta m (1 - m)
//
// The derivatives are the reference's closed forms (they ignore the clip,
// whose effect is below e^-60).  Costs are reduced in two deterministic
// levels (fixed block partials, then a fixed-order sum per image), so the
// same inputs give bit-identical costs, and nothing here is a memset, so the
// chain can be captured into a graph.
#include <cstdint>

#include "lb_kernels.h"

namespace lb {

constexpr int kObjThreads = 256;

// Elementwise passes: each thread handles 16-byte vectors (4 floats or 2
// doubles) when every pointer is 16-byte aligned (VEC), then a scalar tail.
template <typename T>
struct Vec16 {
  static constexpr int K = 16 / sizeof(T);
  T v[K];
};

// Block (blockIdx.x, image blockIdx.y) reduces its share of one image into
// part_cost / part_err [image][block]; d_f and signed_f are optional.  VEC:
// the pixels go through in 16-byte vectors (every pointer 16-byte aligned
// and npix a whole number of vectors).
template <typename T, bool VEC>
__global__ __launch_bounds__(kObjThreads) void k_resist_partial(
    const T* __restrict__ f, const T* __restrict__ target, long long npix, T scale, T level, T a,
    T sign, T* __restrict__ d_f, T* __restrict__ signed_f, double* __restrict__ part_cost,
    int* __restrict__ part_err) {
  const long long base = (long long)blockIdx.y * npix;
  const T dscale = sign * (T)2 / (T)npix * a * scale;
  double c = 0;
  int err = 0;
  auto pixel = [&](T fv, T t, T& d, T& sv) {
    const T s = sign * (fv * scale - level);
    const T z = sigmoid_clipped(a * s);
    const T r = z - t;
    c += (double)(r * r);
    err += (s > (T)0) != (t > (T)0.5);
    d = dscale * r * z * ((T)1 - z);
    sv = s;
  };
  if constexpr (VEC) {
    using V = Vec16<T>;
    const long long nv = npix / V::K;
    const V* fv = reinterpret_cast<const V*>(f + base);
    const V* tv = reinterpret_cast<const V*>(target + base);
    for (long long i = blockIdx.x * (long long)kObjThreads + threadIdx.x; i < nv;
         i += (long long)gridDim.x * kObjThreads) {
      const V x = fv[i], t = tv[i];
      V d, sv;
#pragma unroll
      for (int k = 0; k < V::K; ++k) pixel(x.v[k], t.v[k], d.v[k], sv.v[k]);
      if (d_f) reinterpret_cast<V*>(d_f + base)[i] = d;
      if (signed_f) reinterpret_cast<V*>(signed_f + base)[i] = sv;
    }
  } else {
    for (long long e = blockIdx.x * (long long)kObjThreads + threadIdx.x; e < npix;
         e += (long long)gridDim.x * kObjThreads) {
      T d, sv;
      pixel(f[base + e], target[base + e], d, sv);
      if (d_f) d_f[base + e] = d;
      if (signed_f) signed_f[base + e] = sv;
    }
  }
  __shared__ double sc[kObjThreads / 64];
  __shared__ int se[kObjThreads / 64];
  for (int o = 32; o > 0; o >>= 1) {
    c += __shfl_xor(c, o, 64);
    err += __shfl_xor(err, o, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sc[wave] = c;
    se[wave] = err;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double cs = 0;
    int es = 0;
    for (int w = 0; w < kObjThreads / 64; ++w) {
      cs += sc[w];
      es += se[w];
    }
    const long long k = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    part_cost[k] = cs;
    part_err[k] = es;
  }
}

// One block of kFinalWaves wavefronts: wavefront w sums the block partials of
// images w, w + kFinalWaves, ... (lane-strided, then a fixed shuffle tree);
// thread 0 then adds the per-image costs in image order for the total.
constexpr int kFinalWaves = 16;

template <typename T>
__global__ __launch_bounds__(64 * kFinalWaves) void k_resist_final(
    const double* __restrict__ part_cost, const int* __restrict__ part_err, int n_img, int nblk,
    long long npix, double* __restrict__ img_cost, T* __restrict__ cost, T* __restrict__ total,
    int* __restrict__ mismatch) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int b = wave; b < n_img; b += kFinalWaves) {
    double c = 0;
    int e = 0;
    for (int k = lane; k < nblk; k += 64) {
      c += part_cost[(long long)b * nblk + k];
      e += part_err[(long long)b * nblk + k];
    }
    for (int o = 32; o > 0; o >>= 1) {
      c += __shfl_xor(c, o, 64);
      e += __shfl_xor(e, o, 64);
    }
    if (lane == 0) {
      const double cb = c / (double)npix;
      img_cost[b] = cb;
      if (cost) cost[b] = (T)cb;
      if (mismatch) mismatch[b] = e;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && total) {
    double tot = 0;
    for (int b = 0; b < n_img; ++b) tot += img_cost[b];
    *total = (T)tot;
  }
}

int resist_blocks(long long npix) {
  long long b = (npix + 16LL * kObjThreads - 1) / (16LL * kObjThreads);  // ~16 pixels a thread
  if (b < 1) b = 1;
  if (b > 1024) b = 1024;
  return (int)b;
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// part_cost holds n_img * (nblk + 1) doubles: the block partials, then the
// per-image costs.
template <typename T>
hipError_t resist_objective(const T* f, const T* target, int n_img, long long npix, T scale,
                            T level, T a, T sign, T* cost, T* total, T* d_f, T* signed_f,
                            int* mismatch, double* part_cost, int* part_err, hipStream_t st) {
  const int nblk = resist_blocks(npix);
  const bool vec = npix % (16 / (long long)sizeof(T)) == 0 && aligned16(f) && aligned16(target) &&
                   (!d_f || aligned16(d_f)) && (!signed_f || aligned16(signed_f));
  auto k = vec ? k_resist_partial<T, true> : k_resist_partial<T, false>;
  hipLaunchKernelGGL(k, dim3(nblk, n_img), dim3(kObjThreads), 0, st, f, target, npix, scale,
                     level, a, sign, d_f, signed_f, part_cost, part_err);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return resist_final<T>(part_cost, part_err, n_img, nblk, npix,
                         part_cost + (long long)n_img * nblk, cost, total, mismatch, st);
}

template <typename T>
hipError_t resist_final(const double* part_cost, const int* part_err, int n_img, int nblk,
                        long long npix, double* img_cost, T* cost, T* total, int* mismatch,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_resist_final<T>, dim3(1), dim3(64 * kFinalWaves), 0, st, part_cost,
                     part_err, n_img, nblk, npix, img_cost, cost, total, mismatch);
  return hipGetLastError();
}
template hipError_t resist_final<float>(const double*, const int*, int, int, long long, double*,
                                        float*, float*, int*, hipStream_t);
template hipError_t resist_final<double>(const double*, const int*, int, int, long long, double*,
                                         double*, double*, int*, hipStream_t);

// pbeta (device, optional): beta read from device memory instead of the
// argument, so a captured graph can follow a schedule that changes per call.
template <typename T, bool VEC>
__global__ void k_mask_sigmoid(const T* __restrict__ x, long long n, T beta_arg,
                               const double* __restrict__ pbeta, T* __restrict__ m) {
  const T beta = pbeta ? (T)*pbeta : beta_arg;
  using V = Vec16<T>;
  const long long nv = VEC ? n / V::K : 0;
  const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long i = t0; i < nv; i += st) {
    V a = reinterpret_cast<const V*>(x)[i];
#pragma unroll
    for (int k = 0; k < V::K; ++k) a.v[k] = sigmoid_clipped(beta * a.v[k]);
    reinterpret_cast<V*>(m)[i] = a;
  }
  for (long long i = nv * V::K + t0; i < n; i += st) m[i] = sigmoid_clipped(beta * x[i]);
}

template <typename T, bool VEC>
__global__ void k_mask_sigmoid_grad(const T* __restrict__ gm, const T* __restrict__ m, long long n,
                                    T beta_arg, const double* __restrict__ pbeta,
                                    T* __restrict__ gx) {
  const T beta = pbeta ? (T)*pbeta : beta_arg;
  using V = Vec16<T>;
  const long long nv = VEC ? n / V::K : 0;
  const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long i = t0; i < nv; i += st) {
    V g = reinterpret_cast<const V*>(gm)[i];
    const V a = reinterpret_cast<const V*>(m)[i];
#pragma unroll
    for (int k = 0; k < V::K; ++k) g.v[k] *= beta * a.v[k] * ((T)1 - a.v[k]);
    reinterpret_cast<V*>(gx)[i] = g;
  }
  for (long long i = nv * V::K + t0; i < n; i += st) {
    const T v = m[i];
    gx[i] = gm[i] * (beta * v * ((T)1 - v));
  }
}

// d theta = g * d m * beta m (1 - m) with m = sigmoid(clip(beta theta, +-60)) recomputed
// from theta, g the incoming gradient of a scalar loss read from device memory (null: 1):
// the mask array need not be kept, and the loss gradient scales here instead of in a pass
// over the intensity derivative (the adjoint between them is linear).
template <typename T, bool VEC>
__global__ void k_mask_sigmoid_grad_theta(const T* __restrict__ gm, const T* __restrict__ x,
                                          long long n, T beta_arg, const double* __restrict__ pbeta,
                                          const T* __restrict__ gscale, T* __restrict__ gx) {
  const T beta = pbeta ? (T)*pbeta : beta_arg;
  const T gs = gscale ? *gscale : (T)1;
  using V = Vec16<T>;
  const long long nv = VEC ? n / V::K : 0;
  const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long i = t0; i < nv; i += st) {
    V g = reinterpret_cast<const V*>(gm)[i];
    const V a = reinterpret_cast<const V*>(x)[i];
#pragma unroll
    for (int k = 0; k < V::K; ++k) {
      const T m = sigmoid_clipped(beta * a.v[k]);
      g.v[k] *= beta * m * ((T)1 - m) * gs;
    }
    reinterpret_cast<V*>(gx)[i] = g;
  }
  for (long long i = nv * V::K + t0; i < n; i += st) {
    const T m = sigmoid_clipped(beta * x[i]);
    gx[i] = gm[i] * (beta * m * ((T)1 - m) * gs);
  }
}

static unsigned elementwise_blocks(long long n) {
  long long b = (n + 4 * 256 - 1) / (4 * 256);  // a few vectors per thread
  if (b > 16384) b = 16384;
  return (unsigned)(b < 1 ? 1 : b);
}


template <typename T>
hipError_t mask_sigmoid(const T* x, long long n, T beta, const double* pbeta, T* m,
                        hipStream_t st) {
  if (n <= 0) return hipSuccess;
  auto k = aligned16(x) && aligned16(m) ? k_mask_sigmoid<T, true> : k_mask_sigmoid<T, false>;
  hipLaunchKernelGGL(k, dim3(elementwise_blocks(n)), dim3(256), 0, st, x, n, beta, pbeta, m);
  return hipGetLastError();
}

template <typename T>
hipError_t mask_sigmoid_grad(const T* gm, const T* m, long long n, T beta, const double* pbeta,
                             T* gx, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  auto k = aligned16(gm) && aligned16(m) && aligned16(gx) ? k_mask_sigmoid_grad<T, true>
                                                          : k_mask_sigmoid_grad<T, false>;
  hipLaunchKernelGGL(k, dim3(elementwise_blocks(n)), dim3(256), 0, st, gm, m, n, beta, pbeta, gx);
  return hipGetLastError();
}

template <typename T>
hipError_t mask_sigmoid_grad_theta(const T* gm, const T* x, long long n, T beta,
                                   const double* pbeta, const T* gscale, T* gx, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  auto k = aligned16(gm) && aligned16(x) && aligned16(gx) ? k_mask_sigmoid_grad_theta<T, true>
                                                          : k_mask_sigmoid_grad_theta<T, false>;
  hipLaunchKernelGGL(k, dim3(elementwise_blocks(n)), dim3(256), 0, st, gm, x, n, beta, pbeta,
                     gscale, gx);
  return hipGetLastError();
}

// ---------------------------------------------------------------- penalties
// Mask penalties of opc/ilt.py:339-355 over images [n_img][ny][nx] with
// periodic neighbours: cost_b = wtv sum (gx^2 + gy^2) / npix + wd sum m (1 - m)
// / npix, gx = m - m(x - 1), gy = m - m(y - 1), and its gradient
// wtv (-2 / npix) lap + wd (1 - 2 m) / npix, lap the 5-point Laplacian; the
// costs reduce in the same two deterministic levels as the resist objective.
template <typename T>
__global__ __launch_bounds__(kObjThreads) void k_penalty_partial(
    const T* __restrict__ m, int ny, int nx, T wtv, T wd, T* __restrict__ grad,
    double* __restrict__ part_cost, int* __restrict__ part_zero) {
  const long long npix = (long long)ny * nx;
  const long long base = (long long)blockIdx.y * npix;
  const T inv = (T)1 / (T)npix;
  double c = 0;
  for (long long e = blockIdx.x * (long long)kObjThreads + threadIdx.x; e < npix;
       e += (long long)gridDim.x * kObjThreads) {
    const int y = (int)(e / nx), x = (int)(e - (long long)y * nx);
    const T v = m[base + e];
    T g = 0;
    if (wtv > (T)0) {
      const T l = m[base + (long long)y * nx + (x == 0 ? nx - 1 : x - 1)];
      const T r = m[base + (long long)y * nx + (x == nx - 1 ? 0 : x + 1)];
      const T u = m[base + (long long)(y == 0 ? ny - 1 : y - 1) * nx + x];
      const T d = m[base + (long long)(y == ny - 1 ? 0 : y + 1) * nx + x];
      const T gx = v - l, gy = v - u;
      c += (double)wtv * (double)(gx * gx + gy * gy);
      g += wtv * ((T)-2 * inv) * (u + d + l + r - (T)4 * v);
    }
    if (wd > (T)0) {
      c += (double)wd * (double)(v * ((T)1 - v));
      g += wd * ((T)1 - (T)2 * v) * inv;
    }
    if (grad) grad[base + e] = g;
  }
  __shared__ double sc[kObjThreads / 64];
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sc[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double cs = 0;
    for (int w = 0; w < kObjThreads / 64; ++w) cs += sc[w];
    const long long k = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    part_cost[k] = cs;
    part_zero[k] = 0;
  }
}

template <typename T>
hipError_t mask_penalties(const T* m, int n_img, int ny, int nx, T wtv, T wd, T* cost, T* total,
                          T* grad, double* part_cost, int* part_zero, hipStream_t st) {
  const long long npix = (long long)ny * nx;
  const int nblk = resist_blocks(npix);
  hipLaunchKernelGGL(k_penalty_partial<T>, dim3(nblk, n_img), dim3(kObjThreads), 0, st, m, ny,
                     nx, wtv, wd, grad, part_cost, part_zero);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_resist_final<T>, dim3(1), dim3(64 * kFinalWaves), 0, st, part_cost,
                     part_zero, n_img, nblk, npix, part_cost + (long long)n_img * nblk, cost, total,
                     (int*)nullptr);
  return hipGetLastError();
}

// ---------------------------------------------------------------- Adam
// opc/ilt.py:587-591 in place: m1 = b1 m1 + (1 - b1) g, m2 = b2 m2 + (1 - b2)
// g^2, theta -= step (m1 / c1) / (sqrt(m2 / c2) + eps), c_k = 1 - b_k^it;
// o_k = 1 - b_k and c_k are formed in float64 by the caller (1 - 0.999f in
// float would be off by 1.3e-5).
// pc (device, optional): c1 = pc[0], c2 = pc[1] instead of the arguments.
template <typename T>
__global__ void k_adam(T* __restrict__ theta, T* __restrict__ m1, T* __restrict__ m2,
                       const T* __restrict__ g, long long n, T b1, T b2, T o1, T o2, T eps,
                       T step, T c1_arg, T c2_arg, const double* __restrict__ pc) {
  const T c1 = pc ? (T)pc[0] : c1_arg, c2 = pc ? (T)pc[1] : c2_arg;
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += st) {
    const T gi = g[i];
    const T a = b1 * m1[i] + o1 * gi;
    const T b = b2 * m2[i] + o2 * (gi * gi);
    m1[i] = a;
    m2[i] = b;
    theta[i] = theta[i] - step * (a / c1) / (sqrt(b / c2) + eps);
  }
}

template <typename T>
hipError_t adam_step(T* theta, T* m1, T* m2, const T* g, long long n, T b1, T b2, T o1, T o2,
                     T eps, T step, T c1, T c2, const double* pc, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_adam<T>, dim3(elementwise_blocks(n)), dim3(256), 0, st, theta, m1, m2, g, n,
                     b1, b2, o1, o2, eps, step, c1, c2, pc);
  return hipGetLastError();
}

// The solver's per-iteration scalars on the device (opc/ilt.py:583-592):
// s = [beta, it, 1 - b1^it, 1 - b2^it, beta_next]; one call advances it by
// one, makes beta_next the current beta and grows the next by `growth`.
__global__ void k_ilt_schedule(double* s, double growth, double b1, double b2) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double it = s[1] + 1.0;
  s[1] = it;
  s[0] = s[4];
  s[4] = s[0] * growth;
  s[2] = 1.0 - pow(b1, it);
  s[3] = 1.0 - pow(b2, it);
}

hipError_t ilt_schedule(double* s, double growth, double b1, double b2, hipStream_t st) {
  hipLaunchKernelGGL(k_ilt_schedule, dim3(1), dim3(64), 0, st, s, growth, b1, b2);
  return hipGetLastError();
}

// ---------------------------------------------------------------- volume
// develop/resist3d.py:172-185: voxel z of mask b takes the two optical planes
// j_z, j_z + 1 around its depth, each normalised by its own scale (the image
// normalisation of that plane), and mixes them linearly:
//   a = s_j raw_j,  c = s_{j+1} raw_{j+1},  vol = a + t_z (c - a).
// One read of the two planes and one write per voxel.
template <typename T>
__global__ void k_volume_lerp(const T* __restrict__ raw, int n_planes, long long npix, int nz,
                              const int* __restrict__ jz, const T* __restrict__ tz,
                              const T* __restrict__ scale, T* __restrict__ vol) {
  const int z = blockIdx.y, b = blockIdx.z;
  const int j = jz[z];
  const T t = tz[z];
  const T s0 = scale[(long long)b * n_planes + j], s1 = scale[(long long)b * n_planes + j + 1];
  const T* r0 = raw + ((long long)b * n_planes + j) * npix;
  const T* r1 = r0 + npix;
  T* o = vol + ((long long)b * nz + z) * npix;
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < npix; e += st) {
    const T a = s0 * r0[e], c = s1 * r1[e];
    o[e] = a + t * (c - a);
  }
}

template <typename T>
hipError_t volume_lerp(const T* raw, int n_masks, int n_planes, long long npix, int nz,
                       const int* jz, const T* tz, const T* scale, T* vol, hipStream_t st) {
  if (n_masks <= 0 || nz <= 0 || npix <= 0) return hipSuccess;
  long long bx = (npix + 4LL * 256 - 1) / (4LL * 256);
  if (bx > 1024) bx = 1024;
  hipLaunchKernelGGL(k_volume_lerp<T>, dim3((unsigned)bx, nz, n_masks), dim3(256), 0, st, raw,
                     n_planes, npix, nz, jz, tz, scale, vol);
  return hipGetLastError();
}

#define LB_OBJ_INST(T)                                                                          \
  template hipError_t resist_objective<T>(const T*, const T*, int, long long, T, T, T, T, T*, T*, \
                                          T*, T*, int*, double*, int*, hipStream_t);           \
  template hipError_t mask_sigmoid<T>(const T*, long long, T, const double*, T*, hipStream_t);  \
  template hipError_t mask_sigmoid_grad<T>(const T*, const T*, long long, T, const double*, T*,   \
                                           hipStream_t);                                        \
  template hipError_t mask_sigmoid_grad_theta<T>(const T*, const T*, long long, T, const double*, \
                                                 const T*, T*, hipStream_t);                     \
  template hipError_t mask_penalties<T>(const T*, int, int, int, T, T, T*, T*, T*, double*, int*, \
                                        hipStream_t);                                           \
  template hipError_t adam_step<T>(T*, T*, T*, const T*, long long, T, T, T, T, T, T, T, T,      \
                                   const double*, hipStream_t);                                  \
  template hipError_t volume_lerp<T>(const T*, int, int, long long, int, const int*, const T*,   \
                                     const T*, T*, hipStream_t);
LB_OBJ_INST(float)
LB_OBJ_INST(double)
#undef LB_OBJ_INST

}  // namespace lb
