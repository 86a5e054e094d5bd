#include <map>
#include <mutex>
#include <tuple>

// Fill and deterministic segment sums (partials -> images, band gradients ->
// masks).  A fill kernel is used instead of hipMemsetAsync so the whole
// forward/backward chain can be captured into a graph and replayed.
#include "lb_kernels.h"

namespace lb {

template <typename T>
__global__ void k_fill(T* p, T v, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename T>
hipError_t fill(T* p, T v, long long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long long blocks = (n + 255) / 256;
  if (blocks > 65535) blocks = 65535;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

template <typename T>
__device__ __forceinline__ T tmax(T a, T b) {
  return a > b ? a : b;
}

__device__ __forceinline__ void amax(float* a, float v) {
  atomicMax(reinterpret_cast<unsigned int*>(a), __float_as_uint(v > 0.f ? v : 0.f));
}
__device__ __forceinline__ void amax(double* a, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(a),
            (unsigned long long)__double_as_longlong(v > 0.0 ? v : 0.0));
}

// scratch[g][k][e] = sum of slices [k*per, min((k+1)*per, cnt)) of group g
template <typename T>
__global__ void k_segment_split(const T* __restrict__ in, int cnt, int per, int ns, long long len,
                                T* __restrict__ scratch) {
  const int g = blockIdx.y / ns, k = blockIdx.y % ns;
  const int c0 = k * per, c1 = min(cnt, c0 + per);
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len;
       e += (long long)gridDim.x * blockDim.x) {
    T s = 0;
    for (int c = c0; c < c1; ++c) s += in[((long long)g * cnt + c) * len + e];
    scratch[((long long)g * ns + k) * len + e] = s;
  }
}

template <typename T>
__global__ void k_segment_sum(const T* __restrict__ in, int cnt, long long len, T* __restrict__ out,
                              T* __restrict__ peak) {
  const int g = blockIdx.y;
  T lmax = 0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len;
       e += (long long)gridDim.x * blockDim.x) {
    T s = 0;
    for (int k = 0; k < cnt; ++k) s += in[((long long)g * cnt + k) * len + e];
    out[(long long)g * len + e] = s;
    lmax = tmax(lmax, s);
  }
  if (peak) {
    __shared__ T red[16];
    for (int o = 32; o > 0; o >>= 1) lmax = tmax(lmax, (T)__shfl_xor(lmax, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lmax;
    __syncthreads();
    if (threadIdx.x == 0) {
      T m = red[0];
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = tmax(m, red[w]);
      amax(peak + g, m);
    }
  }
}

template <typename T>
__global__ void k_spread_sum(const T* __restrict__ in, int ipg, int spp, long long total,
                             int blocks, int segs, long long len, T* __restrict__ out,
                             T* __restrict__ peak) {
  const int g = blockIdx.y;
  const long long q0 = (long long)g * ipg * spp, q1 = min(total, q0 + (long long)ipg * spp);
  const int bf = spread_run_of(q0, total, blocks), bl = spread_run_of(q1 - 1, total, blocks);
  T lmax = 0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len;
       e += (long long)gridDim.x * blockDim.x) {
    T s = 0;
    for (int b = bf; b <= bl; ++b) {
      const long long lo = (long long)b * total / blocks, hi = (long long)(b + 1) * total / blocks;
      if (lo >= hi) continue;  // an empty run
      const long long im0 = lo / spp;
      for (int k = 0; k < segs; ++k) {
        const long long im = im0 + k;
        if (im * spp >= hi) break;
        if (im / ipg == g) s += in[((long long)k * blocks + b) * len + e];
      }
    }
    out[(long long)g * len + e] = s;
    lmax = tmax(lmax, s);
  }
  if (peak) {
    __shared__ T red[16];
    for (int o = 32; o > 0; o >>= 1) lmax = tmax(lmax, (T)__shfl_xor(lmax, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lmax;
    __syncthreads();
    if (threadIdx.x == 0) {
      T m = red[0];
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = tmax(m, red[w]);
      amax(peak + g, m);
    }
  }
}

template <typename T>
hipError_t spread_sum(const T* in, int groups, int ipg, int spp, long long total, int blocks,
                      int segs, long long len, T* out, T* peak, hipStream_t st) {
  if (groups <= 0 || len <= 0) return hipSuccess;
  long long bx = (len + 255) / 256;
  if (bx > 4096) bx = 4096;
  hipLaunchKernelGGL(k_spread_sum<T>, dim3((unsigned)bx, groups), dim3(256), 0, st, in, ipg, spp,
                     total, blocks, segs, len, out, peak);
  return hipGetLastError();
}

int segment_splits(int groups, int cnt, long long len) {
  long long bx = (len + 255) / 256;
  if (bx > 4096) bx = 4096;
  int ns = 1;
  if (cnt >= 8)
    while (ns < kSegmentSplits && bx * groups * ns < 2048 && ns * 2 <= cnt) ns *= 2;
  return ns;
}

// Sums `cnt` slices per group.  When the slices are many and the slices
// short (few blocks), the slice axis is split first: each of ns blocks sums
// a contiguous run of slices into scratch, then a second pass sums the ns
// partials.  Order of summation is fixed, so results are deterministic.
template <typename T>
hipError_t segment_sum(const T* in, int groups, int cnt, long long len, T* out, T* peak,
                       T* scratch, hipStream_t st) {
  if (groups <= 0 || len <= 0) return hipSuccess;
  long long bx = (len + 255) / 256;
  if (bx > 4096) bx = 4096;
  const int ns = scratch ? segment_splits(groups, cnt, len) : 1;
  if (ns == 1) {
    hipLaunchKernelGGL(k_segment_sum<T>, dim3((unsigned)bx, groups), dim3(256), 0, st, in, cnt,
                       len, out, peak);
    return hipGetLastError();
  }
  const int per = (cnt + ns - 1) / ns;
  hipLaunchKernelGGL(k_segment_split<T>, dim3((unsigned)bx, groups * ns), dim3(256), 0, st, in,
                     cnt, per, ns, len, scratch);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_segment_sum<T>, dim3((unsigned)bx, groups), dim3(256), 0, st,
                     (const T*)scratch, ns, len, out, peak);
  return hipGetLastError();
}

template hipError_t fill<float>(float*, float, long long, hipStream_t);
template hipError_t fill<double>(double*, double, long long, hipStream_t);
template hipError_t segment_sum<float>(const float*, int, int, long long, float*, float*, float*,
                                       hipStream_t);
template hipError_t spread_sum<float>(const float*, int, int, int, long long, int, int, long long,
                                      float*, float*, hipStream_t);
template hipError_t spread_sum<double>(const double*, int, int, int, long long, int, int,
                                       long long, double*, double*, hipStream_t);
template hipError_t segment_sum<double>(const double*, int, int, long long, double*, double*,
                                        double*, hipStream_t);

}  // namespace lb

namespace lb {

// out[k][bin(nu0+u)][bin(nu0+v)] = in[k][u][v]; the rest of out untouched.
template <typename T>
__global__ void k_scatter_band(const cx<T>* __restrict__ in, int K, int Wb, int nu0, int N,
                               cx<T>* __restrict__ out) {
  const int k = blockIdx.y;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < (long long)Wb * Wb;
       e += (long long)gridDim.x * blockDim.x) {
    const int u = (int)(e / Wb), v = (int)(e % Wb);
    const int i = ((nu0 + u) % N + N) % N, j = ((nu0 + v) % N + N) % N;
    out[((long long)k * N + i) * N + j] = in[(long long)k * Wb * Wb + e];
  }
}

template <typename T>
hipError_t scatter_band(const cx<T>* in, int K, int Wb, int nu0, int N, cx<T>* out,
                        hipStream_t st) {
  if (K <= 0) return hipSuccess;
  long long b = ((long long)Wb * Wb + 255) / 256;
  if (b > 1024) b = 1024;
  hipLaunchKernelGGL(k_scatter_band<T>, dim3((unsigned)b, K), dim3(256), 0, st, in, K, Wb, nu0, N,
                     out);
  return hipGetLastError();
}

// Y[m][k][e] = X[m][e] * Hk[k][e], e < len
template <typename T>
__global__ void k_band_mul(const cx<T>* __restrict__ X, const cx<T>* __restrict__ Hk, int K,
                           long long len, cx<T>* __restrict__ Y) {
  const int mk = blockIdx.y;
  const int m = mk / K, k = mk % K;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len;
       e += (long long)gridDim.x * blockDim.x)
    Y[(long long)mk * len + e] = X[(long long)m * len + e] * Hk[(long long)k * len + e];
}

template <typename T>
hipError_t band_mul(const cx<T>* X, const cx<T>* Hk, int n_masks, int K, long long len, cx<T>* Y,
                    hipStream_t st) {
  if (K <= 0 || n_masks <= 0) return hipSuccess;
  long long b = (len + 255) / 256;
  if (b > 1024) b = 1024;
  hipLaunchKernelGGL(k_band_mul<T>, dim3((unsigned)b, n_masks * K), dim3(256), 0, st, X, Hk, K,
                     len, Y);
  return hipGetLastError();
}

template hipError_t scatter_band<double>(const cx<double>*, int, int, int, int, cx<double>*,
                                         hipStream_t);
template hipError_t band_mul<float>(const cx<float>*, const cx<float>*, int, int, long long,
                                    cx<float>*, hipStream_t);
template hipError_t band_mul<double>(const cx<double>*, const cx<double>*, int, int, long long,
                                     cx<double>*, hipStream_t);

}  // namespace lb

namespace lb {

template <typename T>
__global__ void k_roots(cx<T>* tw, int L) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < L; t += gridDim.x * blockDim.x) {
    double s, c;
    sincospi(2.0 * (double)t / (double)L, &s, &c);
    tw[t] = cx<T>{(T)c, (T)s};
  }
}

namespace {
std::mutex g_roots_mu;
std::map<std::tuple<int, int, int>, void*> g_roots;  // (device, sizeof(T), L) -> table
}  // namespace

template <typename T>
const cx<T>* roots_table(int L) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> g(g_roots_mu);
  auto it = g_roots.find(std::make_tuple(dev, (int)sizeof(T), L));
  return it == g_roots.end() ? nullptr : (const cx<T>*)it->second;
}

template <typename T>
hipError_t ensure_roots(int L) {
  if (roots_table<T>(L)) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  void* p = nullptr;
  e = hipMalloc(&p, sizeof(cx<T>) * (size_t)L);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_roots<T>, dim3((L + 255) / 256), dim3(256), 0, 0, (cx<T>*)p, L);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(p);
    return e;
  }
  std::lock_guard<std::mutex> g(g_roots_mu);
  auto key = std::make_tuple(dev, (int)sizeof(T), L);
  if (g_roots.count(key)) {
    (void)hipFree(p);  // another thread won the race
  } else {
    g_roots[key] = p;  // kept for the life of the process
  }
  return hipSuccess;
}

template hipError_t ensure_roots<float>(int);
template hipError_t ensure_roots<double>(int);
template const cx<float>* roots_table<float>(int);
template const cx<double>* roots_table<double>(int);

}  // namespace lb

namespace lb {

// spec[k][u][v] (band order over the whole grid, nu0 = -N/2) = X[u][v] * H_j
// on system j = j0 + k's box, zero elsewhere (zeroed by the caller).
template <typename T>
__global__ void k_dense_spectra(const cx<T>* __restrict__ X, int N, const cx<T>* __restrict__ H,
                                int W, const int2* __restrict__ box, int j0, int nsys,
                                int nu0m, cx<T>* __restrict__ spec) {
  const int k = blockIdx.y;
  if (k >= nsys) return;
  const int j = j0 + k;
  const int2 b = box[j];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < W * W; e += gridDim.x * blockDim.x) {
    const int a = e / W, c = e - (e / W) * W;
    // box frequency (b.x + a + nu0m, b.y + c + nu0m) -> whole-grid band position
    const int uy = (((b.x + a + nu0m) + N / 2) % N + N) % N;
    const int ux = (((b.y + c + nu0m) + N / 2) % N + N) % N;
    const cx<T> h = H[(long long)j * W * W + e];
    spec[((long long)k * N + uy) * N + ux] = X[(long long)uy * N + ux] * h;
  }
}

// img[e] += sum_k w[j0 + k] |E_k[e]|^2
template <typename T>
__global__ void k_accum_norm2(const cx<T>* __restrict__ E, int nsys, long long len,
                              const T* __restrict__ w, int j0, T* __restrict__ img) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len;
       e += (long long)gridDim.x * blockDim.x) {
    T s = img[e];
    for (int k = 0; k < nsys; ++k) s += w[j0 + k] * norm2(E[(long long)k * len + e]);
    img[e] = s;
  }
}

// E_k[e] *= g[e] for the nsys fields of a batch (the sensitivity times the field).
template <typename T>
__global__ void k_mul_real(cx<T>* __restrict__ E, int nsys, long long len,
                           const T* __restrict__ g) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len;
       e += (long long)gridDim.x * blockDim.x) {
    const T gv = g[e];
    for (int k = 0; k < nsys; ++k) E[(long long)k * len + e] = scale(E[(long long)k * len + e], gv);
  }
}

// A[f] += w_j conj(H_j[f]) G_k[f] over each system's box (whole-grid band
// layout, as k_dense_spectra); boxes of one batch overlap, so the sums are
// atomic.  Used only by the dense comparison adjoint.
template <typename T>
__global__ void k_dense_adjoint_accum(const cx<T>* __restrict__ G, int N,
                                      const cx<T>* __restrict__ H, int W,
                                      const int2* __restrict__ box, const T* __restrict__ w,
                                      int j0, int nsys, int nu0m, cx<T>* __restrict__ A) {
  const int k = blockIdx.y;
  if (k >= nsys) return;
  const int j = j0 + k;
  const int2 b = box[j];
  const T wj = w[j];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < W * W; e += gridDim.x * blockDim.x) {
    const int a = e / W, c = e - (e / W) * W;
    const int uy = (((b.x + a + nu0m) + N / 2) % N + N) % N;
    const int ux = (((b.y + c + nu0m) + N / 2) % N + N) % N;
    const long long f = (long long)uy * N + ux;
    const cx<T> v = scale(mulc(G[(long long)k * N * N + f], H[(long long)j * W * W + e]), wj);
    atomicAdd(&A[f].re, v.re);
    atomicAdd(&A[f].im, v.im);
  }
}

template <typename T>
hipError_t mul_real(cx<T>* E, int nsys, long long len, const T* g, hipStream_t st) {
  long long b = (len + 255) / 256;
  if (b > 8192) b = 8192;
  hipLaunchKernelGGL(k_mul_real<T>, dim3((unsigned)b), dim3(256), 0, st, E, nsys, len, g);
  return hipGetLastError();
}

template <typename T>
hipError_t dense_adjoint_accum(const cx<T>* G, int N, const cx<T>* H, int W, const int2* box,
                               const T* w, int j0, int nsys, int nu0m, cx<T>* A, hipStream_t st) {
  hipLaunchKernelGGL(k_dense_adjoint_accum<T>, dim3((W * W + 255) / 256, nsys), dim3(256), 0, st,
                     G, N, H, W, box, w, j0, nsys, nu0m, A);
  return hipGetLastError();
}

// Y_k = g E_k for a chunk of fields (the fields-consuming adjoint's first step).
template <typename T>
__global__ void k_mul_real_into(const cx<T>* __restrict__ E, int nsys, long long len,
                                const T* __restrict__ g, cx<T>* __restrict__ Y) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len;
       e += (long long)gridDim.x * blockDim.x) {
    const T gv = g[e];
    for (int k = 0; k < nsys; ++k) Y[(long long)k * len + e] = scale(E[(long long)k * len + e], gv);
  }
}

// A[u][v] = (A[u][v] if accumulate) + sum_{k < nsys} w_j conj(H_j[a][c]) X_k[u][v] with
// j = j0 + k and (a, c) the band bin's place in system j's box (modulo wrapN when the
// band spans the whole grid).  One thread per band bin, systems in order: the sums do not
// depend on the launch (no atomics).
template <typename T>
__global__ void k_fields_adj_band(const cx<T>* __restrict__ X, int Wb, int wrapN,
                                  const cx<T>* __restrict__ H, int W, const int2* __restrict__ box,
                                  const T* __restrict__ w, int j0, int nsys, int accumulate,
                                  cx<T>* __restrict__ A) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Wb * Wb) return;
  const int u = e / Wb, v = e - u * Wb;
  cx<T> acc = accumulate ? A[e] : cx<T>{0, 0};
  for (int k = 0; k < nsys; ++k) {
    const int j = j0 + k;
    const int2 b = box[j];
    int a = u - b.x, c = v - b.y;
    if (wrapN > 0) {
      a = ((a % wrapN) + wrapN) % wrapN;
      c = ((c % wrapN) + wrapN) % wrapN;
    }
    if (a < 0 || a >= W || c < 0 || c >= W) continue;
    const cx<T> t = mulc(X[(long long)k * Wb * Wb + e], H[(long long)j * W * W + a * W + c]);
    acc = acc + scale(t, w[j]);
  }
  A[e] = acc;
}

template <typename T>
hipError_t mul_real_into(const cx<T>* E, int nsys, long long len, const T* g, cx<T>* Y,
                         hipStream_t st) {
  long long b = (len + 255) / 256;
  if (b > 8192) b = 8192;
  hipLaunchKernelGGL(k_mul_real_into<T>, dim3((unsigned)b), dim3(256), 0, st, E, nsys, len, g, Y);
  return hipGetLastError();
}

template <typename T>
hipError_t fields_adj_band(const cx<T>* X, int Wb, int wrapN, const cx<T>* H, int W,
                           const int2* box, const T* w, int j0, int nsys, int accumulate,
                           cx<T>* A, hipStream_t st) {
  hipLaunchKernelGGL(k_fields_adj_band<T>, dim3((Wb * Wb + 255) / 256), dim3(256), 0, st, X, Wb,
                     wrapN, H, W, box, w, j0, nsys, accumulate, A);
  return hipGetLastError();
}

template hipError_t mul_real_into<float>(const cx<float>*, int, long long, const float*,
                                         cx<float>*, hipStream_t);
template hipError_t mul_real_into<double>(const cx<double>*, int, long long, const double*,
                                          cx<double>*, hipStream_t);
template hipError_t fields_adj_band<float>(const cx<float>*, int, int, const cx<float>*, int,
                                           const int2*, const float*, int, int, int, cx<float>*,
                                           hipStream_t);
template hipError_t fields_adj_band<double>(const cx<double>*, int, int, const cx<double>*, int,
                                            const int2*, const double*, int, int, int,
                                            cx<double>*, hipStream_t);
template hipError_t mul_real<float>(cx<float>*, int, long long, const float*, hipStream_t);
template hipError_t mul_real<double>(cx<double>*, int, long long, const double*, hipStream_t);
template hipError_t dense_adjoint_accum<float>(const cx<float>*, int, const cx<float>*, int,
                                               const int2*, const float*, int, int, int,
                                               cx<float>*, hipStream_t);
template hipError_t dense_adjoint_accum<double>(const cx<double>*, int, const cx<double>*, int,
                                                const int2*, const double*, int, int, int,
                                                cx<double>*, hipStream_t);

template <typename T>
hipError_t dense_spectra(const cx<T>* X, int N, const cx<T>* H, int W, const int2* box, int j0,
                         int nsys, int nu0m, cx<T>* spec, hipStream_t st) {
  hipError_t e = fill<T>((T*)spec, (T)0, 2LL * nsys * N * N, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_dense_spectra<T>, dim3((W * W + 255) / 256, nsys), dim3(256), 0, st, X, N,
                     H, W, box, j0, nsys, nu0m, spec);
  return hipGetLastError();
}

template <typename T>
hipError_t accum_norm2(const cx<T>* E, int nsys, long long len, const T* w, int j0, T* img,
                       hipStream_t st) {
  long long b = (len + 255) / 256;
  if (b > 8192) b = 8192;
  hipLaunchKernelGGL(k_accum_norm2<T>, dim3((unsigned)b), dim3(256), 0, st, E, nsys, len, w, j0,
                     img);
  return hipGetLastError();
}

template hipError_t dense_spectra<float>(const cx<float>*, int, const cx<float>*, int,
                                         const int2*, int, int, int, cx<float>*, hipStream_t);
template hipError_t dense_spectra<double>(const cx<double>*, int, const cx<double>*, int,
                                          const int2*, int, int, int, cx<double>*, hipStream_t);
template hipError_t accum_norm2<float>(const cx<float>*, int, long long, const float*, int,
                                       float*, hipStream_t);
template hipError_t accum_norm2<double>(const cx<double>*, int, long long, const double*, int,
                                        double*, hipStream_t);

}  // namespace lb
