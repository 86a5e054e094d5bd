// Band transforms: the only kernels that touch full N x N planes.
//
//   analysis  : X[u][v] = s * sum_{y,x} f[y][x] e^{-2 pi i ((nu0+u) y + (nu0+v) x) / N}
//               for a real (optionally summed over slices) or complex plane f,
//               i.e. fft2 restricted to a square band of frequencies.
//   synthesis : f[y][x] = s * Re sum_{u,v} X[u][v] e^{+2 pi i ((nu0+u) y + (nu0+v) x) / N}
//               (real output; X is Hermitianised on load), or the complex sum.
//
// Both are two passes of LDS Stockham FFTs, one line per transform; real
// planes are packed two rows per complex transform.  The intermediate
// (band x N) is stored transposed so both passes read and write contiguous
// runs.  Grid: (lines / lines-per-block, batch).
#include "lb_kernels.h"

namespace lb {

template <int N>
struct Plan;
#define LB_PLAN(n, a, b, c, d)                               \
  template <>                                                \
  struct Plan<n> {                                           \
    static constexpr int R0 = a, R1 = b, R2 = c, R3 = d;     \
  };
LB_PLAN(8, 8, 1, 1, 1)
LB_PLAN(16, 16, 1, 1, 1)
LB_PLAN(32, 16, 2, 1, 1)
LB_PLAN(64, 16, 4, 1, 1)
LB_PLAN(128, 16, 8, 1, 1)
LB_PLAN(256, 16, 16, 1, 1)
LB_PLAN(512, 16, 16, 2, 1)
LB_PLAN(1024, 16, 16, 4, 1)
LB_PLAN(2048, 16, 16, 8, 1)
LB_PLAN(48, 16, 3, 1, 1)
LB_PLAN(96, 16, 2, 3, 1)
LB_PLAN(192, 16, 4, 3, 1)
LB_PLAN(384, 16, 8, 3, 1)
LB_PLAN(768, 16, 16, 3, 1)
LB_PLAN(1536, 16, 16, 2, 3)
#undef LB_PLAN

constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }
constexpr int pow2_floor(int x) {
  int p = 1;
  while (2 * p <= x) p *= 2;
  return p;
}
// Transforms held per block: about 4096 points, and a divisor of N/2 so the
// row kernels (two rows per transform for real planes) tile N exactly.
template <int N>
constexpr int tpb() {
  return gcd_c(pow2_floor(N >= 4096 ? 1 : 4096 / N), N / 2);
}

constexpr int kBandThreads = 256;

template <typename T, int N, int DIR>
__device__ __forceinline__ cx<T>* run_fft(cx<T>* a, cx<T>* b, int ntr, const cx<T>* tw, int tid) {
  using P = Plan<N>;
  return fft_rows_lds<T, DIR, N, P::R0, P::R1, P::R2, P::R3>(a, b, ntr, N, tw, tid, kBandThreads);
}

// ---------------------------------------------------------------- analysis
template <typename T, int N, bool CPLX>
__global__ __launch_bounds__(kBandThreads) void k_analysis_rows(const T* __restrict__ in,
                                                                long long img_stride,
                                                                long long slice_stride, int nsum,
                                                                int nu0, int Wb,
                                                                cx<T>* __restrict__ RT) {
  constexpr int TPB = tpb<N>();
  constexpr int ROWS = CPLX ? TPB : 2 * TPB;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* A = tw + N;
  cx<T>* B = A + TPB * N;
  const int tid = threadIdx.x;
  const int img = blockIdx.y;
  const int y0 = blockIdx.x * ROWS;
  fill_roots(tw, N, tid, kBandThreads);
  const T* base = in + (long long)img * img_stride;
  for (int t = tid; t < TPB * N; t += kBandThreads) {
    const int tr = t / N, x = t - tr * N;
    T re = 0, im = 0;
    if (CPLX) {
      const long long off = ((long long)(y0 + tr) * N + x) * 2;
      for (int c = 0; c < nsum; ++c) {
        re += base[c * slice_stride + off];
        im += base[c * slice_stride + off + 1];
      }
    } else {
      const long long oa = (long long)(y0 + 2 * tr) * N + x;
      for (int c = 0; c < nsum; ++c) {
        re += base[c * slice_stride + oa];
        im += base[c * slice_stride + oa + N];
      }
    }
    A[t] = cx<T>{re, im};
  }
  __syncthreads();
  cx<T>* Z = run_fft<T, N, -1>(A, B, TPB, tw, tid);
  cx<T>* out = RT + (long long)img * Wb * N;
  for (int t = tid; t < Wb * ROWS; t += kBandThreads) {
    const int v = t / ROWS, yy = t - v * ROWS;
    const int nu = nu0 + v;
    if (CPLX) {
      out[(long long)v * N + y0 + yy] = Z[yy * N + bin_of(nu, N)];
    } else {
      const int tr = yy >> 1;
      const cx<T> a = Z[tr * N + bin_of(nu, N)];
      const cx<T> b = conj(Z[tr * N + bin_of(-nu, N)]);
      cx<T> r;
      if ((yy & 1) == 0) {
        r = scale(a + b, (T)0.5);  // transform of the even row
      } else {
        const cx<T> d = a - b;     // (a - b) / (2i)
        r = cx<T>{d.im * (T)0.5, -d.re * (T)0.5};
      }
      out[(long long)v * N + y0 + yy] = r;
    }
  }
}

template <typename T, int N>
__global__ __launch_bounds__(kBandThreads) void k_analysis_cols(const cx<T>* __restrict__ RT,
                                                                int nu0, int Wb, T s,
                                                                cx<T>* __restrict__ X) {
  constexpr int TPB = tpb<N>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* A = tw + N;
  cx<T>* B = A + TPB * N;
  const int tid = threadIdx.x;
  const int img = blockIdx.y;
  const int v0 = blockIdx.x * TPB;
  fill_roots(tw, N, tid, kBandThreads);
  const cx<T>* src = RT + (long long)img * Wb * N;
  for (int t = tid; t < TPB * N; t += kBandThreads) {
    const int tc = t / N, y = t - tc * N;
    A[t] = (v0 + tc < Wb) ? src[(long long)(v0 + tc) * N + y] : cx<T>{0, 0};
  }
  __syncthreads();
  cx<T>* Z = run_fft<T, N, -1>(A, B, TPB, tw, tid);
  cx<T>* out = X + (long long)img * Wb * Wb;
  for (int t = tid; t < Wb * TPB; t += kBandThreads) {
    const int u = t / TPB, tc = t - u * TPB;
    if (v0 + tc < Wb) out[(long long)u * Wb + v0 + tc] = scale(Z[tc * N + bin_of(nu0 + u, N)], s);
  }
}

// --------------------------------------------------------------- synthesis
// mirror index of u in a band that is symmetric (nu0 = -(Wb-1)/2) or spans
// the whole grid (Wb = N, nu0 = -N/2): the band position of -(nu0 + u).
__device__ __forceinline__ int mirror_of(int u, int nu0, int Wb) {
  int m = (-2 * nu0 - u) % Wb;
  return m < 0 ? m + Wb : m;
}

template <typename T, int N>
__global__ __launch_bounds__(kBandThreads) void k_synth_cols(const cx<T>* __restrict__ Y,
                                                             int nu0, int Wb, int herm,
                                                             cx<T>* __restrict__ TT) {
  constexpr int TPB = tpb<N>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* A = tw + N;
  cx<T>* B = A + TPB * N;
  const int tid = threadIdx.x;
  const int img = blockIdx.y;
  const int v0 = blockIdx.x * TPB;
  fill_roots(tw, N, tid, kBandThreads);
  for (int t = tid; t < TPB * N; t += kBandThreads) A[t] = cx<T>{0, 0};
  __syncthreads();
  const cx<T>* src = Y + (long long)img * Wb * Wb;
  for (int t = tid; t < Wb * TPB; t += kBandThreads) {
    const int u = t / TPB, tc = t - u * TPB;
    const int v = v0 + tc;
    if (v < Wb) {
      cx<T> val = src[(long long)u * Wb + v];
      if (herm) {
        const cx<T> m = src[(long long)mirror_of(u, nu0, Wb) * Wb + mirror_of(v, nu0, Wb)];
        val = scale(val + conj(m), (T)0.5);
      }
      A[tc * N + bin_of(nu0 + u, N)] = val;
    }
  }
  __syncthreads();
  cx<T>* Z = run_fft<T, N, +1>(A, B, TPB, tw, tid);
  cx<T>* out = TT + (long long)img * Wb * N;
  for (int t = tid; t < TPB * N; t += kBandThreads) {
    const int tc = t / N, y = t - tc * N;
    if (v0 + tc < Wb) out[(long long)(v0 + tc) * N + y] = Z[t];
  }
}

template <typename T>
__device__ __forceinline__ void atomic_max_nonneg(T* addr, T v);
template <>
__device__ __forceinline__ void atomic_max_nonneg<float>(float* addr, float v) {
  atomicMax(reinterpret_cast<unsigned int*>(addr), __float_as_uint(v));
}
template <>
__device__ __forceinline__ void atomic_max_nonneg<double>(double* addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(addr),
            (unsigned long long)__double_as_longlong(v));
}

template <typename T>
__device__ T block_max(T v, T* red) {
  const int tid = threadIdx.x;
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) {
    T m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, red[w]);
    red[0] = m;
  }
  __syncthreads();
  return red[0];
}

template <typename T, int N, bool COUT>
__global__ __launch_bounds__(kBandThreads) void k_synth_rows(const cx<T>* __restrict__ TT,
                                                             int nu0, int Wb, T s,
                                                             T* __restrict__ out,
                                                             T* __restrict__ peak) {
  constexpr int TPB = tpb<N>();
  constexpr int ROWS = COUT ? TPB : 2 * TPB;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* A = tw + N;
  cx<T>* B = A + TPB * N;
  T* red = reinterpret_cast<T*>(B + TPB * N);
  const int tid = threadIdx.x;
  const int img = blockIdx.y;
  const int y0 = blockIdx.x * ROWS;
  fill_roots(tw, N, tid, kBandThreads);
  for (int t = tid; t < TPB * N; t += kBandThreads) A[t] = cx<T>{0, 0};
  __syncthreads();
  const cx<T>* src = TT + (long long)img * Wb * N;
  for (int t = tid; t < Wb * TPB; t += kBandThreads) {
    const int v = t / TPB, tr = t - v * TPB;
    const int b = bin_of(nu0 + v, N);
    if (COUT) {
      A[tr * N + b] = src[(long long)v * N + y0 + tr];
    } else {
      const cx<T> p = src[(long long)v * N + y0 + 2 * tr];
      const cx<T> q = src[(long long)v * N + y0 + 2 * tr + 1];
      A[tr * N + b] = cx<T>{p.re - q.im, p.im + q.re};  // p + i q
    }
  }
  __syncthreads();
  cx<T>* Z = run_fft<T, N, +1>(A, B, TPB, tw, tid);
  T lmax = 0;
  if (COUT) {
    cx<T>* o = reinterpret_cast<cx<T>*>(out) + (long long)img * N * N;
    for (int t = tid; t < TPB * N; t += kBandThreads) {
      const int tr = t / N, x = t - tr * N;
      o[(long long)(y0 + tr) * N + x] = scale(Z[t], s);
    }
  } else {
    T* o = out + (long long)img * N * N;
    for (int t = tid; t < TPB * N; t += kBandThreads) {
      const int tr = t / N, x = t - tr * N;
      const T a = Z[t].re * s, b = Z[t].im * s;
      o[(long long)(y0 + 2 * tr) * N + x] = a;
      o[(long long)(y0 + 2 * tr + 1) * N + x] = b;
      lmax = fmax(lmax, fmax(a, b));
    }
    if (peak) {
      const T m = block_max(lmax, red);
      if (tid == 0) atomic_max_nonneg(peak + img, m);
    }
  }
}

// ---------------------------------------------------------------- launchers
template <typename T, int N>
static size_t smem_bytes() {
  return sizeof(cx<T>) * (size_t)(N + 2 * tpb<N>() * N) + 16 * sizeof(T);
}

template <typename K>
static hipError_t set_smem(K kern, size_t bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <typename T, int N>
static hipError_t analysis_n(const void* in, int cplx, long long img_stride,
                             long long slice_stride, int nsum, int nimg, int nu0, int Wb, T s,
                             cx<T>* RT, cx<T>* X, hipStream_t st) {
  constexpr int TPB = tpb<N>();
  const size_t sm = smem_bytes<T, N>();
  hipError_t e;
  if (cplx) {
    auto k = k_analysis_rows<T, N, true>;
    if ((e = set_smem(k, sm)) != hipSuccess) return e;
    dim3 g((N + TPB - 1) / TPB, nimg);
    hipLaunchKernelGGL(k, g, dim3(kBandThreads), sm, st, (const T*)in, img_stride, slice_stride,
                       nsum, nu0, Wb, RT);
  } else {
    auto k = k_analysis_rows<T, N, false>;
    if ((e = set_smem(k, sm)) != hipSuccess) return e;
    dim3 g((N + 2 * TPB - 1) / (2 * TPB), nimg);
    hipLaunchKernelGGL(k, g, dim3(kBandThreads), sm, st, (const T*)in, img_stride, slice_stride,
                       nsum, nu0, Wb, RT);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  auto k2 = k_analysis_cols<T, N>;
  if ((e = set_smem(k2, sm)) != hipSuccess) return e;
  hipLaunchKernelGGL(k2, dim3((Wb + TPB - 1) / TPB, nimg), dim3(kBandThreads), sm, st,
                     (const cx<T>*)RT, nu0, Wb, s, X);
  return hipGetLastError();
}

template <typename T, int N>
static hipError_t synthesis_n(const cx<T>* Y, int nimg, int nu0, int Wb, int herm, T s, int cout,
                              cx<T>* TT, T* out, T* peak, hipStream_t st) {
  constexpr int TPB = tpb<N>();
  const size_t sm = smem_bytes<T, N>();
  hipError_t e;
  auto k1 = k_synth_cols<T, N>;
  if ((e = set_smem(k1, sm)) != hipSuccess) return e;
  hipLaunchKernelGGL(k1, dim3((Wb + TPB - 1) / TPB, nimg), dim3(kBandThreads), sm, st, Y, nu0, Wb,
                     herm, TT);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (cout) {
    auto k2 = k_synth_rows<T, N, true>;
    if ((e = set_smem(k2, sm)) != hipSuccess) return e;
    hipLaunchKernelGGL(k2, dim3((N + TPB - 1) / TPB, nimg), dim3(kBandThreads), sm, st,
                       (const cx<T>*)TT, nu0, Wb, s, out, (T*)nullptr);
  } else {
    auto k2 = k_synth_rows<T, N, false>;
    if ((e = set_smem(k2, sm)) != hipSuccess) return e;
    hipLaunchKernelGGL(k2, dim3((N + 2 * TPB - 1) / (2 * TPB), nimg), dim3(kBandThreads), sm, st,
                       (const cx<T>*)TT, nu0, Wb, s, out, peak);
  }
  return hipGetLastError();
}

#define LB_BAND_SIZES(X) \
  X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(48) X(96) X(192) X(384) X(768) X(1536)

bool band_size_supported(int n) {
  switch (n) {
#define LB_CASE(n) case n:
    LB_BAND_SIZES(LB_CASE)
#undef LB_CASE
    return true;
    default:
      return false;
  }
}

template <typename T>
hipError_t band_analysis(int N, const void* in, int cplx, long long img_stride,
                         long long slice_stride, int nsum, int nimg, int nu0, int Wb, T s,
                         cx<T>* RT, cx<T>* X, hipStream_t st) {
  switch (N) {
#define LB_CASE(n)                                                                             \
  case n:                                                                                      \
    return analysis_n<T, n>(in, cplx, img_stride, slice_stride, nsum, nimg, nu0, Wb, s, RT, X, \
                            st);
    LB_BAND_SIZES(LB_CASE)
#undef LB_CASE
    default:
      return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t band_synthesis(int N, const cx<T>* Y, int nimg, int nu0, int Wb, int herm, T s,
                          int cout, cx<T>* TT, T* out, T* peak, hipStream_t st) {
  switch (N) {
#define LB_CASE(n) \
  case n:          \
    return synthesis_n<T, n>(Y, nimg, nu0, Wb, herm, s, cout, TT, out, peak, st);
    LB_BAND_SIZES(LB_CASE)
#undef LB_CASE
    default:
      return hipErrorInvalidValue;
  }
}

template hipError_t band_analysis<float>(int, const void*, int, long long, long long, int, int,
                                         int, int, float, cx<float>*, cx<float>*, hipStream_t);
template hipError_t band_analysis<double>(int, const void*, int, long long, long long, int, int,
                                          int, int, double, cx<double>*, cx<double>*,
                                          hipStream_t);
template hipError_t band_synthesis<float>(int, const cx<float>*, int, int, int, int, float, int,
                                          cx<float>*, float*, float*, hipStream_t);
template hipError_t band_synthesis<double>(int, const cx<double>*, int, int, int, int, double, int,
                                           cx<double>*, double*, double*, hipStream_t);

}  // namespace lb
