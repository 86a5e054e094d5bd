// Band transforms for any grid side.
//
// The radix plans of lb_band.hip cover the sides 2^k (8..2048), 3 * 2^k
// (48..1536) and 80, and bands no wider than the grid.  Every other case --
// sides such as 100, 250, 1000 or a prime, and image bands wider than a small
// grid (whose frequencies then alias, which the sums below take modulo N) --
// runs here: each of the four one-dimensional passes is a direct DFT over a
// line,
//
//     out[line][k] = s * sum_j in[line][j] * e^{SG 2 pi i ((k + ko)(j + jo) mod N) / N},
//
// with the exponent index advanced by one addition per term and the factors
// read from the exact roots table (no recurrence error).  A block owns 16
// lines x 64 outputs: the 16 lines' inputs are staged in LDS 64 terms at a
// time, each thread keeps 4 lines of one output and sums each staged chunk on
// its own before adding it (the FP32 rounding grows with the chunk and chunk
// counts, not with N).  Same results as the radix path to rounding, at
// O(N^2 Wb) work per image instead of O(N Wb log N).
#include "lb_kernels.h"

namespace lb {

namespace {

constexpr int kAnyK = 64;     // outputs per block
constexpr int kAnyL = 16;     // lines per block
constexpr int kAnyJ = 64;     // terms staged per chunk
constexpr int kAnyNT = 256;   // threads: kAnyK outputs x 4 line groups
constexpr int kAnyLPT = kAnyL / (kAnyNT / kAnyK);  // lines per thread (4)

__device__ __forceinline__ int mod_n(long long v, int n) {
  long long r = v % n;
  return (int)(r < 0 ? r + n : r);
}

// IN_CPLX: complex input (else real, summed over nsum slices); OUT: 0 complex,
// 1 real part (with an optional running maximum per image).
template <typename T, bool IN_CPLX, int OUT, int SG>
__global__ __launch_bounds__(kAnyNT) void k_line_dft(
    const T* __restrict__ in, long long in_img, long long in_slice, int nsum, long long in_line,
    long long in_el, int J, int jo, T* __restrict__ out, long long out_img, long long out_line,
    long long out_el, int K, int ko, int n_lines, int N, T s, const cx<T>* __restrict__ twg,
    T* __restrict__ peak) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);  // N roots e^{+2 pi i t / N}
  cx<T>* tile = tw + N;                             // kAnyL x (kAnyJ + 1)
  const int tid = threadIdx.x;
  const int img = blockIdx.z;
  for (int t = tid; t < N; t += kAnyNT) tw[t] = twg[t];
  const int tk = tid % kAnyK, lg = tid / kAnyK;
  const int k = blockIdx.x * kAnyK + tk;
  const int l0 = blockIdx.y * kAnyL;
  const int fk = mod_n((long long)k + ko, N);  // the output's exponent factor
  cx<T> acc[kAnyLPT];
#pragma unroll
  for (int i = 0; i < kAnyLPT; ++i) acc[i] = cx<T>{0, 0};
  const long long base_in = (long long)img * in_img;
  for (int j0 = 0; j0 < J; j0 += kAnyJ) {
    __syncthreads();  // the previous chunk has been read (and, first time, tw written)
    for (int e = tid; e < kAnyL * kAnyJ; e += kAnyNT) {
      const int l = e / kAnyJ, jj = e - l * kAnyJ;
      const int line = l0 + l, j = j0 + jj;
      cx<T> v = cx<T>{0, 0};
      if (line < n_lines && j < J) {
        const long long off = base_in + (long long)line * in_line + (long long)j * in_el;
        if constexpr (IN_CPLX) {
          v = reinterpret_cast<const cx<T>*>(in)[off];
        } else {
          T r = 0;
          for (int q = 0; q < nsum; ++q) r += in[off + (long long)q * in_slice];
          v = cx<T>{r, (T)0};
        }
      }
      tile[l * (kAnyJ + 1) + jj] = v;
    }
    __syncthreads();
    int idx = mod_n((long long)fk * mod_n((long long)j0 + jo, N), N);
    cx<T> part[kAnyLPT];
#pragma unroll
    for (int i = 0; i < kAnyLPT; ++i) part[i] = cx<T>{0, 0};
    const int jn = min(kAnyJ, J - j0);
    for (int jj = 0; jj < jn; ++jj) {
      cx<T> w = tw[idx];
      if (SG < 0) w.im = -w.im;
#pragma unroll
      for (int i = 0; i < kAnyLPT; ++i)
        part[i] = part[i] + tile[(lg * kAnyLPT + i) * (kAnyJ + 1) + jj] * w;
      idx += fk;
      if (idx >= N) idx -= N;
    }
#pragma unroll
    for (int i = 0; i < kAnyLPT; ++i) acc[i] = acc[i] + part[i];
  }
  T lmax = 0;
  if (k < K) {
#pragma unroll
    for (int i = 0; i < kAnyLPT; ++i) {
      const int line = l0 + lg * kAnyLPT + i;
      if (line >= n_lines) continue;
      const long long o = (long long)img * out_img + (long long)line * out_line + (long long)k * out_el;
      if constexpr (OUT == 0) {
        reinterpret_cast<cx<T>*>(out)[o] = scale(acc[i], s);
      } else {
        const T v = acc[i].re * s;
        out[o] = v;
        lmax = fmax(lmax, v);
      }
    }
  }
  if (OUT == 1 && peak) {
    for (int o = 32; o > 0; o >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, o, 64));
    __shared__ T wmax[kAnyNT / 64];
    if ((tid & 63) == 0) wmax[tid >> 6] = lmax;
    __syncthreads();
    if (tid == 0) {
      T m = 0;
      for (int w = 0; w < kAnyNT / 64; ++w) m = fmax(m, wmax[w]);
      if constexpr (sizeof(T) == 4)
        atomicMax(reinterpret_cast<unsigned int*>(peak + img), __float_as_uint(m));
      else
        atomicMax(reinterpret_cast<unsigned long long*>(peak + img),
                  (unsigned long long)__double_as_longlong((double)m));
    }
  }
}

template <typename T, bool IN_CPLX, int OUT, int SG>
hipError_t line_dft(const void* in, long long in_img, long long in_slice, int nsum,
                    long long in_line, long long in_el, int J, int jo, void* out,
                    long long out_img, long long out_line, long long out_el, int K, int ko,
                    int n_lines, int nimg, int N, T s, const cx<T>* tw, T* peak,
                    hipStream_t st) {
  auto kern = k_line_dft<T, IN_CPLX, OUT, SG>;
  const size_t sm = sizeof(cx<T>) * ((size_t)N + (size_t)kAnyL * (kAnyJ + 1));
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
  if (e != hipSuccess) return e;
  dim3 grid((K + kAnyK - 1) / kAnyK, (n_lines + kAnyL - 1) / kAnyL, nimg);
  hipLaunchKernelGGL(kern, grid, dim3(kAnyNT), sm, st, (const T*)in, in_img, in_slice, nsum,
                     in_line, in_el, J, jo, (T*)out, out_img, out_line, out_el, K, ko, n_lines, N,
                     s, tw, peak);
  return hipGetLastError();
}

}  // namespace

bool band_any_supported(int n) { return n >= 1 && n <= kBandAnyMax; }

template <typename T>
hipError_t band_analysis_any(int N, const void* in, int cplx, long long img_stride,
                             long long slice_stride, int nsum, int nimg, int nu0, int Wb, T s,
                             cx<T>* RT, cx<T>* X, hipStream_t st) {
  if (!band_any_supported(N)) return hipErrorInvalidValue;
  const cx<T>* tw = nullptr;
  hipError_t e = table_for<T>(N, &tw);
  if (e != hipSuccess) return e;
  const int ld = band_ld(Wb);
  const long long rt_img = (long long)N * ld;
  // rows: RT[img][y][v] = sum_x f[y][x] e^{-2 pi i (nu0 + v) x / N}
  if (cplx) {
    // complex planes: in elements are complex, strides in complex units
    e = line_dft<T, true, 0, -1>(in, img_stride / 2, 0, 1, N, 1, N, 0, RT, rt_img, ld, 1, Wb,
                                 nu0, N, nimg, N, (T)1, tw, nullptr, st);
  } else {
    e = line_dft<T, false, 0, -1>(in, img_stride, slice_stride, nsum, N, 1, N, 0, RT, rt_img,
                                  ld, 1, Wb, nu0, N, nimg, N, (T)1, tw, nullptr, st);
  }
  if (e != hipSuccess) return e;
  // columns: X[img][u][v] = s sum_y RT[y][v] e^{-2 pi i (nu0 + u) y / N}
  return line_dft<T, true, 0, -1>(RT, rt_img, 0, 1, 1, ld, N, 0, X, (long long)Wb * Wb, 1, Wb,
                                  Wb, nu0, Wb, nimg, N, s, tw, nullptr, st);
}

template <typename T>
hipError_t band_synthesis_any(int N, const cx<T>* Y, int nimg, int nu0, int Wb, int herm, T s,
                              int cout, cx<T>* TT, T* out, T* peak, hipStream_t st) {
  (void)herm;  // the real output is Re of the sum, Hermitian band or not
  if (!band_any_supported(N)) return hipErrorInvalidValue;
  const cx<T>* tw = nullptr;
  hipError_t e = table_for<T>(N, &tw);
  if (e != hipSuccess) return e;
  const int ld = band_ld(Wb);
  const long long tt_img = (long long)N * ld;
  // columns: TT[img][y][v] = sum_u Y[u][v] e^{+2 pi i (nu0 + u) y / N}
  e = line_dft<T, true, 0, +1>(Y, (long long)Wb * Wb, 0, 1, 1, Wb, Wb, nu0, TT, tt_img, 1, ld,
                               N, 0, Wb, nimg, N, (T)1, tw, nullptr, st);
  if (e != hipSuccess) return e;
  // rows: out[img][y][x] = s (Re) sum_v TT[y][v] e^{+2 pi i (nu0 + v) x / N}
  const long long nn = (long long)N * N;
  if (cout)
    return line_dft<T, true, 0, +1>(TT, tt_img, 0, 1, ld, 1, Wb, nu0, out, nn, N, 1, N, 0, N,
                                    nimg, N, s, tw, nullptr, st);
  return line_dft<T, true, 1, +1>(TT, tt_img, 0, 1, ld, 1, Wb, nu0, out, nn, N, 1, N, 0, N, nimg,
                                  N, s, tw, peak, st);
}

template hipError_t band_analysis_any<float>(int, const void*, int, long long, long long, int,
                                             int, int, int, float, cx<float>*, cx<float>*,
                                             hipStream_t);
template hipError_t band_analysis_any<double>(int, const void*, int, long long, long long, int,
                                              int, int, int, double, cx<double>*, cx<double>*,
                                              hipStream_t);
template hipError_t band_synthesis_any<float>(int, const cx<float>*, int, int, int, int, float,
                                              int, cx<float>*, float*, float*, hipStream_t);
template hipError_t band_synthesis_any<double>(int, const cx<double>*, int, int, int, int,
                                               double, int, cx<double>*, double*, double*,
                                               hipStream_t);

}  // namespace lb
