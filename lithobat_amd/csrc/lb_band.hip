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
// Transforms per wavefront: about 1024 points of wave-private LDS.
template <int N>
constexpr int tpw() {
  return N >= 1024 ? 1 : pow2_floor(1024 / N);
}
constexpr int kBandThreads = 256;
constexpr int kWaves = kBandThreads / 64;

// One Stockham stage over the TPW transforms of one wavefront, in place in
// wave-private LDS `wb` (transform tr at wb + tr * L).  Every lane reads all
// inputs of all its tasks before any lane writes: a wavefront's LDS
// operations complete in issue order, so no barrier is needed.  The first
// stage gathers its inputs through `load(tr, index)`; the last stage hands
// its outputs to `store(tr, index, value)` instead of writing LDS.
template <typename T, int DIR, int L, int LD, int TPW, int R, bool FIRST, bool LAST, class Load,
          class Store>
__device__ __forceinline__ void wave_stage(cx<T>* wb, const cx<T>* tw, int lane, int ntr, int Ns,
                                           Load& load, Store& store) {
  constexpr int S = L / R;
  constexpr int TASKS = TPW * S;
  constexpr int TPL = (TASKS + 63) / 64;
  cx<T> v[TPL][R];
#pragma unroll
  for (int i = 0; i < TPL; ++i) {
    const int t = lane + 64 * i;
    const int tr = t / S, j = t - (t / S) * S;
    if (t < TASKS && tr < ntr) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = FIRST ? load(tr, j + r * S) : wb[tr * LD + j + r * S];
      if (!FIRST) {
        const int k = j % Ns;
        const int step = L / (Ns * R);
#pragma unroll
        for (int r = 1; r < R; ++r) v[i][r] = v[i][r] * tw_get<DIR>(tw, r * k * step);
      }
      dft<R, DIR>(v[i]);
    }
  }
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int i = 0; i < TPL; ++i) {
    const int t = lane + 64 * i;
    const int tr = t / S, j = t - (t / S) * S;
    if (t < TASKS && tr < ntr) {
      const int k = j % Ns;
      const int base = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (LAST)
          store(tr, base + r * Ns, v[i][r]);
        else
          wb[tr * LD + base + r * Ns] = v[i][r];
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
}

// A length-L transform of the wavefront's `ntr` (<= TPW) transforms, held at
// wb + tr * LD.
template <typename T, int DIR, int L, int TPW, int LD = L, class Load, class Store>
__device__ __forceinline__ void wave_fft(cx<T>* wb, const cx<T>* tw, int lane, int ntr, Load& load,
                                         Store& store) {
  using P = Plan<L>;
  constexpr int NST = (P::R0 > 1) + (P::R1 > 1) + (P::R2 > 1) + (P::R3 > 1);
  int Ns = 1;
  wave_stage<T, DIR, L, LD, TPW, P::R0, true, NST == 1>(wb, tw, lane, ntr, Ns, load, store);
  Ns *= P::R0;
  if constexpr (P::R1 > 1) {
    wave_stage<T, DIR, L, LD, TPW, P::R1, false, NST == 2>(wb, tw, lane, ntr, Ns, load, store);
    Ns *= P::R1;
  }
  if constexpr (P::R2 > 1) {
    wave_stage<T, DIR, L, LD, TPW, P::R2, false, NST == 3>(wb, tw, lane, ntr, Ns, load, store);
    Ns *= P::R2;
  }
  if constexpr (P::R3 > 1) {
    wave_stage<T, DIR, L, LD, TPW, P::R3, false, NST == 4>(wb, tw, lane, ntr, Ns, load, store);
  }
}

// Column kernels: CPB columns per block (block-cooperative transposition
// through LDS, so the band x N intermediate is read and written in
// contiguous runs of a row), TPW per wavefront, padded LDS stride.
template <typename T, int N>
constexpr int cpb() {
  constexpr int room = (160 * 1024) / (int)sizeof(cx<T>) - N;  // LDS after the roots table
  constexpr int by_lds = room < N + 1 ? 1 : pow2_floor(room / (N + 1));
  return by_lds < 1 ? 1 : (by_lds > 16 ? 16 : by_lds);
}
template <typename T, int N>
constexpr int col_waves() {
  return cpb<T, N>() / tpw<N>() < 1 ? 1 : cpb<T, N>() / tpw<N>();
}
template <typename T, int N>
constexpr int col_tpw() {  // transforms per wave in the column kernels
  return cpb<T, N>() < tpw<N>() ? cpb<T, N>() : tpw<N>();
}

// Band position of FFT bin b, or -1 when the bin is outside the band.
__device__ __forceinline__ int band_pos(int b, int n, int nu0, int Wb) {
  const int u = nu_of_bin(b, n) - nu0;
  return (u >= 0 && u < Wb) ? u : -1;
}

// mirror index of u in a band that is symmetric (nu0 = -(Wb-1)/2) or spans
// the whole grid (Wb = N, nu0 = -N/2): the band position of -(nu0 + u).
__device__ __forceinline__ int mirror_of(int u, int nu0, int Wb) {
  int m = (-2 * nu0 - u) % Wb;
  return m < 0 ? m + Wb : m;
}

#define LB_BAND_PROLOGUE(T, N)                                       \
  constexpr int TPW = tpw<N>();                                      \
  constexpr int TPB = kWaves * TPW;                                  \
  extern __shared__ __align__(16) unsigned char smem_raw[];          \
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);                    \
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;     \
  cx<T>* wb = tw + N + wave * TPW * N;                               \
  const int img = blockIdx.y;                                        \
  load_roots(tw, twg, N, tid, kBandThreads);                         \
  __syncthreads();

// ---------------------------------------------------------------- analysis
// Pass 1: one forward transform per row (complex planes) or per row pair
// (real planes, packed as row 2p + i row 2p+1); the band columns are pulled
// out of the packed spectrum, RT[img][y][v] (rows of Wb, contiguous).
template <typename T, int N, bool CPLX>
__global__ __launch_bounds__(kBandThreads) void k_analysis_rows(const T* __restrict__ in,
                                                                long long img_stride,
                                                                long long slice_stride, int nsum,
                                                                int nu0, int Wb,
                                                                const cx<T>* __restrict__ twg,
                                                                cx<T>* __restrict__ RT) {
  LB_BAND_PROLOGUE(T, N)
  constexpr int NTR = CPLX ? N : N / 2;
  const int tr_blk = blockIdx.x * TPB;
  const int tr0 = tr_blk + wave * TPW;
  const int ntr = max(0, min(TPW, NTR - tr0));
  const T* base = in + (long long)img * img_stride;
  auto load = [&](int tr, int x) -> cx<T> {
    const int p = tr0 + tr;
    T re = 0, im = 0;
    if (CPLX) {
      const long long off = ((long long)p * N + x) * 2;
      for (int c = 0; c < nsum; ++c) {
        re += base[c * slice_stride + off];
        im += base[c * slice_stride + off + 1];
      }
    } else {
      const long long off = (long long)(2 * p) * N + x;
      for (int c = 0; c < nsum; ++c) {
        re += base[c * slice_stride + off];
        im += base[c * slice_stride + off + N];
      }
    }
    return cx<T>{re, im};
  };
  auto store = [&](int tr, int x, cx<T> v) { wb[tr * N + x] = v; };
  wave_fft<T, -1, N, TPW>(wb, tw, lane, ntr, load, store);
  __syncthreads();
  // block-cooperative extraction of the band columns
  constexpr int ROWS = CPLX ? TPB : 2 * TPB;
  const int y0 = CPLX ? tr_blk : 2 * tr_blk;
  const int nrows = min(ROWS, N - y0);
  cx<T>* out = RT + (long long)img * Wb * N;
  const cx<T>* Zb = tw + N;
  for (int t = tid; t < Wb * ROWS; t += kBandThreads) {
    const int yy = t / Wb, v = t - yy * Wb;
    if (yy >= nrows) continue;
    const int nu = nu0 + v;
    if (CPLX) {
      out[(long long)(y0 + yy) * Wb + v] = Zb[yy * N + bin_of(nu, N)];
    } else {
      const int tr = yy >> 1;
      const cx<T> a = Zb[tr * N + bin_of(nu, N)];
      const cx<T> b = conj(Zb[tr * N + bin_of(-nu, N)]);
      cx<T> r;
      if ((yy & 1) == 0) {
        r = scale(a + b, (T)0.5);  // transform of the even row
      } else {
        const cx<T> d = a - b;     // (a - b) / (2i)
        r = cx<T>{d.im * (T)0.5, -d.re * (T)0.5};
      }
      out[(long long)(y0 + yy) * Wb + v] = r;
    }
  }
}

// Pass 2: one forward transform per band column.  The block loads CPB
// columns of RT[img][y][v] (contiguous runs of a row) into padded LDS; only
// the band rows of each result are stored, X[img][u][v] = s * Z[bin(nu0 + u)].
template <typename T, int N>
__global__ __launch_bounds__((64 * col_waves<T, N>())) void k_analysis_cols(const cx<T>* __restrict__ RT, int nu0, int Wb, T s,
                                const cx<T>* __restrict__ twg, cx<T>* __restrict__ X) {
  constexpr int CPB = cpb<T, N>(), TPW = col_tpw<T, N>(), LD = N + 1;
  constexpr int NT = 64 * col_waves<T, N>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* buf = tw + N;  // CPB x LD
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int img = blockIdx.y;
  const int vb = blockIdx.x * CPB;
  const int ncol = max(0, min(CPB, Wb - vb));
  load_roots(tw, twg, N, tid, NT);
  const cx<T>* src = RT + (long long)img * N * Wb;
  for (int t = tid; t < N * CPB; t += NT) {
    const int y = t / CPB, c = t - y * CPB;
    if (c < ncol) buf[c * LD + y] = src[(long long)y * Wb + vb + c];
  }
  __syncthreads();
  const int c0 = wave * TPW;
  const int ntr = max(0, min(TPW, ncol - c0));
  cx<T>* wb = buf + c0 * LD;
  cx<T>* out = X + (long long)img * Wb * Wb;
  auto load = [&](int tr, int y) -> cx<T> { return wb[tr * LD + y]; };
  auto store = [&](int tr, int b, cx<T> val) {
    const int u = band_pos(b, N, nu0, Wb);
    if (u >= 0) out[(long long)u * Wb + vb + c0 + tr] = scale(val, s);
  };
  wave_fft<T, -1, N, TPW, LD>(wb, tw, lane, ntr, load, store);
}

// --------------------------------------------------------------- synthesis
// Pass 1: one inverse transform per band column v; the first radix stage
// gathers only the in-band inputs from Y (Hermitianised on load when herm).
// The CPB columns of a block are written out cooperatively as row runs,
// TT[img][y][v].
template <typename T, int N>
__global__ __launch_bounds__((64 * col_waves<T, N>())) void k_synth_cols(const cx<T>* __restrict__ Y, int nu0, int Wb, int herm,
                             const cx<T>* __restrict__ twg, cx<T>* __restrict__ TT) {
  constexpr int CPB = cpb<T, N>(), TPW = col_tpw<T, N>(), LD = N + 1;
  constexpr int NT = 64 * col_waves<T, N>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* buf = tw + N;  // CPB x LD
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int img = blockIdx.y;
  const int vb = blockIdx.x * CPB;
  const int ncol = max(0, min(CPB, Wb - vb));
  load_roots(tw, twg, N, tid, NT);
  __syncthreads();
  const int c0 = wave * TPW;
  const int ntr = max(0, min(TPW, ncol - c0));
  cx<T>* wb = buf + c0 * LD;
  const cx<T>* src = Y + (long long)img * Wb * Wb;
  auto load = [&](int tr, int b) -> cx<T> {
    const int u = band_pos(b, N, nu0, Wb);
    if (u < 0) return cx<T>{0, 0};
    const int v = vb + c0 + tr;
    cx<T> val = src[(long long)u * Wb + v];
    if (herm) {
      const cx<T> m = src[(long long)mirror_of(u, nu0, Wb) * Wb + mirror_of(v, nu0, Wb)];
      val = scale(val + conj(m), (T)0.5);
    }
    return val;
  };
  auto store = [&](int tr, int y, cx<T> val) { wb[tr * LD + y] = val; };
  wave_fft<T, +1, N, TPW, LD>(wb, tw, lane, ntr, load, store);
  __syncthreads();
  cx<T>* out = TT + (long long)img * N * Wb;
  for (int t = tid; t < N * CPB; t += NT) {
    const int y = t / CPB, c = t - y * CPB;
    if (c < ncol) out[(long long)y * Wb + vb + c] = buf[c * LD + y];
  }
}

template <typename T>
__device__ __forceinline__ void atomic_max_nonneg(T* addr, T v);
template <>
__device__ __forceinline__ void atomic_max_nonneg<float>(float* addr, float v) {
  atomicMax(reinterpret_cast<unsigned int*>(addr), __float_as_uint(v > 0.f ? v : 0.f));
}
template <>
__device__ __forceinline__ void atomic_max_nonneg<double>(double* addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(addr),
            (unsigned long long)__double_as_longlong(v > 0.0 ? v : 0.0));
}

// Pass 2: one inverse transform per output row (complex output) or row pair
// (real output: row y + i row y+1, valid because the band is Hermitian);
// sparse first stage as above, outputs scaled and stored straight to HBM,
// per-image running maximum for the peak.
template <typename T, int N, bool COUT>
__global__ __launch_bounds__(kBandThreads) void k_synth_rows(const cx<T>* __restrict__ TT,
                                                             int nu0, int Wb, T s,
                                                             const cx<T>* __restrict__ twg,
                                                             T* __restrict__ out,
                                                             T* __restrict__ peak) {
  LB_BAND_PROLOGUE(T, N)
  constexpr int NTR = COUT ? N : N / 2;
  const int tr0 = blockIdx.x * TPB + wave * TPW;
  const int ntr = max(0, min(TPW, NTR - tr0));
  const cx<T>* src = TT + (long long)img * N * Wb;
  auto load = [&](int tr, int b) -> cx<T> {
    const int v = band_pos(b, N, nu0, Wb);
    if (v < 0) return cx<T>{0, 0};
    const int p = tr0 + tr;
    if (COUT) return src[(long long)p * Wb + v];
    const cx<T> a = src[(long long)(2 * p) * Wb + v];
    const cx<T> c = src[(long long)(2 * p + 1) * Wb + v];
    return cx<T>{a.re - c.im, a.im + c.re};  // a + i c
  };
  T lmax = 0;
  auto store = [&](int tr, int x, cx<T> val) {
    const int p = tr0 + tr;
    if (COUT) {
      reinterpret_cast<cx<T>*>(out)[((long long)img * N + p) * N + x] = scale(val, s);
    } else {
      const T a = val.re * s, b = val.im * s;
      T* o = out + (long long)img * N * N;
      o[(long long)(2 * p) * N + x] = a;
      o[(long long)(2 * p + 1) * N + x] = b;
      lmax = fmax(lmax, fmax(a, b));
    }
  };
  wave_fft<T, +1, N, TPW>(wb, tw, lane, ntr, load, store);
  if (!COUT && peak) {
    for (int o = 32; o > 0; o >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, o, 64));
    if (lane == 0) atomic_max_nonneg(peak + img, lmax);
  }
}

// ---------------------------------------------------------------- launchers
template <typename T, int N>
static size_t smem_bytes() {
  return sizeof(cx<T>) * (size_t)(N + kWaves * tpw<N>() * N);
}

template <typename T, int N>
static size_t col_smem_bytes() {
  return sizeof(cx<T>) * (size_t)(N + cpb<T, N>() * (N + 1));
}

template <typename K>
static hipError_t set_smem(K kern, size_t bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

static unsigned blocks_for(int lines, int per_block) {
  return (unsigned)((lines + per_block - 1) / per_block);
}

template <typename T, int N>
static hipError_t analysis_n(const void* in, int cplx, long long img_stride,
                             long long slice_stride, int nsum, int nimg, int nu0, int Wb, T s,
                             cx<T>* RT, cx<T>* X, hipStream_t st) {
  constexpr int TPB = kWaves * tpw<N>();
  const size_t sm = smem_bytes<T, N>();
  const cx<T>* twg = nullptr;
  hipError_t e = table_for<T>(N, &twg);
  if (e != hipSuccess) return e;
  if (cplx) {
    auto k = k_analysis_rows<T, N, true>;
    if ((e = set_smem(k, sm)) != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(blocks_for(N, TPB), nimg), dim3(kBandThreads), sm, st,
                       (const T*)in, img_stride, slice_stride, nsum, nu0, Wb, twg, RT);
  } else {
    auto k = k_analysis_rows<T, N, false>;
    if ((e = set_smem(k, sm)) != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(blocks_for(N / 2, TPB), nimg), dim3(kBandThreads), sm, st,
                       (const T*)in, img_stride, slice_stride, nsum, nu0, Wb, twg, RT);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  auto k2 = k_analysis_cols<T, N>;
  const size_t smc = col_smem_bytes<T, N>();
  if ((e = set_smem(k2, smc)) != hipSuccess) return e;
  hipLaunchKernelGGL(k2, dim3(blocks_for(Wb, cpb<T, N>()), nimg), dim3(64 * col_waves<T, N>()),
                     smc, st, (const cx<T>*)RT, nu0, Wb, s, twg, X);
  return hipGetLastError();
}

template <typename T, int N>
static hipError_t synthesis_n(const cx<T>* Y, int nimg, int nu0, int Wb, int herm, T s, int cout,
                              cx<T>* TT, T* out, T* peak, hipStream_t st) {
  constexpr int TPB = kWaves * tpw<N>();
  const size_t sm = smem_bytes<T, N>();
  const cx<T>* twg = nullptr;
  hipError_t e = table_for<T>(N, &twg);
  if (e != hipSuccess) return e;
  auto k1 = k_synth_cols<T, N>;
  const size_t smc = col_smem_bytes<T, N>();
  if ((e = set_smem(k1, smc)) != hipSuccess) return e;
  hipLaunchKernelGGL(k1, dim3(blocks_for(Wb, cpb<T, N>()), nimg), dim3(64 * col_waves<T, N>()),
                     smc, st, Y, nu0, Wb, herm, twg, TT);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (cout) {
    auto k2 = k_synth_rows<T, N, true>;
    if ((e = set_smem(k2, sm)) != hipSuccess) return e;
    hipLaunchKernelGGL(k2, dim3(blocks_for(N, TPB), nimg), dim3(kBandThreads), sm, st,
                       (const cx<T>*)TT, nu0, Wb, s, twg, out, (T*)nullptr);
  } else {
    auto k2 = k_synth_rows<T, N, false>;
    if ((e = set_smem(k2, sm)) != hipSuccess) return e;
    hipLaunchKernelGGL(k2, dim3(blocks_for(N / 2, TPB), nimg), dim3(kBandThreads), sm, st,
                       (const cx<T>*)TT, nu0, Wb, s, twg, out, peak);
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
