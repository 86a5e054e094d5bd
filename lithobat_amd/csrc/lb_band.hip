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
#include <algorithm>

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
LB_PLAN(80, 16, 5, 1, 1)
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

// LDS layout: one pad slot after every 256 bytes of complex values, so the
// strided writes of a Stockham stage (stride R) and the strided twiddle reads
// (stride r * step) fall in distinct banks.  Transforms, the roots table and
// the transposition buffers all use this indexing.
template <typename T>
constexpr int kPadShift = sizeof(cx<T>) == 8 ? 5 : 4;
template <typename T>
__device__ __forceinline__ int pidx(int i) {
  return i + (i >> kPadShift<T>);
}
template <typename T, int L>
constexpr int padded() {
  return L + (L >> kPadShift<T>);
}

// Wavefronts per block of the row kernels: the most resident waves per CU
// (160 KB of LDS, each block also holding the N-point roots table), ties
// going to the larger block; at most 8 waves so two blocks can overlap.
template <typename T, int N>
constexpr int row_waves() {
  constexpr int cap = (160 * 1024) / (int)sizeof(cx<T>);
  int best = 1, best_waves = 0;
  for (int w = 1; w <= 8; w *= 2) {
    const int per_block = padded<T, N>() + w * tpw<N>() * padded<T, N>();
    int blocks = cap / per_block;
    int waves = blocks * w;
    if (waves > 32) waves = 32;
    if (blocks >= 1 && waves >= best_waves) {
      best_waves = waves;
      best = w;
    }
  }
  return best;
}

// One Stockham stage over the TPW transforms of one wavefront, in place in
// wave-private LDS `wb` (transform tr at wb + tr * L).  Every lane reads all
// inputs of all its tasks before any lane writes: a wavefront's LDS
// operations complete in issue order, so no barrier is needed.  The first
// stage gathers its inputs through `load(tr, index)`; the last stage hands
// its outputs to `store(tr, index, value)` instead of writing LDS.
//
// SPARSE (first stage only): the inputs are known to vanish outside the
// first and the last stride-S run of the line (a band of frequencies
// |nu| < S, S = L / R), so each radix-R DFT has two inputs and reduces to
// X_k = v_0 + v_{R-1} e^{DIR 2 pi i (R-1) k / R}.
//
// RTW: the stage's twiddles come from registers, rtw[i * (R - 1) + r - 1] for
// task slot i and input r (see load_stage_tw), instead of the roots table.
template <typename T, int DIR, int L, int LD, int TPW, int R, int NS, bool FIRST, bool LAST,
          class Load, class Store, bool SPARSE = false, bool RTW = false>
__device__ __forceinline__ void wave_stage(cx<T>* wb, const cx<T>* tw, int lane, int ntr,
                                           Load& load, Store& store,
                                           const cx<T>* rtw = nullptr) {
  constexpr int S = L / R;
  constexpr int TASKS = TPW * S;
  constexpr int TPL = (TASKS + 63) / 64;
  // Padded LDS offsets that are linear in r, so they fold into the
  // instruction's immediate offset: reads j + r S when S is a whole number of
  // 256-byte rows (BANK values); writes base + r NS when the task's R outputs
  // stay in one BANK block (NS R divides BANK) or NS is a power of two and
  // NS R a multiple of BANK (then the base's low part plus r NS never carries).
  constexpr int SH = kPadShift<T>, BANK = 1 << SH;
  constexpr bool RD_LIN = S % BANK == 0;
  constexpr bool NS_POW2 = (NS & (NS - 1)) == 0;
  constexpr bool WR_LIN = NS_POW2 && ((NS * R) % BANK == 0 || BANK % (NS * R) == 0);
  cx<T> v[TPL][R];
#pragma unroll
  for (int i = 0; i < TPL; ++i) {
    const int t = lane + 64 * i;
    const int tr = t / S, j = t - (t / S) * S;
    if (t < TASKS && tr < ntr) {
      if constexpr (FIRST && SPARSE) {
        const cx<T> a = load(tr, j), b = load(tr, j + (R - 1) * S);
#pragma unroll
        for (int k = 0; k < R; ++k) v[i][k] = a + mul_root<R, DIR, T>(b, (R - 1) * k);
        continue;
      }
      const cx<T>* rd = wb + tr * LD + pidx<T>(j);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (FIRST)
          v[i][r] = load(tr, j + r * S);
        else if constexpr (RD_LIN)
          v[i][r] = rd[r * (S + (S >> SH))];
        else
          v[i][r] = wb[tr * LD + pidx<T>(j + r * S)];
      }
      if (!FIRST) {
        const int k = j % NS;
        constexpr int step = L / (NS * R);
#pragma unroll
        for (int r = 1; r < R; ++r) {
          if constexpr (RTW)
            v[i][r] = v[i][r] * rtw[i * (R - 1) + r - 1];
          else
            v[i][r] = v[i][r] * tw_get<DIR>(tw, pidx<T>(r * k * step));
        }
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
      const int k = j % NS;
      const int base = (j / NS) * NS * R + k;
      cx<T>* wr = wb + tr * LD + pidx<T>(base);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (LAST)
          store(tr, base + r * NS, v[i][r]);
        else if constexpr (WR_LIN)
          wr[r * NS + ((r * NS) >> SH)] = v[i][r];
        else
          wb[tr * LD + pidx<T>(base + r * NS)] = v[i][r];
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
}

// A length-L transform of the wavefront's `ntr` (<= TPW) transforms, held at
// wb + tr * LD in padded indexing; tw is the padded roots table.
// SPARSE: the input vanishes outside the first and last L / R0 points.
template <typename T, int DIR, int L, int TPW, int LD = padded<T, L>(), bool SPARSE = false,
          class Load, class Store>
__device__ __forceinline__ void wave_fft(cx<T>* wb, const cx<T>* tw, int lane, int ntr, Load& load,
                                         Store& store) {
  using P = Plan<L>;
  constexpr int NST = (P::R0 > 1) + (P::R1 > 1) + (P::R2 > 1) + (P::R3 > 1);
  wave_stage<T, DIR, L, LD, TPW, P::R0, 1, true, NST == 1, Load, Store, SPARSE>(wb, tw, lane, ntr,
                                                                                load, store);
  if constexpr (P::R1 > 1)
    wave_stage<T, DIR, L, LD, TPW, P::R1, P::R0, false, NST == 2>(wb, tw, lane, ntr, load, store);
  if constexpr (P::R2 > 1)
    wave_stage<T, DIR, L, LD, TPW, P::R2, P::R0 * P::R1, false, NST == 3>(wb, tw, lane, ntr, load,
                                                                          store);
  if constexpr (P::R3 > 1)
    wave_stage<T, DIR, L, LD, TPW, P::R3, P::R0 * P::R1 * P::R2, false, NST == 4>(
        wb, tw, lane, ntr, load, store);
}

// Twiddle count of one non-first stage for one lane: task slots x (R - 1).
template <int L, int TPW, int R>
constexpr int stage_ntw() {
  return R > 1 ? (TPW * (L / R) + 63) / 64 * (R - 1) : 0;
}

// The twiddles of one non-first stage for this lane's task slots, in the
// order wave_stage<..., RTW = true> reads them.  They depend on the lane
// only, so a persistent kernel loads them once and reuses them per unit.
template <typename T, int DIR, int L, int TPW, int R, int NS>
__device__ __forceinline__ void load_stage_tw(cx<T>* w, const cx<T>* tw, int lane) {
  constexpr int S = L / R, TPL = (TPW * S + 63) / 64, step = L / (NS * R);
#pragma unroll
  for (int i = 0; i < TPL; ++i) {
    const int t = lane + 64 * i;
    const int k = (t - (t / S) * S) % NS;  // r * k * step < L: inside the table
#pragma unroll
    for (int r = 1; r < R; ++r) w[i * (R - 1) + r - 1] = tw_get<DIR>(tw, pidx<T>(r * k * step));
  }
}

// Register copy of the twiddles of stages 2.. of a length-L transform, in
// single precision: stage after stage while the running count stays within
// BUDGET complex values per lane; the other stages read the roots table.
template <typename T, int DIR, int L, int TPW, int BUDGET>
struct TailTw {
  using P = Plan<L>;
  static constexpr int N1 = stage_ntw<L, TPW, P::R1>(), N2 = stage_ntw<L, TPW, P::R2>(),
                       N3 = stage_ntw<L, TPW, P::R3>();
  static constexpr bool ON1 = sizeof(T) == 4 && N1 > 0 && N1 <= BUDGET;
  static constexpr bool ON2 = ON1 && N2 > 0 && N1 + N2 <= BUDGET;
  static constexpr bool ON3 = ON2 && N3 > 0 && N1 + N2 + N3 <= BUDGET;
  cx<T> w1[ON1 ? N1 : 1], w2[ON2 ? N2 : 1], w3[ON3 ? N3 : 1];
  __device__ __forceinline__ void load(const cx<T>* tw, int lane) {
    if constexpr (ON1) load_stage_tw<T, DIR, L, TPW, P::R1, P::R0>(w1, tw, lane);
    if constexpr (ON2) load_stage_tw<T, DIR, L, TPW, P::R2, P::R0 * P::R1>(w2, tw, lane);
    if constexpr (ON3) load_stage_tw<T, DIR, L, TPW, P::R3, P::R0 * P::R1 * P::R2>(w3, tw, lane);
  }
};

// Stages 2.. of a length-L transform whose first stage the caller has
// written into wb (Stockham order, span R0); twiddles from `rt` when it holds
// them, else from the roots table.
template <typename T, int DIR, int L, int TPW, int LD, class Store, int BUDGET>
__device__ __forceinline__ void wave_fft_tail(cx<T>* wb, const cx<T>* tw, int lane, int ntr,
                                              Store& store,
                                              const TailTw<T, DIR, L, TPW, BUDGET>& rt) {
  using P = Plan<L>;
  using Q = TailTw<T, DIR, L, TPW, BUDGET>;
  constexpr int NST = (P::R0 > 1) + (P::R1 > 1) + (P::R2 > 1) + (P::R3 > 1);
  static_assert(NST >= 2, "the tail needs a second stage");
  auto none = [](int, int) { return cx<T>{0, 0}; };
  wave_stage<T, DIR, L, LD, TPW, P::R1, P::R0, false, NST == 2, decltype(none), Store, false,
             Q::ON1>(wb, tw, lane, ntr, none, store, rt.w1);
  if constexpr (P::R2 > 1)
    wave_stage<T, DIR, L, LD, TPW, P::R2, P::R0 * P::R1, false, NST == 3, decltype(none), Store,
               false, Q::ON2>(wb, tw, lane, ntr, none, store, rt.w2);
  if constexpr (P::R3 > 1)
    wave_stage<T, DIR, L, LD, TPW, P::R3, P::R0 * P::R1 * P::R2, false, NST == 4,
               decltype(none), Store, false, Q::ON3>(wb, tw, lane, ntr, none, store, rt.w3);
}

// Column kernels: CPB columns per block (block-cooperative transposition
// through LDS, so the band x N intermediate is read and written in
// contiguous runs of a row), TPW per wavefront.  Column c sits at c * LDC:
// LDC = padded length + 64 / CPB puts the 64 lanes of a cooperative copy
// (CPB columns x 64 / CPB rows) in distinct banks.
// CPB: the most columns (a power of two, at most 16) for which two blocks fit
// in a CU's LDS, so that one block's transforms overlap the other's
// write-out (at 1024 points, 8 columns: 61 us against 78 us for one block of
// 16 on a 320-image synthesis), unless that leaves fewer than 8 columns; then
// as many as one block holds.
template <typename T, int N>
constexpr int cpb_for(int lds_elems) {
  const int room = lds_elems - padded<T, N>();  // after the roots
  const int per = padded<T, N>() + 64;
  const int by_lds = room < per ? 1 : pow2_floor(room / per);
  return by_lds < 1 ? 1 : (by_lds > 16 ? 16 : by_lds);
}
template <typename T, int N>
constexpr int cpb() {
  constexpr int cap = (160 * 1024) / (int)sizeof(cx<T>);
  return cpb_for<T, N>(cap / 2) >= 8 ? cpb_for<T, N>(cap / 2) : cpb_for<T, N>(cap);
}
template <typename T, int N>
constexpr int col_ld() {
  return padded<T, N>() + 64 / cpb<T, N>();
}
template <typename T, int N>
constexpr int col_waves() {
  return cpb<T, N>() / tpw<N>() < 1 ? 1 : cpb<T, N>() / tpw<N>();
}
template <typename T, int N>
constexpr int col_tpw() {  // transforms per wave in the column kernels
  return cpb<T, N>() < tpw<N>() ? cpb<T, N>() : tpw<N>();
}

template <typename T>
__device__ __forceinline__ void load_roots_padded(cx<T>* tw, const cx<T>* __restrict__ twg, int L,
                                                  int tid, int nt) {
  for (int t = tid; t < L; t += nt) tw[pidx<T>(t)] = twg[t];
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
  constexpr int NT = 64 * row_waves<T, N>();                         \
  constexpr int TPB = row_waves<T, N>() * TPW;                       \
  constexpr int PL = padded<T, N>();                                 \
  extern __shared__ __align__(16) unsigned char smem_raw[];          \
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);                    \
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;     \
  cx<T>* wb = tw + PL + wave * TPW * PL;                             \
  const int img = blockIdx.y;                                        \
  load_roots_padded(tw, twg, N, tid, NT);                            \
  __syncthreads();

// ---------------------------------------------------------------- analysis
// Pass 1: one forward transform per row (complex planes) or per row pair
// (real planes, packed as row 2p + i row 2p+1); the band columns are pulled
// out of the packed spectrum, RT[img][y][v] (rows of Wb, contiguous).
template <typename T, int N, bool CPLX>
__global__ __launch_bounds__((64 * row_waves<T, N>())) void k_analysis_rows(const T* __restrict__ in,
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
  auto store = [&](int tr, int x, cx<T> v) { wb[tr * PL + pidx<T>(x)] = v; };
  wave_fft<T, -1, N, TPW>(wb, tw, lane, ntr, load, store);
  __syncthreads();
  // block-cooperative extraction of the band columns
  constexpr int ROWS = CPLX ? TPB : 2 * TPB;
  const int y0 = CPLX ? tr_blk : 2 * tr_blk;
  const int nrows = min(ROWS, N - y0);
  const int ld = band_ld(Wb);
  cx<T>* out = RT + (long long)img * ld * N;
  const cx<T>* Zb = tw + PL;
  for (int t = tid; t < Wb * ROWS; t += NT) {
    const int yy = t / Wb, v = t - yy * Wb;
    if (yy >= nrows) continue;
    const int nu = nu0 + v;
    if (CPLX) {
      out[(long long)(y0 + yy) * ld + v] = Zb[yy * PL + pidx<T>(bin_of(nu, N))];
    } else {
      const int tr = yy >> 1;
      const cx<T> a = Zb[tr * PL + pidx<T>(bin_of(nu, N))];
      const cx<T> b = conj(Zb[tr * PL + pidx<T>(bin_of(-nu, N))]);
      cx<T> r;
      if ((yy & 1) == 0) {
        r = scale(a + b, (T)0.5);  // transform of the even row
      } else {
        const cx<T> d = a - b;     // (a - b) / (2i)
        r = cx<T>{d.im * (T)0.5, -d.re * (T)0.5};
      }
      out[(long long)(y0 + yy) * ld + v] = r;
    }
  }
}

// Waves per SIMD that the LDS of the row kernels allows, at most 4 (the
// register budget is held to it, so registers never cut the residency).
template <typename T, int N>
constexpr int row_waves_per_simd() {
  constexpr int cap = (160 * 1024) / (int)sizeof(cx<T>);
  constexpr int per_block = padded<T, N>() * (1 + row_waves<T, N>() * tpw<N>());
  constexpr int w = (cap / per_block) * row_waves<T, N>() / 4;
  return w < 1 ? 1 : (w > 4 ? 4 : w);
}

// Row pass of the analysis for one input slice, persistent per wavefront:
// each wavefront loops over transform units (TPW rows, or row pairs of a
// real plane), loading the next unit's inputs while it transforms the
// current one, and extracts its own units' band columns (no block barrier
// after the roots table).
// Wavefronts per block of the persistent analysis rows (at most 4), and the
// blocks per CU the LDS allows, at most 3 (the register budget is held to
// it; more would spill).
template <typename T, int N>
constexpr int ap_waves() {
  return row_waves<T, N>() < 4 ? row_waves<T, N>() : 4;
}
template <typename T, int N>
constexpr int ap_blocks_per_cu() {
  constexpr int cap = (160 * 1024) / (int)sizeof(cx<T>);
  constexpr int b = cap / (padded<T, N>() * (1 + ap_waves<T, N>() * tpw<N>()));
  return b < 1 ? 1 : (b > 3 ? 3 : b);
}
template <typename T, int N>
static size_t ap_smem_bytes() {
  return sizeof(cx<T>) * (size_t)(padded<T, N>() * (1 + ap_waves<T, N>() * tpw<N>()));
}

// SIG: the real input is the ILT parameter theta and the row pass transforms
// the mask sigmoid(clip(beta theta, +-60)) (applied as the unit's values are
// taken from the prefetch registers).
// HALF (real input, band symmetric about 0, Wb = 2Q + 1): a real row's
// spectrum at -nu is the conjugate of that at nu, so only nu = 0..Q is
// stored, RT[img][y] holding Q + 1 values (k_analysis_cols<..., HALF>).
template <typename T, int N, bool CPLX, bool SIG = false, bool HALF = false>
__global__ __launch_bounds__((64 * ap_waves<T, N>()),
                             (ap_blocks_per_cu<T, N>() * ap_waves<T, N>() / 4 < 1
                                  ? 1
                                  : ap_blocks_per_cu<T, N>() * ap_waves<T, N>() /
                                        4)) void k_analysis_rows_p(
    const T* __restrict__ in, long long img_stride, int nu0, int Wb,
    const cx<T>* __restrict__ twg, cx<T>* __restrict__ RT, T beta) {
  static_assert(!(SIG && CPLX), "the sigmoid applies to a real parameter");
  static_assert(!(HALF && CPLX), "the half band is for real input");
  constexpr int TPW = tpw<N>();
  constexpr int NW = ap_waves<T, N>();
  constexpr int PL = padded<T, N>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  cx<T>* wb = tw + PL + wave * TPW * PL;
  const int img = blockIdx.y;
  load_roots_padded(tw, twg, N, tid, 64 * NW);
  __syncthreads();
  constexpr int R = Plan<N>::R0, S = N / R;
  constexpr int NTR = CPLX ? N : N / 2;
  constexpr int NU = (NTR + TPW - 1) / TPW;
  constexpr int TASKS = TPW * S, TPL = (TASKS + 63) / 64;
  constexpr int ROWS_U = CPLX ? TPW : 2 * TPW;  // image rows per unit
  const T* base = in + (long long)img * img_stride;
  int trs[TPL], jj[TPL];
#pragma unroll
  for (int i = 0; i < TPL; ++i) {
    const int t = lane + 64 * i;
    trs[i] = t < TASKS ? t / S : TPW;
    jj[i] = t - (t / S) * S;
  }
  // the unit's inputs x = j + r S: real part / imaginary part (complex plane)
  // or even row / odd row (real plane, packed as even + i odd)
  T ra[TPL][R], rb[TPL][R];
  auto fetch = [&](int u) {
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
      const int p = u * TPW + trs[i];
      const bool ok = trs[i] < TPW && p < NTR;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int x = jj[i] + r * S;
        if (CPLX) {
          const cx<T> c = ok ? reinterpret_cast<const cx<T>*>(base)[(long long)p * N + x]
                             : cx<T>{0, 0};
          ra[i][r] = c.re;
          rb[i][r] = c.im;
        } else {
          const long long off = (long long)(2 * p) * N + x;
          ra[i][r] = ok ? base[off] : (T)0;
          rb[i][r] = ok ? base[off + N] : (T)0;
        }
      }
    }
  };
  const int stride = gridDim.x * NW;
  int u = blockIdx.x * NW + wave;
  if (u < NU) fetch(u);
  wait_vmem_all();  // no loads in flight at the loop head (see wait_vmem_all)
  TailTw<T, -1, N, TPW, 32> rt;  // 3 waves per SIMD: room for 64 more VGPRs
  rt.load(tw, lane);
  const int ncols = HALF ? (Wb - 1) / 2 + 1 : Wb;  // stored column vv is nu = vv when HALF
  const int ld = band_ld(ncols);
  cx<T>* out = RT + (long long)img * ld * N;
  for (; u < NU; u += stride) {
    cx<T> v[TPL][R];
#pragma unroll
    for (int i = 0; i < TPL; ++i)
#pragma unroll
      for (int r = 0; r < R; ++r)
        v[i][r] = SIG ? cx<T>{sigmoid_clipped(beta * ra[i][r]), sigmoid_clipped(beta * rb[i][r])}
                      : cx<T>{ra[i][r], rb[i][r]};
    if (u + stride < NU) fetch(u + stride);
    const int ntr = min(TPW, NTR - u * TPW);
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
      if (trs[i] < ntr) {
        dft<R, -1>(v[i]);
        cx<T>* w = wb + trs[i] * PL;
        cx<T>* w0 = w + pidx<T>(jj[i] * R);
        constexpr bool LIN = (1 << kPadShift<T>) % R == 0;  // the R outputs share a bank row
#pragma unroll
        for (int k = 0; k < R; ++k) {
          if constexpr (LIN)
            w0[k] = v[i][k];
          else
            w[pidx<T>(jj[i] * R + k)] = v[i][k];
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    auto store = [&](int tr, int x, cx<T> val) { wb[tr * PL + pidx<T>(x)] = val; };
    wave_fft_tail<T, -1, N, TPW, PL>(wb, tw, lane, ntr, store, rt);
    // band columns of this unit's rows, RT[img][y][v]
    const int y0 = u * ROWS_U;
    const int nrows = min(ROWS_U, N - y0);
    for (int t = lane; t < ncols * nrows; t += 64) {
      const int yy = t / ncols, vv = t - yy * ncols;
      const int nu = HALF ? vv : nu0 + vv;
      cx<T> r;
      if (CPLX) {
        r = wb[yy * PL + pidx<T>(bin_of(nu, N))];
      } else {
        const int tr = yy >> 1;
        const cx<T> a = wb[tr * PL + pidx<T>(bin_of(nu, N))];
        const cx<T> b = conj(wb[tr * PL + pidx<T>(bin_of(-nu, N))]);
        if ((yy & 1) == 0) {
          r = scale(a + b, (T)0.5);  // transform of the even row
        } else {
          const cx<T> d = a - b;     // (a - b) / (2i)
          r = cx<T>{d.im * (T)0.5, -d.re * (T)0.5};
        }
      }
      out[(long long)(y0 + yy) * ld + vv] = r;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Pass 2: one forward transform per band column.  The block loads CPB
// columns of RT[img][y][v] (contiguous runs of a row) into padded LDS; only
// the band rows of each result are stored, X[img][u][v] = s * Z[bin(nu0 + u)].
// HALF: RT holds the columns nu = 0..Q of a real plane; column nu > 0 also
// gives the mirrored column, X[2Q - u][Q - nu] = conj(X[u][Q + nu]).
template <typename T, int N, bool HALF = false>
__global__ __launch_bounds__((64 * col_waves<T, N>())) void k_analysis_cols(const cx<T>* __restrict__ RT, int nu0, int Wb, T s,
                                const cx<T>* __restrict__ twg, cx<T>* __restrict__ X) {
  constexpr int CPB = cpb<T, N>(), TPW = col_tpw<T, N>(), LD = col_ld<T, N>();
  constexpr int NT = 64 * col_waves<T, N>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* buf = tw + padded<T, N>();  // CPB x LD
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int img = blockIdx.y;
  const int vb = blockIdx.x * CPB;
  const int Qh = (Wb - 1) / 2;
  const int nstored = HALF ? Qh + 1 : Wb;
  const int ncol = max(0, min(CPB, nstored - vb));
  load_roots_padded(tw, twg, N, tid, NT);
  const int ld = band_ld(nstored);
  const cx<T>* src = RT + (long long)img * N * ld;
  for (int t = tid; t < N * CPB; t += NT) {
    const int y = t / CPB, c = t - y * CPB;
    if (c < ncol) buf[c * LD + pidx<T>(y)] = src[(long long)y * ld + vb + c];
  }
  __syncthreads();
  const int c0 = wave * TPW;
  const int ntr = max(0, min(TPW, ncol - c0));
  cx<T>* wb = buf + c0 * LD;
  auto load = [&](int tr, int y) -> cx<T> { return wb[tr * LD + pidx<T>(y)]; };
  auto store = [&](int tr, int b, cx<T> val) { wb[tr * LD + pidx<T>(b)] = val; };
  wave_fft<T, -1, N, TPW, LD>(wb, tw, lane, ntr, load, store);
  __syncthreads();
  // the band rows of the block's columns, written out cooperatively as runs
  // of ncol values (a lane per band row would scatter single values Wb apart)
  cx<T>* out = X + (long long)img * Wb * Wb;
  for (int t = tid; t < Wb * CPB; t += NT) {
    const int u = t / CPB, c = t - u * CPB;
    const int b = bin_of(nu0 + u, N);
    if (c >= ncol || band_pos(b, N, nu0, Wb) != u) continue;
    const cx<T> y = scale(buf[c * LD + pidx<T>(b)], s);
    const int v = vb + c;
    if (HALF) {
      out[(long long)u * Wb + Qh + v] = y;
      if (v > 0) out[(long long)(2 * Qh - u) * Wb + Qh - v] = conj(y);
    } else {
      out[(long long)u * Wb + v] = y;
    }
  }
}

// --------------------------------------------------------------- synthesis
// Pass 1: one inverse transform per band column v; the first radix stage
// gathers only the in-band inputs from Y (Hermitianised on load when herm).
// The CPB columns of a block are written out cooperatively as row runs,
// TT[img][y][v].
//
// HALF (real output of a Hermitian band symmetric about 0, Wb = 2Q + 1): each
// row of the output is real, so its spectrum at -nu is the conjugate of that
// at nu and only the columns nu = 0..Q are transformed and stored, TT[img][y]
// holding Q + 1 values (the row pass takes the others as conjugates).
template <typename T, int N, bool SP, bool HALF = false>
__global__ __launch_bounds__((64 * col_waves<T, N>())) void k_synth_cols(const cx<T>* __restrict__ Y, int nu0, int Wb, int herm,
                             const cx<T>* __restrict__ twg, cx<T>* __restrict__ TT) {
  constexpr int CPB = cpb<T, N>(), TPW = col_tpw<T, N>(), LD = col_ld<T, N>();
  constexpr int NT = 64 * col_waves<T, N>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* buf = tw + padded<T, N>();  // CPB x LD
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int img = blockIdx.y;
  const int vb = blockIdx.x * CPB;
  const int Qh = (Wb - 1) / 2;
  const int nstored = HALF ? Qh + 1 : Wb;  // stored column v is band column v + Qh when HALF
  const int ncol = max(0, min(CPB, nstored - vb));
  load_roots_padded(tw, twg, N, tid, NT);
  __syncthreads();
  const int c0 = wave * TPW;
  const int ntr = max(0, min(TPW, ncol - c0));
  cx<T>* wb = buf + c0 * LD;
  const cx<T>* src = Y + (long long)img * Wb * Wb;
  auto load = [&](int tr, int b) -> cx<T> {
    const int u = band_pos(b, N, nu0, Wb);
    if (u < 0) return cx<T>{0, 0};
    const int v = vb + c0 + tr + (HALF ? Qh : 0);
    cx<T> val = src[(long long)u * Wb + v];
    if (herm) {
      const cx<T> m = src[(long long)mirror_of(u, nu0, Wb) * Wb + mirror_of(v, nu0, Wb)];
      val = scale(val + conj(m), (T)0.5);
    }
    return val;
  };
  auto store = [&](int tr, int y, cx<T> val) { wb[tr * LD + pidx<T>(y)] = val; };
  wave_fft<T, +1, N, TPW, LD, SP>(wb, tw, lane, ntr, load, store);
  __syncthreads();
  const int ld = band_ld(nstored);
  cx<T>* out = TT + (long long)img * N * ld;
  for (int t = tid; t < N * CPB; t += NT) {
    const int y = t / CPB, c = t - y * CPB;
    if (c < ncol) out[(long long)y * ld + vb + c] = buf[c * LD + pidx<T>(y)];
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

// The block's running maximum into *peak with one atomic per block: the
// image's blocks all update one address (and the whole batch's peaks share a
// few cache lines), so one atomic per wavefront serialises many of them on the
// same L2 channel.  Each wave leaves its maximum in the first slot of its own
// LDS buffer (no longer in use), at bufs + wave * stride.
template <typename T, int NW, int STRIDE>
__device__ __forceinline__ void block_peak(T lmax, cx<T>* bufs, int lane, int wave, int tid,
                                           T* peak) {
  for (int o = 32; o > 0; o >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, o, 64));
  if (lane == 0) bufs[wave * STRIDE].re = lmax;
  __syncthreads();
  if (tid == 0) {
    T m = bufs[0].re;
    for (int w = 1; w < NW; ++w) m = fmax(m, bufs[w * STRIDE].re);
    atomic_max_nonneg(peak, m);
  }
}

// Pass 2: one inverse transform per output row (complex output) or row pair
// (real output: row y + i row y+1, valid because the band is Hermitian);
// sparse first stage as above, outputs scaled and stored straight to HBM,
// per-image running maximum for the peak.
// HALF (real output): TT holds the columns nu = 0..Q of a band symmetric
// about 0 (see k_synth_cols); the input at nu < 0 is the conjugate of the
// stored column -nu, for each row of the pair.
// STREAM (real output): the image leaves in streaming stores, for batches
// too large to be read back from the caches.
template <typename T, int N, bool COUT, bool SP, bool HALF = false, bool STREAM = false>
__global__ __launch_bounds__((64 * row_waves<T, N>())) void k_synth_rows(const cx<T>* __restrict__ TT,
                                                             int nu0, int Wb, T s,
                                                             const cx<T>* __restrict__ twg,
                                                             T* __restrict__ out,
                                                             T* __restrict__ peak) {
  static_assert(!(HALF && COUT), "the half band is for real output");
  LB_BAND_PROLOGUE(T, N)
  constexpr int NTR = COUT ? N : N / 2;
  const int tr0 = blockIdx.x * TPB + wave * TPW;
  const int ntr = max(0, min(TPW, NTR - tr0));
  const int Qh = (Wb - 1) / 2;
  const int ld = band_ld(HALF ? Qh + 1 : Wb);
  const cx<T>* src = TT + (long long)img * N * ld;
  auto load = [&](int tr, int b) -> cx<T> {
    int v;
    bool neg = false;
    if constexpr (HALF) {
      const int nu = nu_of_bin(b, N);
      neg = nu < 0;
      v = neg ? -nu : nu;
      if (v > Qh) return cx<T>{0, 0};
    } else {
      v = band_pos(b, N, nu0, Wb);
      if (v < 0) return cx<T>{0, 0};
    }
    const int p = tr0 + tr;
    if (COUT) return src[(long long)p * ld + v];
    cx<T> a = src[(long long)(2 * p) * ld + v];
    cx<T> c = src[(long long)(2 * p + 1) * ld + v];
    if (HALF && neg) {
      a = conj(a);
      c = conj(c);
    }
    return cx<T>{a.re - c.im, a.im + c.re};  // a + i c
  };
  T lmax = 0;
  auto store = [&](int tr, int x, cx<T> val) {
    const int p = tr0 + tr;
    if (COUT) {
      reinterpret_cast<cx<T>*>(out)[((long long)img * N + p) * N + x] = scale(val, s);
    } else {
      const T a = val.re * s, b = val.im * s;
      T* o = out + (long long)img * N * N + (long long)(2 * p) * N + x;
      if constexpr (STREAM) {
        __builtin_nontemporal_store(a, o);
        __builtin_nontemporal_store(b, o + N);
      } else {
        o[0] = a;
        o[N] = b;
      }
      lmax = fmax(lmax, fmax(a, b));
    }
  };
  wave_fft<T, +1, N, TPW, PL, SP>(wb, tw, lane, ntr, load, store);
  if (!COUT && peak)
    block_peak<T, row_waves<T, N>(), TPW * PL>(lmax, tw + PL, lane, wave, tid, peak + img);
}

// Row pass for a band inside |nu| < S = N / R0 (the usual case): only two
// inputs per first-stage task are non-zero.  Wavefronts are persistent over
// transform units (TPW transforms each, `stride` apart): the next unit's band
// values are loaded while the current one is transformed, the roots table is
// loaded once per block, and the peak is one atomic per wavefront.
//
// HALF: TT holds the columns nu = 0..Q of a band symmetric about 0 (see
// k_synth_cols); the input at nu = j - S < 0 is the conjugate of the stored
// column S - j.
// STAGED (real output, `out` 16-byte aligned): the last stage's values go to
// the wave's LDS buffer as the two image rows and are written out in 16-byte
// runs, four times fewer store instructions than the stage's own outputs,
// which lie N / R apart.
// EPI (staged): the rows are a mask gradient d/dm and leave as the gradient of
// the ILT parameter, d/dtheta = d/dm (beta m (1 - m) g) with m =
// sigmoid(clip(beta theta, +-60)) recomputed from te.theta (16-byte aligned,
// like out) and g = *te.g (1 when null): lb_mask_sigmoid_grad_theta's pass
// fused into the synthesis.
template <typename T, int N, bool COUT, bool HALF = false, bool STAGED = false, bool EPI = false>
__global__ __launch_bounds__((64 * row_waves<T, N>()),
                             (row_waves_per_simd<T, N>())) void k_synth_rows_sp(
    const cx<T>* __restrict__ TT, int nu0, int Wb, T s, const cx<T>* __restrict__ twg,
    T* __restrict__ out, T* __restrict__ peak, ThetaEpi<T> te) {
  static_assert(!((HALF || STAGED) && COUT), "the half band and staged rows are for real output");
  static_assert(!EPI || STAGED, "the theta epilogue works on staged rows");
  LB_BAND_PROLOGUE(T, N)
  const T gsc = EPI && te.g ? te.g[0] : (T)1;
  constexpr int R = Plan<N>::R0, S = N / R;
  constexpr int NTR = COUT ? N : N / 2;
  constexpr int NU = (NTR + TPW - 1) / TPW;
  constexpr int TASKS = TPW * S, TPL = (TASKS + 63) / 64;
  const int Qh = (Wb - 1) / 2;
  const int ld = band_ld(HALF ? Qh + 1 : Wb);
  const cx<T>* src = TT + (long long)img * N * ld;
  // per task slot: transform within the unit, and the band positions of its
  // two inputs (nu = j and nu = j - S).  The loads below always use an
  // in-range position (clamped) and are masked afterwards, so the prefetch
  // has no branches and its registers carry straight into the next unit.
  int trs[TPL], plo[TPL], phi[TPL], jj[TPL];
  bool vlo[TPL], vhi[TPL];
#pragma unroll
  for (int i = 0; i < TPL; ++i) {
    const int t = lane + 64 * i;
    const int tr = t / S, j = t - (t / S) * S;
    const bool ok = t < TASKS;
    trs[i] = ok ? tr : TPW;
    jj[i] = j;
    int lo = j - nu0, hi = j - S - nu0;
    if constexpr (HALF) {
      lo = j <= Qh ? j : -1;          // nu = j >= 0: stored column j
      hi = S - j <= Qh ? S - j : -1;  // nu = j - S < 0: conjugate of stored column S - j
    }
    vlo[i] = ok && lo >= 0 && lo < Wb;
    vhi[i] = ok && hi >= 0 && hi < Wb;
    plo[i] = vlo[i] ? lo : 0;
    phi[i] = vhi[i] ? hi : 0;
  }
  // raw band values of one unit: even / odd rows (real output) at both inputs
  cx<T> a0[TPL], a1[TPL], b0[TPL], b1[TPL];
  auto fetch = [&](int u) {  // any u < NU
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
      const int p = min(u * TPW + min(trs[i], TPW - 1), NTR - 1);
      const long long r0 = COUT ? (long long)p * ld : (long long)(2 * p) * ld;
      a0[i] = src[r0 + plo[i]];
      b0[i] = src[r0 + phi[i]];
      if (!COUT) {
        a1[i] = src[r0 + ld + plo[i]];
        b1[i] = src[r0 + ld + phi[i]];
      }
    }
  };
  // the two first-stage inputs of every task slot of a unit, scaled
  cx<T> lo[TPL], hi[TPL];
  auto take = [&]() {
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
      // real output: row 2p + i row 2p+1 (the band is Hermitian per row pair)
      const cx<T> l = COUT ? a0[i] : cx<T>{a0[i].re - a1[i].im, a0[i].im + a1[i].re};
      const cx<T> h = COUT   ? b0[i]
                      : HALF ? cx<T>{b0[i].re + b1[i].im, b1[i].re - b0[i].im}  // conj b0 + i conj b1
                             : cx<T>{b0[i].re - b1[i].im, b0[i].im + b1[i].re};
      lo[i] = vlo[i] ? scale(l, s) : cx<T>{0, 0};
      hi[i] = vhi[i] ? scale(h, s) : cx<T>{0, 0};
    }
  };
  const int nw = row_waves<T, N>();
  const int stride = gridDim.x * nw;
  int u = blockIdx.x * nw + wave;
  // The next unit's values are loaded at the top of an iteration and taken
  // into lo / hi at its end, after the unit's image stores: a wave then waits
  // for those loads only, never for its stores, and nothing is in flight at
  // the loop head.
  fetch(min(u, NU - 1));
  take();
  TailTw<T, +1, N, TPW, 16> rt;  // 4 waves per SIMD: 128 VGPRs in all
  rt.load(tw, lane);
  T lmax = 0;
  for (; u < NU; u += stride) {
    fetch(min(u + stride, NU - 1));
    // every unit is full when TPW divides the transform count (a constant then)
    const int ntr = NTR % TPW == 0 ? TPW : min(TPW, NTR - u * TPW);
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
      if (trs[i] < ntr) {
        cx<T>* w = wb + trs[i] * PL;
        cx<T>* w0 = w + pidx<T>(jj[i] * R);
        constexpr bool LIN = (1 << kPadShift<T>) % R == 0;  // the R outputs share a bank row
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const cx<T> val = lo[i] + mul_root<R, +1, T>(hi[i], (R - 1) * k);
          if constexpr (LIN)
            w0[k] = val;
          else
            w[pidx<T>(jj[i] * R + k)] = val;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    const int p0 = u * TPW;
    if constexpr (STAGED) {
      // the two image rows (real parts, then imaginary parts) in LDS
      T* rows = reinterpret_cast<T*>(wb);
      auto store = [&](int tr, int x, cx<T> val) {
        rows[tr * 2 * PL + x] = val.re;
        rows[tr * 2 * PL + N + x] = val.im;
        lmax = fmax(lmax, fmax(val.re, val.im));
      };
      wave_fft_tail<T, +1, N, TPW, PL>(wb, tw, lane, ntr, store, rt);
      using V = T __attribute__((ext_vector_type(16 / sizeof(T))));
      constexpr int K = 16 / (int)sizeof(T), NV = 2 * N / K;
      for (int tr = 0; tr < ntr; ++tr) {
        const V* src = reinterpret_cast<const V*>(rows + tr * 2 * PL);
        const long long r0 = (long long)img * N * N + (long long)(2 * (p0 + tr)) * N;
        V* dst = reinterpret_cast<V*>(out + r0);
        if constexpr (EPI) {
          const V* th = reinterpret_cast<const V*>(te.theta + r0);
#pragma unroll
          for (int k = lane; k < NV; k += 64) {
            V d = src[k];
            const V x = th[k];
#pragma unroll
            for (int c = 0; c < K; ++c) {
              const T m = sigmoid_clipped(te.beta * x[c]);
              d[c] *= te.beta * m * ((T)1 - m) * gsc;
            }
            __builtin_nontemporal_store(d, dst + k);
          }
        } else {
#pragma unroll
          for (int k = lane; k < NV; k += 64) __builtin_nontemporal_store(src[k], dst + k);
        }
      }
      __builtin_amdgcn_wave_barrier();
    } else {
      auto store = [&](int tr, int x, cx<T> val) {
        const int p = p0 + tr;
        if (COUT) {
          reinterpret_cast<cx<T>*>(out)[((long long)img * N + p) * N + x] = val;
        } else {
          // the image is written once and not read back here: streaming stores
          T* o = out + (long long)img * N * N + (long long)(2 * p) * N + x;
          __builtin_nontemporal_store(val.re, o);
          __builtin_nontemporal_store(val.im, o + N);
          lmax = fmax(lmax, fmax(val.re, val.im));
        }
      };
      wave_fft_tail<T, +1, N, TPW, PL>(wb, tw, lane, ntr, store, rt);
    }
    take();
  }
  if (!COUT && peak)
    block_peak<T, row_waves<T, N>(), TPW * PL>(lmax, tw + PL, lane, wave, tid, peak + img);
}

// ------------------------------------------- synthesis + resist objective
// The synthesis rows of a real image over a Hermitian half band (as
// k_synth_rows_sp<T, N, false, true>) fused with the relaxed-resist objective
// and the row pass of the band analysis of the intensity derivative.  Per row
// pair the image rows land in the wave's LDS buffer; the objective of
// opc/ilt.py:311-330 turns them in place into the derivative rows
//   v = sign (scale I - level), z = sigmoid(a v), cost += (z - t)^2,
//   mismatch += (v > 0) != (t > 0.5), dI = dscale (z - t) z (1 - z),
// (cost and mismatch summed per block in a fixed order), and one forward
// transform of (even + i odd) gives the derivative's band columns
// nu = 0..Q, RT[img][y][nu] (the input of k_analysis_cols<..., HALF>).
// Neither the image nor its derivative goes to HBM unless image / dI are
// given.  Four-wave blocks, three per CU.
constexpr int kObjRowWaves = 4;

template <typename T, int N>
__global__ __launch_bounds__(64 * kObjRowWaves, 3) void k_synth_objective_rows(
    const cx<T>* __restrict__ TT, int Wb, T s, const cx<T>* __restrict__ twg,
    const T* __restrict__ target, ResistPix<T> rp, T* __restrict__ image, T* __restrict__ dI,
    cx<T>* __restrict__ RT, double* __restrict__ part_cost, int* __restrict__ part_err) {
  constexpr int NW = kObjRowWaves, TPW = tpw<N>(), PL = padded<T, N>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  cx<T>* wb = tw + PL + wave * TPW * PL;
  const int img = blockIdx.y;
  load_roots_padded(tw, twg, N, tid, 64 * NW);
  __syncthreads();
  constexpr int R = Plan<N>::R0, S = N / R;
  constexpr int NTR = N / 2, NU = (NTR + TPW - 1) / TPW;
  constexpr int TASKS = TPW * S, TPL = (TASKS + 63) / 64;
  using V = T __attribute__((ext_vector_type(16 / sizeof(T))));
  constexpr int VK = 16 / (int)sizeof(T), NV = (2 * TPW * N / VK + 63) / 64;
  const int Qh = (Wb - 1) / 2, ld = band_ld(Qh + 1);
  const cx<T>* src = TT + (long long)img * N * ld;
  const long long ibase = (long long)img * N * N;
  int trs[TPL], plo[TPL], phi[TPL], jj[TPL];
  bool vlo[TPL], vhi[TPL];
#pragma unroll
  for (int i = 0; i < TPL; ++i) {
    const int t = lane + 64 * i;
    const int tr = t / S, j = t - (t / S) * S;
    const bool ok = t < TASKS;
    trs[i] = ok ? tr : TPW;
    jj[i] = j;
    vlo[i] = ok && j <= Qh;      // nu = j >= 0: stored column j
    vhi[i] = ok && S - j <= Qh;  // nu = j - S < 0: conjugate of stored column S - j
    plo[i] = vlo[i] ? j : 0;
    phi[i] = vhi[i] ? S - j : 0;
  }
  cx<T> a0[TPL], a1[TPL], b0[TPL], b1[TPL];
  auto fetch = [&](int u) {  // any u < NU
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
      const int p = min(u * TPW + min(trs[i], TPW - 1), NTR - 1);
      const long long r0 = (long long)(2 * p) * ld;
      a0[i] = src[r0 + plo[i]];
      b0[i] = src[r0 + phi[i]];
      a1[i] = src[r0 + ld + plo[i]];
      b1[i] = src[r0 + ld + phi[i]];
    }
  };
  cx<T> lo[TPL], hi[TPL];
  auto take = [&]() {
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
      const cx<T> l = cx<T>{a0[i].re - a1[i].im, a0[i].im + a1[i].re};
      const cx<T> h = cx<T>{b0[i].re + b1[i].im, b1[i].re - b0[i].im};  // conj b0 + i conj b1
      lo[i] = vlo[i] ? scale(l, s) : cx<T>{0, 0};
      hi[i] = vhi[i] ? scale(h, s) : cx<T>{0, 0};
    }
  };
  const int stride = gridDim.x * NW;
  int u = blockIdx.x * NW + wave;
  fetch(min(u, NU - 1));
  take();
  TailTw<T, +1, N, TPW, 0> rt;  // twiddles from the roots table: registers go to the rows
  double cost = 0;
  int err = 0;
  T* rows = reinterpret_cast<T*>(wb);  // transform tr: rows 2 tr, 2 tr + 1 at tr * 2 PL
  for (; u < NU; u += stride) {
    fetch(min(u + stride, NU - 1));
    const int ntr = NTR % TPW == 0 ? TPW : min(TPW, NTR - u * TPW);
    const int nvu = 2 * ntr * N / VK;  // the unit's rows are contiguous in the image
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
      if (trs[i] < ntr) {
        cx<T>* w = wb + trs[i] * PL;
        cx<T>* w0 = w + pidx<T>(jj[i] * R);
        constexpr bool LIN = (1 << kPadShift<T>) % R == 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const cx<T> val = lo[i] + mul_root<R, +1, T>(hi[i], (R - 1) * k);
          if constexpr (LIN)
            w0[k] = val;
          else
            w[pidx<T>(jj[i] * R + k)] = val;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    auto put = [&](int tr, int x, cx<T> val) {
      rows[tr * 2 * PL + x] = val.re;
      rows[tr * 2 * PL + N + x] = val.im;
    };
    wave_fft_tail<T, +1, N, TPW, PL>(wb, tw, lane, ntr, put, rt);
    // the unit's target rows
    V tv[NV];
    const V* tsrc = reinterpret_cast<const V*>(target + ibase + (long long)(2 * u * TPW) * N);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int e = lane + 64 * k;
      tv[k] = e < nvu ? tsrc[e] : V{};
    }
    // the objective, in place: LDS row x of transform tr sits at tr * 2 PL + x,
    // the image row at (2 tr + {0, 1}) N + x: vector e covers both the same way
    // when each transform's two rows are whole vectors (2 N % VK == 0)
    V* rv = reinterpret_cast<V*>(rows);
    V* img_v = image ? reinterpret_cast<V*>(image + ibase + (long long)(2 * u * TPW) * N) : nullptr;
    V* d_v = dI ? reinterpret_cast<V*>(dI + ibase + (long long)(2 * u * TPW) * N) : nullptr;
    constexpr int VPT = 2 * N / VK;  // vectors per transform's row pair
#pragma unroll 2  // fully unrolled, the loop needs about 240 registers (spills at three waves)
    for (int k = 0; k < NV; ++k) {
      const int e = lane + 64 * k;
      if (e < nvu) {
        const int tr = e / VPT, q = e - tr * VPT;
        V x = rv[tr * (2 * PL / VK) + q];
        if (img_v) __builtin_nontemporal_store(x, img_v + e);
        V d;
#pragma unroll
        for (int c = 0; c < VK; ++c) {
          const T t = tv[k][c];
          const T v = rp.sign * (x[c] * rp.scale - rp.level);
          const T z = sigmoid_clipped(rp.a * v);
          const T r = z - t;
          cost += (double)(r * r);
          err += (v > (T)0) != (t > (T)0.5);
          d[c] = rp.dscale * r * z * ((T)1 - z);
        }
        rv[tr * (2 * PL / VK) + q] = d;
        if (d_v) __builtin_nontemporal_store(d, d_v + e);
      }
    }
    __builtin_amdgcn_wave_barrier();
    // forward transform of (dI even + i dI odd), band columns nu = 0..Q
    auto load = [&](int tr, int x) -> cx<T> {
      return cx<T>{rows[tr * 2 * PL + x], rows[tr * 2 * PL + N + x]};
    };
    auto store = [&](int tr, int x, cx<T> val) { wb[tr * PL + pidx<T>(x)] = val; };
    wave_fft<T, -1, N, TPW>(wb, tw, lane, ntr, load, store);
    const int y0 = 2 * u * TPW;
    cx<T>* out = RT + (long long)img * ld * N;
    for (int t = lane; t < (Qh + 1) * 2 * ntr; t += 64) {
      const int yy = t / (Qh + 1), nu = t - yy * (Qh + 1), tr = yy >> 1;
      const cx<T> a = wb[tr * PL + pidx<T>(nu)];
      const cx<T> b = conj(wb[tr * PL + pidx<T>(bin_of(-nu, N))]);
      cx<T> r;
      if ((yy & 1) == 0) {
        r = scale(a + b, (T)0.5);
      } else {
        const cx<T> d = a - b;
        r = cx<T>{d.im * (T)0.5, -d.re * (T)0.5};
      }
      out[(long long)(y0 + yy) * ld + nu] = r;
    }
    __builtin_amdgcn_wave_barrier();
    take();
  }
  // block partials in a fixed order: lanes (shuffle tree), then waves
  for (int o = 32; o > 0; o >>= 1) {
    cost += __shfl_xor(cost, o, 64);
    err += __shfl_xor(err, o, 64);
  }
  __syncthreads();  // every wave is done with its buffer
  double* pc = reinterpret_cast<double*>(tw + PL);
  int* pe = reinterpret_cast<int*>(pc + NW);
  if (lane == 0) {
    pc[wave] = cost;
    pe[wave] = err;
  }
  __syncthreads();
  if (tid == 0) {
    double c = 0;
    int e = 0;
    for (int w = 0; w < NW; ++w) {
      c += pc[w];
      e += pe[w];
    }
    part_cost[(long long)img * gridDim.x + blockIdx.x] = c;
    part_err[(long long)img * gridDim.x + blockIdx.x] = e;
  }
}

// ------------------------------------------------------- coarse image band
// One block per image: the sum of the image's coarse partials (the spread or
// segment layout of the coherent-system pass) and its band spectrum
//   Y[u][v] = s * sum_{y,x} I[y][x] e^{-2 pi i ((u - Q) y + (v - Q) x) / M},
// u, v < 2Q + 1, in one pass, the summed image never leaving the registers of
// the row transforms.  Rows are packed in pairs (even + i odd); since every
// row is real, its spectrum at -nu is the conjugate of that at nu, so only
// nu = 0..Q is extracted (RT, M x (Q + 1) in LDS) and column-transformed, and
// the columns v < Q are written as conjugate mirrors of v > Q.  The partials
// are added in the order the separate sum pass adds them, so the image is the
// same to the bit.
constexpr int kCoarseWaves = 6;
constexpr int kCoarseMaxParts = 64;

__host__ __device__ constexpr int coarse_ldr(int Q) { return (Q + 1) % 2 == 0 ? Q + 2 : Q + 1; }

template <typename T, int M>
__global__ __launch_bounds__(64 * kCoarseWaves) void k_coarse_band(
    const T* __restrict__ parts, int spread, int cnt, int spp, long long total, int blocks,
    int segs, int Q, T s, const cx<T>* __restrict__ twg, cx<T>* __restrict__ Y,
    T* __restrict__ peak) {
  constexpr int TPW = tpw<M>(), PL = padded<T, M>(), NT = 64 * kCoarseWaves;
  constexpr long long LEN = (long long)M * M;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  cx<T>* wb = tw + PL + wave * TPW * PL;
  const int LDR = coarse_ldr(Q);
  cx<T>* rt = tw + PL + kCoarseWaves * TPW * PL;  // M x LDR
  int* plist = reinterpret_cast<int*>(rt + M * LDR);
  const int g = blockIdx.x;
  if (tid == 0) {
    int n = 0;
    if (spread) {
      // the runs that overlap image g, each holding one segment of it
      const long long q0 = (long long)g * spp, q1 = min(total, q0 + spp);
      const int bf = spread_run_of(q0, total, blocks), bl = spread_run_of(q1 - 1, total, blocks);
      for (int b = bf; b <= bl; ++b) {
        const long long lo = (long long)b * total / blocks, hi = (long long)(b + 1) * total / blocks;
        if (lo >= hi) continue;
        const long long k = g - lo / spp;
        if (k >= 0 && k < segs && n < kCoarseMaxParts) plist[n++] = (int)(k * blocks + b);
      }
    } else {
      for (int c = 0; c < cnt && n < kCoarseMaxParts; ++c) plist[n++] = g * cnt + c;
    }
    plist[kCoarseMaxParts] = n;
    if (peak) peak[g] = 0;  // the synthesis rows' running maximum starts here
  }
  load_roots_padded(tw, twg, M, tid, NT);
  __syncthreads();
  const int np = plist[kCoarseMaxParts];
  // row pairs: forward transforms of (even + i odd), spectra at nu = 0..Q.
  // A unit's 2 TPW rows are contiguous in every partial: the wave sums them
  // with 16-byte loads (one round trip per partial), packs the sums into its
  // LDS buffer and transforms them there.
  using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
  constexpr int VE = 16 / (int)sizeof(T);  // values per 16-byte load
  constexpr int NPAIR = M / 2, NUR = (NPAIR + TPW - 1) / TPW;
  constexpr int VPL = (2 * TPW * M / VE + 63) / 64;  // loads per lane per partial
  static_assert(M % VE == 0, "rows are whole 16-byte runs");
  for (int u = wave; u < NUR; u += kCoarseWaves) {
    const int p0 = u * TPW, ntr = min(TPW, NPAIR - p0);
    const int nv = 2 * ntr * M / VE;
    V acc[VPL];
#pragma unroll
    for (int k = 0; k < VPL; ++k) acc[k] = V{};
    for (int i = 0; i < np; ++i) {
      const V* src = reinterpret_cast<const V*>(parts + (long long)plist[i] * LEN +
                                                (long long)(2 * p0) * M);
#pragma unroll
      for (int k = 0; k < VPL; ++k) {
        const int e = lane + 64 * k;
        if (e < nv) {
          const V a = src[e];
          if constexpr (VE == 4) {
            acc[k].x += a.x, acc[k].y += a.y, acc[k].z += a.z, acc[k].w += a.w;
          } else {
            acc[k].x += a.x, acc[k].y += a.y;
          }
        }
      }
    }
    T* wr = reinterpret_cast<T*>(wb);
#pragma unroll
    for (int k = 0; k < VPL; ++k) {
      const int e = lane + 64 * k;
      if (e < nv) {
        T vals[VE];
        if constexpr (VE == 4) {
          vals[0] = acc[k].x, vals[1] = acc[k].y, vals[2] = acc[k].z, vals[3] = acc[k].w;
        } else {
          vals[0] = acc[k].x, vals[1] = acc[k].y;
        }
#pragma unroll
        for (int c = 0; c < VE; ++c) {
          const int f = e * VE + c, rr = f / M, x = f - rr * M;
          wr[2 * ((rr >> 1) * PL + pidx<T>(x)) + (rr & 1)] = vals[c];
        }
      }
    }
    auto load = [&](int tr, int x) -> cx<T> { return wb[tr * PL + pidx<T>(x)]; };
    auto store = [&](int tr, int x, cx<T> v) { wb[tr * PL + pidx<T>(x)] = v; };
    wave_fft<T, -1, M, TPW>(wb, tw, lane, ntr, load, store);
    for (int t = lane; t < 2 * ntr * (Q + 1); t += 64) {
      const int yy = t / (Q + 1), nu = t - yy * (Q + 1), tr = yy >> 1;
      const cx<T> a = wb[tr * PL + pidx<T>(nu)];
      const cx<T> b = conj(wb[tr * PL + pidx<T>(bin_of(-nu, M))]);
      cx<T> r;
      if ((yy & 1) == 0) {
        r = scale(a + b, (T)0.5);  // transform of the even row
      } else {
        const cx<T> d = a - b;     // (a - b) / (2i)
        r = cx<T>{d.im * (T)0.5, -d.re * (T)0.5};
      }
      rt[(2 * p0 + yy) * LDR + nu] = r;
    }
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // columns nu_v = 0..Q: forward transforms along y, band rows kept
  const int Wb = 2 * Q + 1, ncol = Q + 1, NUC = (ncol + TPW - 1) / TPW;
  cx<T>* out = Y + (long long)g * Wb * Wb;
  for (int u = wave; u < NUC; u += kCoarseWaves) {
    const int c0 = u * TPW, ntr = min(TPW, ncol - c0);
    auto load = [&](int tr, int y) -> cx<T> { return rt[y * LDR + c0 + tr]; };
    auto store = [&](int tr, int b, cx<T> val) {
      const int nu = nu_of_bin(b, M);
      if (nu < -Q || nu > Q) return;
      const int nv = c0 + tr;
      const cx<T> y = scale(val, s);
      out[(long long)(nu + Q) * Wb + (nv + Q)] = y;
      if (nv > 0) out[(long long)(Q - nu) * Wb + (Q - nv)] = conj(y);
    };
    wave_fft<T, -1, M, TPW>(wb, tw, lane, ntr, load, store);
  }
}

// ---------------------------------------------------------------- launchers
template <typename T, int N>
static size_t smem_bytes() {
  return sizeof(cx<T>) * (size_t)(padded<T, N>() * (1 + row_waves<T, N>() * tpw<N>()));
}

template <typename T, int N>
static size_t col_smem_bytes() {
  static_assert(sizeof(cx<T>) * (padded<T, N>() + cpb<T, N>() * col_ld<T, N>()) <= 160 * 1024,
                "column block exceeds the LDS");
  return sizeof(cx<T>) * (size_t)(padded<T, N>() + cpb<T, N>() * col_ld<T, N>());
}

template <typename K>
static hipError_t set_smem(K kern, size_t bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// Persistent row kernels: units per wavefront, and the wavefront count
// (about 8 resident per SIMD over 256 CUs) below which fewer units per
// wavefront are used.  The synthesis rows take 2 units per wavefront: with
// the half band (about 40 stored columns) their reads are light, and short
// blocks balance the store-bound tail better than long loops (a 320-image
// batch at 1024^2: the whole forward 0.962 ms at 2, 0.973 at 4, 0.988 at 8,
// 1.012 at 1).
constexpr int kUnitsPerWave = 4;
constexpr int kRowWavesTarget = 8192;
constexpr int kSynthUnitsPerWave = 2;
constexpr int kSynthWavesTarget = 1024;

static unsigned blocks_for(int lines, int per_block) {
  return (unsigned)((lines + per_block - 1) / per_block);
}

template <typename T, int N>
static hipError_t analysis_n(const void* in, int cplx, long long img_stride,
                             long long slice_stride, int nsum, int nimg, int nu0, int Wb, T s,
                             cx<T>* RT, cx<T>* X, hipStream_t st) {
  constexpr int TPB = row_waves<T, N>() * tpw<N>();
  constexpr int NT = 64 * row_waves<T, N>();
  const size_t sm = smem_bytes<T, N>();
  const cx<T>* twg = nullptr;
  hipError_t e = table_for<T>(N, &twg);
  if (e != hipSuccess) return e;
  bool done = false, half = false;
  if constexpr (Plan<N>::R1 > 1) {
    if (nsum == 1) {
      const int units = ((cplx ? N : N / 2) + tpw<N>() - 1) / tpw<N>();
      const int nw = ap_waves<T, N>();
      const int upw = (int)std::min<long long>(
          kUnitsPerWave, std::max<long long>(1, (long long)units * nimg / (nw * kRowWavesTarget)));
      const unsigned gx = (unsigned)((units + nw * upw - 1) / (nw * upw));
      // a real plane over a band symmetric about 0: half the columns
      half = !cplx && Wb % 2 == 1 && 2 * nu0 + Wb - 1 == 0;
      auto k = cplx   ? k_analysis_rows_p<T, N, true>
               : half ? k_analysis_rows_p<T, N, false, false, true>
                      : k_analysis_rows_p<T, N, false>;
      const size_t smp = ap_smem_bytes<T, N>();
      if ((e = set_smem(k, smp)) != hipSuccess) return e;
      hipLaunchKernelGGL(k, dim3(gx, nimg), dim3(64 * nw), smp, st, (const T*)in, img_stride,
                         nu0, Wb, twg, RT, (T)0);
      done = true;
    }
  }
  if (done) {
  } else if (cplx) {
    auto k = k_analysis_rows<T, N, true>;
    if ((e = set_smem(k, sm)) != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(blocks_for(N, TPB), nimg), dim3(NT), sm, st,
                       (const T*)in, img_stride, slice_stride, nsum, nu0, Wb, twg, RT);
  } else {
    auto k = k_analysis_rows<T, N, false>;
    if ((e = set_smem(k, sm)) != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(blocks_for(N / 2, TPB), nimg), dim3(NT), sm, st,
                       (const T*)in, img_stride, slice_stride, nsum, nu0, Wb, twg, RT);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  auto k2 = half ? k_analysis_cols<T, N, true> : k_analysis_cols<T, N>;
  const size_t smc = col_smem_bytes<T, N>();
  if ((e = set_smem(k2, smc)) != hipSuccess) return e;
  hipLaunchKernelGGL(k2, dim3(blocks_for(half ? (Wb + 1) / 2 : Wb, cpb<T, N>()), nimg),
                     dim3(64 * col_waves<T, N>()), smc, st, (const cx<T>*)RT, nu0, Wb, s, twg, X);
  return hipGetLastError();
}

template <typename T, int N>
static hipError_t analysis_sigmoid_n(const T* theta, int nimg, int nu0, int Wb, T s, T beta,
                                     cx<T>* RT, cx<T>* X, hipStream_t st) {
  if constexpr (Plan<N>::R1 > 1) {
    const cx<T>* twg = nullptr;
    hipError_t e = table_for<T>(N, &twg);
    if (e != hipSuccess) return e;
    const int units = (N / 2 + tpw<N>() - 1) / tpw<N>();
    const int nw = ap_waves<T, N>();
    const int upw = (int)std::min<long long>(
        kUnitsPerWave, std::max<long long>(1, (long long)units * nimg / (nw * kRowWavesTarget)));
    const unsigned gx = (unsigned)((units + nw * upw - 1) / (nw * upw));
    const bool half = Wb % 2 == 1 && 2 * nu0 + Wb - 1 == 0;
    auto k = half ? k_analysis_rows_p<T, N, false, true, true> : k_analysis_rows_p<T, N, false, true>;
    const size_t smp = ap_smem_bytes<T, N>();
    if ((e = set_smem(k, smp)) != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(gx, nimg), dim3(64 * nw), smp, st, theta, (long long)N * N, nu0,
                       Wb, twg, RT, beta);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    auto k2 = half ? k_analysis_cols<T, N, true> : k_analysis_cols<T, N>;
    const size_t smc = col_smem_bytes<T, N>();
    if ((e = set_smem(k2, smc)) != hipSuccess) return e;
    hipLaunchKernelGGL(k2, dim3(blocks_for(half ? (Wb + 1) / 2 : Wb, cpb<T, N>()), nimg),
                       dim3(64 * col_waves<T, N>()), smc, st, (const cx<T>*)RT, nu0, Wb, s, twg, X);
    return hipGetLastError();
  } else {
    (void)theta, (void)nimg, (void)nu0, (void)Wb, (void)s, (void)beta, (void)RT, (void)X,
        (void)st;
    return hipErrorNotSupported;
  }
}

template <typename T, int N>
static hipError_t synthesis_n(const cx<T>* Y, int nimg, int nu0, int Wb, int herm, T s, int cout,
                              cx<T>* TT, T* out, T* peak, hipStream_t st) {
  constexpr int TPB = row_waves<T, N>() * tpw<N>();
  constexpr int NT = 64 * row_waves<T, N>();
  const size_t sm = smem_bytes<T, N>();
  const cx<T>* twg = nullptr;
  hipError_t e = table_for<T>(N, &twg);
  if (e != hipSuccess) return e;
  // the band lies within |nu| < N / R0: sparse first radix stage
  constexpr int S0 = N / Plan<N>::R0;
  const bool sp = nu0 > -S0 && nu0 + Wb - 1 < S0;
  // real output of a Hermitian band symmetric about 0: half the columns
  const bool half = herm && !cout && Wb % 2 == 1 && 2 * nu0 + Wb - 1 == 0;
  auto k1 = half ? (sp ? k_synth_cols<T, N, true, true> : k_synth_cols<T, N, false, true>)
                 : (sp ? k_synth_cols<T, N, true> : k_synth_cols<T, N, false>);
  const size_t smc = col_smem_bytes<T, N>();
  if ((e = set_smem(k1, smc)) != hipSuccess) return e;
  hipLaunchKernelGGL(k1, dim3(blocks_for(half ? (Wb + 1) / 2 : Wb, cpb<T, N>()), nimg),
                     dim3(64 * col_waves<T, N>()), smc, st, Y, nu0, Wb, herm, twg, TT);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if constexpr (Plan<N>::R1 > 1) {
    if (sp) {
      // persistent rows: about kUnitsPerWave units per wavefront, fewer when
      // the batch is too small to fill the device
      const int units = ((cout ? N : N / 2) + tpw<N>() - 1) / tpw<N>();
      const int nw = row_waves<T, N>();
      int upw = (int)std::min<long long>(
          kSynthUnitsPerWave,
          std::max<long long>(1, (long long)units * nimg / (nw * kSynthWavesTarget)));
      const unsigned gx = (unsigned)((units + nw * upw - 1) / (nw * upw));
      // staging pays once the batch gives every wavefront its full share of
      // units (a store-bound pass); small batches are latency-bound and skip
      // the extra LDS round trip (16 images at 1024^2: 15 us direct, 20 staged)
      const bool staged = !cout && ((uintptr_t)out & 15) == 0 && upw == kSynthUnitsPerWave;
      auto k2 = cout     ? k_synth_rows_sp<T, N, true>
                : staged ? (half ? k_synth_rows_sp<T, N, false, true, true>
                                 : k_synth_rows_sp<T, N, false, false, true>)
                         : (half ? k_synth_rows_sp<T, N, false, true> : k_synth_rows_sp<T, N, false>);
      if ((e = set_smem(k2, sm)) != hipSuccess) return e;
      hipLaunchKernelGGL(k2, dim3(gx, nimg), dim3(NT), sm, st, (const cx<T>*)TT, nu0, Wb, s, twg,
                         out, cout ? (T*)nullptr : peak, ThetaEpi<T>{});
      return hipGetLastError();
    }
  }
  if (cout) {
    auto k2 = sp ? k_synth_rows<T, N, true, true> : k_synth_rows<T, N, true, false>;
    if ((e = set_smem(k2, sm)) != hipSuccess) return e;
    hipLaunchKernelGGL(k2, dim3(blocks_for(N, TPB), nimg), dim3(NT), sm, st,
                       (const cx<T>*)TT, nu0, Wb, s, twg, out, (T*)nullptr);
  } else {
    // a batch of 64 MB or more of images is not read back from the caches:
    // streaming stores (SOCS, 256 images at 512^2: 1.11 -> 1.08 ms per forward;
    // a single small image read back at once by an ILT step is faster cached)
    const bool stream = (long long)nimg * N * N * (long long)sizeof(T) >= (64LL << 20);
    auto k2 = half ? (sp ? (stream ? k_synth_rows<T, N, false, true, true, true>
                                   : k_synth_rows<T, N, false, true, true>)
                         : (stream ? k_synth_rows<T, N, false, false, true, true>
                                   : k_synth_rows<T, N, false, false, true>))
                   : (sp ? (stream ? k_synth_rows<T, N, false, true, false, true>
                                   : k_synth_rows<T, N, false, true>)
                         : (stream ? k_synth_rows<T, N, false, false, false, true>
                                   : k_synth_rows<T, N, false, false>));
    if ((e = set_smem(k2, sm)) != hipSuccess) return e;
    hipLaunchKernelGGL(k2, dim3(blocks_for(N / 2, TPB), nimg), dim3(NT), sm, st,
                       (const cx<T>*)TT, nu0, Wb, s, twg, out, peak);
  }
  return hipGetLastError();
}

template <typename T, int M>
static hipError_t coarse_band_n(const T* parts, int spread, int cnt, int spp, long long total,
                                int blocks, int segs, int nimg, int Q, T s, cx<T>* Y, T* peak,
                                hipStream_t st) {
  const size_t sm = sizeof(cx<T>) * ((size_t)padded<T, M>() * (1 + kCoarseWaves * tpw<M>()) +
                                     (size_t)M * coarse_ldr(Q)) +
                    sizeof(int) * (kCoarseMaxParts + 1);
  if (sm > 160 * 1024) return hipErrorNotSupported;
  const cx<T>* twg = nullptr;
  hipError_t e = table_for<T>(M, &twg);
  if (e != hipSuccess) return e;
  auto k = k_coarse_band<T, M>;
  if ((e = set_smem(k, sm)) != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3((unsigned)nimg), dim3(64 * kCoarseWaves), sm, st, parts, spread, cnt,
                     spp, total, blocks, segs, Q, s, twg, Y, peak);
  return hipGetLastError();
}

// Real-output synthesis of a mask gradient over a symmetric half band with the
// theta epilogue (k_synth_rows_sp<..., EPI>).
template <typename T, int N>
static hipError_t synthesis_theta_n(const cx<T>* Y, int nimg, int nu0, int Wb, T s, cx<T>* TT,
                                    ThetaEpi<T> te, T* grad_theta, hipStream_t st) {
  if constexpr (Plan<N>::R1 > 1) {
    constexpr int S0 = N / Plan<N>::R0, VK = 16 / (int)sizeof(T);
    const bool sp = nu0 > -S0 && nu0 + Wb - 1 < S0;
    const bool half = Wb % 2 == 1 && 2 * nu0 + Wb - 1 == 0;
    auto al = [](const void* q) { return ((uintptr_t)q & 15) == 0; };
    if (!sp || !half || N % VK != 0 || !al(te.theta) || !al(grad_theta))
      return hipErrorNotSupported;
    const cx<T>* twg = nullptr;
    hipError_t e = table_for<T>(N, &twg);
    if (e != hipSuccess) return e;
    auto k1 = k_synth_cols<T, N, true, true>;
    const size_t smc = col_smem_bytes<T, N>();
    if ((e = set_smem(k1, smc)) != hipSuccess) return e;
    hipLaunchKernelGGL(k1, dim3(blocks_for((Wb + 1) / 2, cpb<T, N>()), nimg),
                       dim3(64 * col_waves<T, N>()), smc, st, Y, nu0, Wb, 1, twg, TT);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int units = (N / 2 + tpw<N>() - 1) / tpw<N>();
    const int nw = row_waves<T, N>();
    const int upw = (int)std::min<long long>(
        kSynthUnitsPerWave,
        std::max<long long>(1, (long long)units * nimg / (nw * kSynthWavesTarget)));
    const unsigned gx = (unsigned)((units + nw * upw - 1) / (nw * upw));
    auto k2 = k_synth_rows_sp<T, N, false, true, true, true>;
    const size_t sm = smem_bytes<T, N>();
    if ((e = set_smem(k2, sm)) != hipSuccess) return e;
    hipLaunchKernelGGL(k2, dim3(gx, nimg), dim3(64 * nw), sm, st, (const cx<T>*)TT, nu0, Wb, s,
                       twg, grad_theta, (T*)nullptr, te);
    return hipGetLastError();
  } else {
    (void)Y, (void)nimg, (void)nu0, (void)Wb, (void)s, (void)TT, (void)te, (void)grad_theta,
        (void)st;
    return hipErrorNotSupported;
  }
}

// Worst-case blocks per image of the fused synthesis-objective rows (one unit
// per wave), for sizing the cost partials.
template <int N>
static int synth_objective_blocks() {
  const int units = (N / 2 + tpw<N>() - 1) / tpw<N>();
  return (units + kObjRowWaves - 1) / kObjRowWaves;
}

template <typename T, int N>
static hipError_t synth_objective_n(const cx<T>* Y, int nimg, int nu0, int Wb, cx<T>* TT,
                                    const T* target, ResistPix<T> rp, T* image, T* dI,
                                    cx<T>* RT, cx<T>* GY, double* part_cost, int* part_err,
                                    T* cost, T* total, int* mismatch, hipStream_t st) {
  if constexpr (Plan<N>::R1 > 1) {
    constexpr int S0 = N / Plan<N>::R0, VK = 16 / (int)sizeof(T);
    const bool sp = nu0 > -S0 && nu0 + Wb - 1 < S0;
    const bool half = Wb % 2 == 1 && 2 * nu0 + Wb - 1 == 0;
    auto al = [](const void* q) { return ((uintptr_t)q & 15) == 0; };
    if (!sp || !half || N % VK != 0 || !al(target) || (image && !al(image)) || (dI && !al(dI)))
      return hipErrorNotSupported;
    const cx<T>* twg = nullptr;
    hipError_t e = table_for<T>(N, &twg);
    if (e != hipSuccess) return e;
    auto k1 = k_synth_cols<T, N, true, true>;
    const size_t smc = col_smem_bytes<T, N>();
    if ((e = set_smem(k1, smc)) != hipSuccess) return e;
    hipLaunchKernelGGL(k1, dim3(blocks_for((Wb + 1) / 2, cpb<T, N>()), nimg),
                       dim3(64 * col_waves<T, N>()), smc, st, Y, nu0, Wb, 1, twg, TT);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int units = (N / 2 + tpw<N>() - 1) / tpw<N>();
    const int upw = (int)std::min<long long>(
        kSynthUnitsPerWave,
        std::max<long long>(1, (long long)units * nimg / (kObjRowWaves * kSynthWavesTarget)));
    const unsigned gx = (unsigned)((units + kObjRowWaves * upw - 1) / (kObjRowWaves * upw));
    auto k2 = k_synth_objective_rows<T, N>;
    const size_t sm = sizeof(cx<T>) * (size_t)padded<T, N>() * (1 + kObjRowWaves * tpw<N>());
    if ((e = set_smem(k2, sm)) != hipSuccess) return e;
    hipLaunchKernelGGL(k2, dim3(gx, nimg), dim3(64 * kObjRowWaves), sm, st, (const cx<T>*)TT, Wb,
                       (T)1, twg, target, rp, image, dI, RT, part_cost, part_err);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = resist_final<T>(part_cost, part_err, nimg, (int)gx, (long long)N * N,
                             part_cost + (long long)nimg * synth_objective_blocks<N>(), cost,
                             total, mismatch, st)) != hipSuccess)
      return e;
    auto k3 = k_analysis_cols<T, N, true>;
    if ((e = set_smem(k3, smc)) != hipSuccess) return e;
    hipLaunchKernelGGL(k3, dim3(blocks_for((Wb + 1) / 2, cpb<T, N>()), nimg),
                       dim3(64 * col_waves<T, N>()), smc, st, (const cx<T>*)RT, nu0, Wb, (T)1, twg,
                       GY);
    return hipGetLastError();
  } else {
    (void)Y, (void)nimg, (void)nu0, (void)Wb, (void)TT, (void)target, (void)rp, (void)image,
        (void)dI, (void)RT, (void)GY, (void)part_cost, (void)part_err, (void)cost, (void)total,
        (void)mismatch, (void)st;
    return hipErrorNotSupported;
  }
}

#define LB_BAND_SIZES(X) \
  X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(48) X(96) X(192) X(384) X(768) \
      X(1536) X(80)

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
  if (!band_size_supported(N) || Wb > N)
    return band_analysis_any<T>(N, in, cplx, img_stride, slice_stride, nsum, nimg, nu0, Wb, s,
                                RT, X, st);
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
  if (!band_size_supported(N) || Wb > N)
    return band_synthesis_any<T>(N, Y, nimg, nu0, Wb, herm, s, cout, TT, out, peak, st);
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

template <typename T>
hipError_t band_analysis_sigmoid(int N, const T* theta, int nimg, int nu0, int Wb, T s, T beta,
                                 cx<T>* RT, cx<T>* X, hipStream_t st) {
  if (!band_size_supported(N) || Wb > N) return hipErrorNotSupported;
  switch (N) {
#define LB_CASE(n) \
  case n:          \
    return analysis_sigmoid_n<T, n>(theta, nimg, nu0, Wb, s, beta, RT, X, st);
    LB_BAND_SIZES(LB_CASE)
#undef LB_CASE
    default:
      return hipErrorNotSupported;
  }
}

template <typename T>
hipError_t coarse_band(int M, const T* parts, int spread, int cnt, int spp, long long total,
                       int blocks, int segs, int nimg, int Q, T s, cx<T>* Y, T* peak,
                       hipStream_t st) {
  if (nimg <= 0) return hipSuccess;
  if (Q < 0 || 2 * Q + 1 > M) return hipErrorNotSupported;
  // every image's partials must fit the block's list: a spread image meets at
  // most (spp - 1) * blocks / total + 2 runs
  const long long most = spread ? (long long)(spp - 1) * blocks / total + 2 : cnt;
  if (most > kCoarseMaxParts) return hipErrorNotSupported;
  switch (M) {
#define LB_CASE(m) \
  case m:          \
    return coarse_band_n<T, m>(parts, spread, cnt, spp, total, blocks, segs, nimg, Q, s, Y, peak, \
                               st);
    LB_CASE(16) LB_CASE(32) LB_CASE(48) LB_CASE(64) LB_CASE(80) LB_CASE(96) LB_CASE(128)
        LB_CASE(256)
#undef LB_CASE
    default:
      return hipErrorNotSupported;
  }
}

template hipError_t coarse_band<float>(int, const float*, int, int, int, long long, int, int, int,
                                       int, float, cx<float>*, float*, hipStream_t);
template hipError_t coarse_band<double>(int, const double*, int, int, int, long long, int, int,
                                        int, int, double, cx<double>*, double*, hipStream_t);
template <typename T>
hipError_t band_synthesis_theta(int N, const cx<T>* Y, int nimg, int nu0, int Wb, T s, cx<T>* TT,
                                ThetaEpi<T> te, T* grad_theta, hipStream_t st) {
  if (!band_size_supported(N) || Wb > N) return hipErrorNotSupported;
  switch (N) {
#define LB_CASE(n) \
  case n:          \
    return synthesis_theta_n<T, n>(Y, nimg, nu0, Wb, s, TT, te, grad_theta, st);
    LB_BAND_SIZES(LB_CASE)
#undef LB_CASE
    default:
      return hipErrorNotSupported;
  }
}

template hipError_t band_synthesis_theta<float>(int, const cx<float>*, int, int, int, float,
                                                cx<float>*, ThetaEpi<float>, float*, hipStream_t);
template hipError_t band_synthesis_theta<double>(int, const cx<double>*, int, int, int, double,
                                                 cx<double>*, ThetaEpi<double>, double*,
                                                 hipStream_t);

bool synth_objective_supported(int N, int nu0, int Wb) {
  if (!band_size_supported(N) || Wb > N || Wb % 2 == 0 || 2 * nu0 + Wb - 1 != 0) return false;
  switch (N) {
#define LB_CASE(n) \
  case n:          \
    return Plan<n>::R1 > 1 && nu0 > -(n / Plan<n>::R0) && nu0 + Wb - 1 < n / Plan<n>::R0;
    LB_BAND_SIZES(LB_CASE)
#undef LB_CASE
    default:
      return false;
  }
}

int synth_objective_part_count(int N, int nimg) {
  switch (N) {
#define LB_CASE(n) \
  case n:          \
    return nimg * (synth_objective_blocks<n>() + 1);
    LB_BAND_SIZES(LB_CASE)
#undef LB_CASE
    default:
      return 0;
  }
}

template <typename T>
hipError_t synth_objective(int N, const cx<T>* Y, int nimg, int nu0, int Wb, cx<T>* TT,
                           const T* target, ResistPix<T> rp, T* image, T* dI, cx<T>* RT,
                           cx<T>* GY, double* part_cost, int* part_err, T* cost, T* total,
                           int* mismatch, hipStream_t st) {
  if (!band_size_supported(N) || Wb > N) return hipErrorNotSupported;
  switch (N) {
#define LB_CASE(n)                                                                            \
  case n:                                                                                     \
    return synth_objective_n<T, n>(Y, nimg, nu0, Wb, TT, target, rp, image, dI, RT, GY,       \
                                   part_cost, part_err, cost, total, mismatch, st);
    LB_BAND_SIZES(LB_CASE)
#undef LB_CASE
    default:
      return hipErrorNotSupported;
  }
}

template hipError_t synth_objective<float>(int, const cx<float>*, int, int, int, cx<float>*,
                                           const float*, ResistPix<float>, float*, float*,
                                           cx<float>*, cx<float>*, double*, int*, float*, float*,
                                           int*, hipStream_t);
template hipError_t synth_objective<double>(int, const cx<double>*, int, int, int, cx<double>*,
                                            const double*, ResistPix<double>, double*, double*,
                                            cx<double>*, cx<double>*, double*, int*, double*,
                                            double*, int*, hipStream_t);
template hipError_t band_analysis_sigmoid<float>(int, const float*, int, int, int, float, float,
                                                 cx<float>*, cx<float>*, hipStream_t);
template hipError_t band_analysis_sigmoid<double>(int, const double*, int, int, int, double,
                                                  double, cx<double>*, cx<double>*, hipStream_t);
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
