// Coherent-system kernels: the Abbe source-point sum (one system per source
// point x plane x field component) and the SOCS sum (one system per kernel).
//
// A system j images the band spectrum X of a mask through its transfer table
// H_j, held on a W x W box of frequencies starting at box[j]:
//
//     C[a][b]   = X[box_y + a][box_x + b] * H_j[a][b]
//     E[y'][x'] = sum_{a,b < W} C[a][b] e^{+2 pi i (a y' + b x') / M}
//     I[y'][x'] += w_j |E[y'][x']|^2
//
// on an M x M grid with M >= 2W - 1.  |E|^2 only holds frequency differences
// inside the box, so M >= 2W - 1 samples it without aliasing and the fine
// N x N image is its exact band-limited interpolation (done by the band
// synthesis afterwards); the box-origin phase ramp is a unit-modulus factor
// and cancels in |E|^2.  When M == N the coarse grid is the image grid.
//
// One block owns one work item (a chunk of systems of one image) and loops
// over its systems; the intensity accumulates in registers, so each item
// writes its M x M partial once.  Per system: the row transforms (W rows,
// input pruned to W points) run in blocks of CPR rows through LDS into the
// W x (M/CSPLIT) table Tb; the column transforms (input pruned to W rows) run
// in rounds of CPR columns; the last radix stage's outputs are squared and
// accumulated in place.  Both 1-D transforms are two Stockham stages,
// M = R1 * R2, with the R1 stage first.
//
// The adjoint kernel is the exact transpose of that linear chain: it
// recomputes E, multiplies by the coarse sensitivity gM, and runs the
// transposed stages (forward DFTs, conjugate twiddles, outputs pruned to the
// box) to get G_C = 2 w FFT_M(gM E) on the box, then accumulates
// G_C * conj(H) into the item's band-gradient partial.
#include "lb_kernels.h"

namespace lb {

template <int M>
struct FieldRadix;
template <>
struct FieldRadix<16> {
  static constexpr int R1 = 4, R2 = 4;
};
template <>
struct FieldRadix<32> {
  static constexpr int R1 = 8, R2 = 4;
};
template <>
struct FieldRadix<48> {
  static constexpr int R1 = 12, R2 = 4;
};
template <>
struct FieldRadix<80> {
  static constexpr int R1 = 10, R2 = 8;
};
template <>
struct FieldRadix<96> {
  static constexpr int R1 = 12, R2 = 8;
};
template <>
struct FieldRadix<64> {
  static constexpr int R1 = 8, R2 = 8;
};
template <>
struct FieldRadix<128> {
  static constexpr int R1 = 16, R2 = 8;
};
template <>
struct FieldRadix<256> {
  static constexpr int R1 = 16, R2 = 16;
};

template <typename T, int M, int NT, int CSPLIT>
struct FieldShape {
  static constexpr int R1 = FieldRadix<M>::R1, R2 = FieldRadix<M>::R2;
  static constexpr int T1 = M / R1;  // stage-1 tasks per line
  static constexpr int T2 = M / R2;  // stage-2 tasks per line (== R1)
  static constexpr int CPR = NT / T1;
  static constexpr int MC = M / CSPLIT;
  static constexpr int ROUNDS = MC / CPR;
  static constexpr int NT2 = CPR * T2;             // stage-2 tasks per round
  static constexpr int Q2 = (NT2 + NT - 1) / NT;   // per thread (the last pass may idle)
  static constexpr int NACC = ROUNDS * Q2 * R2;
  static constexpr int LDC = M + 1;  // padded line stride of the round buffer
  // stage-1 slots r >= NRW address box rows / columns >= (M+1)/2 >= W: always zero
  static constexpr int NRW = ((M + 1) / 2 + T1 - 1) / T1 < R1 ? ((M + 1) / 2 + T1 - 1) / T1 : R1;
  static_assert(T2 == R1, "two-stage plan");
  static_assert(CPR * T1 == NT, "stage 1 maps one task per thread");
  static_assert(CPR <= MC && MC % CPR == 0, "rounds tile the split");
  // TWR: the row passes' stage-2 twiddles e^{2 pi i r j2 / M} come from an
  // (R2 - 1) x R1 table by (r, j2) rather than from the roots table at r j2:
  // the R1 tasks of a row read R1 distinct entries there, while r j2 puts
  // several of them on one bank (up to 8 at M = 128).  On at M = 96 and 128.
  // The table (672 / 896 B in FP32) leaves the resident block count of the
  // 96 x 256-thread and 128 x 512-thread kernels unchanged at every box
  // width; a few other shapes lose a block at some widths (128 x 256 threads
  // x 4 splits at W = 41-43, 96 x 128 threads at W = 34, several FP64 ones),
  // which the autotuner's timings include.
  static constexpr bool TWR = M == 96 || M == 128;
  static constexpr int TWR_N = TWR ? (R2 - 1) * R1 : 0;
};

// The by-(r, j2) table of a shape with TWR, filled from the global roots.
template <typename F, typename T>
__device__ __forceinline__ void load_row_tw(cx<T>* twr, const cx<T>* __restrict__ twg, int tid,
                                            int nt) {
  for (int i = tid; i < F::TWR_N; i += nt) twr[i] = twg[(i / F::R1 + 1) * (i % F::R1)];
}

// Row-pass stage-2 twiddle for r >= 1, j2 < R1.
template <typename F, typename T>
__device__ __forceinline__ cx<T> row_tw(const cx<T>* tw, const cx<T>* twr, int r, int j2) {
  if constexpr (F::TWR)
    return twr[(r - 1) * F::R1 + j2];
  else
    return tw[r * j2];
}

// Row-pass scratch layout per coarse grid: stage-1 task j's R1 outputs at
// j * SP, rows LDR apart (SP >= R1, LDR >= T1 * SP), chosen so that a wave's
// stage-1 writes and stage-2 reads spread over the LDS banks (unpadded, a row
// with R1 = 16 puts its stage-1 tasks on two banks, and rows M + 1 apart
// overlap the stage-2 reads of neighbouring rows).  Element x of a row, in
// stage-1 output order x = j * R1 + r, sits at rpos(x).
template <int M>
struct RowLayout;
template <>
struct RowLayout<16> {
  static constexpr int SP = 5, LDR = 20;
  static constexpr bool SWZ = false;
};
template <>
struct RowLayout<32> {
  static constexpr int SP = 9, LDR = 36;
  static constexpr bool SWZ = false;
};
template <>
struct RowLayout<48> {
  static constexpr int SP = 13, LDR = 60;
  static constexpr bool SWZ = false;
};
template <>
struct RowLayout<64> {
  static constexpr int SP = 9, LDR = 72;
  static constexpr bool SWZ = false;
};
template <>
struct RowLayout<80> {
  static constexpr int SP = 10, LDR = 81;
  static constexpr bool SWZ = false;
};
template <>
struct RowLayout<96> {
  static constexpr int SP = 13, LDR = 108;
  static constexpr bool SWZ = true;
};
template <>
struct RowLayout<128> {
  static constexpr int SP = 17, LDR = 136;
  static constexpr bool SWZ = false;
};
template <>
struct RowLayout<256> {
  static constexpr int SP = 17, LDR = 272;
  static constexpr bool SWZ = false;
};

// Row taken by thread group g = tid / T1 in the stage-1 mapping (T1 threads
// per row).  The two rows that share a 16-lane half of a 64-bit LDS access
// write {a LDR + j SP} for j < T1; with SWZ those rows are two apart rather
// than adjacent (bits 0 and 1 of g swapped), which at M = 96 (LDR = 108,
// SP = 13, T1 = 8) puts their writes on disjoint banks where adjacent rows
// overlap in half of them.
template <int M>
__device__ __forceinline__ int stage1_row(int g) {
  if constexpr (RowLayout<M>::SWZ)
    return (g & ~3) | ((g & 1) << 1) | ((g >> 1) & 1);
  else
    return g;
}

template <int M, int R1>
__device__ __forceinline__ int rpos(int x) {
  return (x / R1) * RowLayout<M>::SP + x % R1;
}

__device__ __forceinline__ int wrap_idx(int i, int n) {
  if (n <= 0) return i;
  int r = i % n;
  return r < 0 ? r + n : r;
}

// Row pass of one system for one split: W rows of C, inverse length-M
// transforms, columns of the split stored into Tb (W x MC).
template <typename T, int M, int NT, int CSPLIT>
__device__ __forceinline__ void row_pass(const cx<T>* __restrict__ Xm, int Wb, int wrapN,
                                         const cx<T>* __restrict__ Hj, int W, int2 bx,
                                         const cx<T>* tw, cx<T>* S, cx<T>* Tb, int cs, int tid,
                                         const cx<T>* twr) {
  using F = FieldShape<T, M, NT, CSPLIT>;
  using RL = RowLayout<M>;
  const int rl1 = stage1_row<M>(tid / F::T1), j1 = tid - (tid / F::T1) * F::T1;
  for (int a0 = 0; a0 < W; a0 += F::CPR) {
    {
      const int a = a0 + rl1;
      if (a < W) {
        cx<T> v[F::R1];
        const int xr = wrap_idx(bx.x + a, wrapN);
#pragma unroll
        for (int r = 0; r < F::R1; ++r) {
          const int b = j1 + r * F::T1;
          if (r < F::NRW && b < W) {
            const cx<T> x = Xm[(long long)xr * Wb + wrap_idx(bx.y + b, wrapN)];
            v[r] = x * Hj[a * W + b];
          } else {
            v[r] = cx<T>{0, 0};
          }
        }
        dft<F::R1, +1>(v);
#pragma unroll
        for (int r = 0; r < F::R1; ++r) S[rl1 * RL::LDR + j1 * RL::SP + r] = v[r];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < F::Q2; ++q) {
      const int t = tid + q * NT;
      const int rl = t / F::T2, j2 = t - (t / F::T2) * F::T2;
      const int a = a0 + rl;
      if (t < F::NT2 && a < W) {
        cx<T> v[F::R2];
#pragma unroll
        for (int r = 0; r < F::R2; ++r) {
          v[r] = S[rl * RL::LDR + r * RL::SP + j2];
          if (r > 0) v[r] = v[r] * row_tw<F>(tw, twr, r, j2);
        }
        dft<F::R2, +1>(v);
#pragma unroll
        for (int r = 0; r < F::R2; ++r) {
          const int x = j2 + r * F::R1 - cs * F::MC;
          if (CSPLIT == 1 || (x >= 0 && x < F::MC)) Tb[a * F::MC + x] = v[r];
        }
      }
    }
    __syncthreads();
  }
}

// Row-pass inputs of one system held in registers, so the next system's
// mask-band and transfer-table values can be loaded while the current one
// runs its column transforms.  Thread (rl1, j1) of the stage-1 mapping needs
// rows a = c * CPR + rl1 and columns b = j1 + r * T1 (a, b < W <= (M+1)/2).
// Used when that is at most 16 pairs per thread (4 in double precision).
template <typename T, int M, int NT, int CSPLIT>
struct RowPre {
  using F = FieldShape<T, M, NT, CSPLIT>;
  static constexpr int MAXW = (M + 1) / 2;
  static constexpr int NCH = (MAXW + F::CPR - 1) / F::CPR;
  static constexpr int NR = (MAXW + F::T1 - 1) / F::T1 < F::R1 ? (MAXW + F::T1 - 1) / F::T1 : F::R1;
  static constexpr bool ON = NCH * NR <= (sizeof(T) == 4 ? 16 : 4);
  // the adjoint keeps its coarse sensitivity in registers as well: prefetch
  // there only when it costs at most half as much
  static constexpr bool ON_ADJ = NCH * NR <= (sizeof(T) == 4 ? 8 : 2);
  // With the inputs in registers the forward row pass transforms all W rows
  // at once, so its scratch buffer Cb holds up to MAXW rows; the adjoint (and
  // the row pass without RowPre) runs in chunks of CPR rows.
  // DB: the forward column rounds alternate between two CPR-row halves of Cb,
  // so a round needs one barrier (between its two radix stages) instead of
  // two; used when the second half costs at most 16 rows beyond the row pass's
  // own scratch.
  static constexpr bool DB = ON && F::ROUNDS >= 2 && 2 * F::CPR <= MAXW + 16;
  static constexpr int CB_ROWS_FWD =
      DB ? (MAXW > 2 * F::CPR ? MAXW : 2 * F::CPR) : (ON && MAXW > F::CPR ? MAXW : F::CPR);
  // Cb in complex elements: the column rounds' rows (LDC apart) or the row
  // pass's rows (RowLayout), whichever needs more
  static constexpr int ROW_ROWS_FWD = ON ? MAXW : F::CPR;
  static constexpr int CB_FWD = CB_ROWS_FWD * F::LDC > ROW_ROWS_FWD * RowLayout<M>::LDR
                                    ? CB_ROWS_FWD * F::LDC
                                    : ROW_ROWS_FWD * RowLayout<M>::LDR;
  static constexpr int CB_ADJ = F::LDC > RowLayout<M>::LDR ? F::CPR * F::LDC
                                                           : F::CPR * RowLayout<M>::LDR;
  static_assert(!RowLayout<M>::SWZ || F::CPR % 4 == 0, "the row swizzle permutes rows in fours");
  cx<T> x[NCH][NR], h[NCH][NR];

  __device__ __forceinline__ void fetch(const cx<T>* __restrict__ Xm, int Wb, int wrapN,
                                        const cx<T>* __restrict__ Hj, int W, int2 bx, int rl1,
                                        int j1) {
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const int a = c * F::CPR + rl1;
      const bool ra = a < W;
      const int xr = ra ? wrap_idx(bx.x + a, wrapN) : 0;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int b = j1 + r * F::T1;
        if (ra && b < W) {
          x[c][r] = Xm[(long long)xr * Wb + wrap_idx(bx.y + b, wrapN)];
          h[c][r] = Hj[a * W + b];
        } else {
          x[c][r] = cx<T>{0, 0};
          h[c][r] = cx<T>{0, 0};
        }
      }
    }
  }
};

// row_pass with the stage-1 inputs taken from a RowPre.  ALL: stage 1 of all
// W rows (thread (rl1, j1) takes rows c * CPR + rl1), one barrier, stage 2 of
// all W rows (W * T2 tasks over the block), one barrier, with S holding
// CB_ROWS_FWD rows; else chunks of CPR rows, two barriers each.
template <bool ALL, typename T, int M, int NT, int CSPLIT>
__device__ __forceinline__ void row_pass_pre(const RowPre<T, M, NT, CSPLIT>& pre, int W,
                                             const cx<T>* tw, cx<T>* S, cx<T>* Tb, int cs,
                                             int tid, const cx<T>* twr) {
  using F = FieldShape<T, M, NT, CSPLIT>;
  using P = RowPre<T, M, NT, CSPLIT>;
  using RL = RowLayout<M>;
  const int rl1 = stage1_row<M>(tid / F::T1), j1 = tid - (tid / F::T1) * F::T1;
  if constexpr (!ALL) {
#pragma unroll
    for (int c = 0; c < P::NCH; ++c) {
      const int a0 = c * F::CPR;
      if (a0 >= W) break;  // uniform over the block
      if (a0 + rl1 < W) {
        cx<T> v[F::R1];
#pragma unroll
        for (int r = 0; r < F::R1; ++r) v[r] = r < P::NR ? pre.x[c][r] * pre.h[c][r] : cx<T>{0, 0};
        dft<F::R1, +1>(v);
#pragma unroll
        for (int r = 0; r < F::R1; ++r) S[rl1 * RL::LDR + j1 * RL::SP + r] = v[r];
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < F::Q2; ++q) {
        const int t = tid + q * NT;
        const int rl = t / F::T2, j2 = t - (t / F::T2) * F::T2;
        const int a = a0 + rl;
        if (t < F::NT2 && a < W) {
          cx<T> v[F::R2];
#pragma unroll
          for (int r = 0; r < F::R2; ++r) {
            v[r] = S[rl * RL::LDR + r * RL::SP + j2];
            if (r > 0) v[r] = v[r] * row_tw<F>(tw, twr, r, j2);
          }
          dft<F::R2, +1>(v);
#pragma unroll
          for (int r = 0; r < F::R2; ++r) {
            const int x = j2 + r * F::R1 - cs * F::MC;
            if (CSPLIT == 1 || (x >= 0 && x < F::MC)) Tb[a * F::MC + x] = v[r];
          }
        }
      }
      __syncthreads();
    }
    return;
  }
#pragma unroll
  for (int c = 0; c < P::NCH; ++c) {
    const int a = c * F::CPR + rl1;
    if (a < W) {
      cx<T> v[F::R1];
#pragma unroll
      for (int r = 0; r < F::R1; ++r) v[r] = r < P::NR ? pre.x[c][r] * pre.h[c][r] : cx<T>{0, 0};
      dft<F::R1, +1>(v);
#pragma unroll
      for (int r = 0; r < F::R1; ++r) S[a * RL::LDR + j1 * RL::SP + r] = v[r];
    }
  }
  __syncthreads();
  constexpr int Q2R = (P::MAXW * F::T2 + NT - 1) / NT;
#pragma unroll
  for (int q = 0; q < Q2R; ++q) {
    const int t = tid + q * NT;
    const int a = t / F::T2, j2 = t - (t / F::T2) * F::T2;
    if (a < W) {
      cx<T> v[F::R2];
#pragma unroll
      for (int r = 0; r < F::R2; ++r) {
        v[r] = S[a * RL::LDR + r * RL::SP + j2];
        if (r > 0) v[r] = v[r] * row_tw<F>(tw, twr, r, j2);
      }
      dft<F::R2, +1>(v);
#pragma unroll
      for (int r = 0; r < F::R2; ++r) {
        const int x = j2 + r * F::R1 - cs * F::MC;
        if (CSPLIT == 1 || (x >= 0 && x < F::MC)) Tb[a * F::MC + x] = v[r];
      }
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------- work items
// What one block needs for one work item: a mask's band spectrum, the
// systems [s0, s1) of one image, and where its partial goes.
template <typename T>
struct ItemView {
  const cx<T>* X;    // band spectrum of the item's mask, Wb x Wb
  int Wb, wrapN;
  const cx<T>* H;    // transfer tables (system j at H + j W W)
  int W;
  const int2* box;
  const T* wsys;
  int s0, s1;
  T* part;           // forward: M x M coarse partial
  const T* gM;       // adjoint: M x M coarse sensitivity of the item's image
  cx<T>* gX;         // adjoint: Wb x Wb band-gradient partial
};

// Items of one plan, by arithmetic (ItemMap).
template <typename T, int M>
struct MapSource {
  const cx<T>* X;
  int Wb, wrapN;
  const cx<T>* H;
  int W;
  const int2* box;
  const T* wsys;
  ItemMap map;
  T* partial;
  const T* gM;
  cx<T>* gX;
  int* queue;
  __device__ int count() const { return map.n_items; }
  __device__ ItemView<T> get(int item) const {
    int img, s0, s1;
    const int spp = map.sys_per_plane;
    if (map.spread) {
      // block b's run [lo, hi) of the image-ordered systems, its k-th image
      const int b = item % map.blocks, k = item / map.blocks;
      const long long lo = (long long)b * map.total / map.blocks;
      const long long hi = (long long)(b + 1) * map.total / map.blocks;
      const long long im = lo / spp + k;
      const long long a0 = max(lo, im * spp), a1 = min(hi, (im + 1) * spp);
      if (a0 < a1 && im * spp < map.total) {
        img = (int)im;
        s0 = (int)(a0 - im * spp);
        s1 = (int)(a1 - im * spp);
      } else {
        img = 0;  // empty segment: a valid image, no systems
        s0 = s1 = 0;
      }
    } else {
      img = item / map.n_chunks;
      const int chunk = item - img * map.n_chunks;
      s0 = chunk * map.chunk_sz;
      s1 = min(s0 + map.chunk_sz, spp);
    }
    const int plane = img % map.n_planes, mask = img / map.n_planes;
    ItemView<T> v;
    v.X = X + (long long)mask * Wb * Wb;
    v.Wb = Wb;
    v.wrapN = wrapN;
    v.H = H;
    v.W = W;
    v.box = box;
    v.wsys = wsys;
    v.s0 = plane * spp + s0;
    v.s1 = plane * spp + s1;
    v.part = partial ? partial + (long long)item * M * M : nullptr;
    v.gM = gM ? gM + (long long)img * M * M : nullptr;
    v.gX = gX ? gX + (long long)item * Wb * Wb : nullptr;
    return v;
  }
};

// Items of several plans (uneven clips) from a descriptor table; buffer
// offsets are relative to the workspace base passed at launch.
template <typename T>
struct TableSource {
  const ItemDesc* items;
  int n;
  char* ws;
  int* queue;
  __device__ int count() const { return n; }
  __device__ ItemView<T> get(int i) const {
    const ItemDesc d = items[i];
    ItemView<T> v;
    v.X = (const cx<T>*)(ws + d.X_off);
    v.Wb = d.Wb;
    v.wrapN = d.wrapN;
    v.H = (const cx<T>*)d.H;
    v.W = d.W;
    v.box = d.box;
    v.wsys = (const T*)d.wsys;
    v.s0 = d.s0;
    v.s1 = d.s1;
    v.part = d.part_off >= 0 ? (T*)(ws + d.part_off) : nullptr;
    v.gM = d.gM_off >= 0 ? (const T*)(ws + d.gM_off) : nullptr;
    v.gX = d.gX_off >= 0 ? (cx<T>*)(ws + d.gX_off) : nullptr;
    return v;
  }
};

// The block's next work item: the first is blockIdx.x, later ones come from
// an atomic counter (items differ in cost -- the last chunk of an image, clips
// of different boxes -- so blocks that finish early take more), or in a fixed
// stride when there is no counter.  Called by every thread of the block.
__device__ __forceinline__ int next_item(int* queue, int item, int* s_next) {
  if (!queue) return item + (int)gridDim.x;
  __syncthreads();
  if (threadIdx.x == 0) *s_next = (int)gridDim.x + atomicAdd(queue, 1);
  __syncthreads();
  return *s_next;
}

// Column rounds taken in pairs: where a round's stage-2 tasks (CPR columns x
// T2) do not fill whole passes of the block but two rounds' do (M = 96 at 256
// threads: 384 tasks, 1.5 passes; two rounds 3), both rounds' stage 1 write
// the two halves of the round buffer and their stage 2 runs as one group of
// 2 CPR columns, so no pass runs on half the waves while the others wait at
// the next barrier.  Group g covers rounds 2g and 2g + 1 (the last alone when
// the round count is odd).
template <class F, int NT>
struct RoundPairs {
  static constexpr int G = (F::ROUNDS + 1) / 2;
  static constexpr int cols(int g) { return (2 * g + 1 < F::ROUNDS ? 2 : 1) * F::CPR; }
  static constexpr int passes(int g) { return (cols(g) * F::T2 + NT - 1) / NT; }
  static constexpr int off(int g) {
    int o = 0;
    for (int i = 0; i < g; ++i) o += passes(i) * F::R2;
    return o;
  }
  static constexpr int NACC = off(G);
};

template <typename T, int M, int NT, int CSPLIT, typename Src>
__global__ __launch_bounds__(NT) void k_field_fwd(Src src, const cx<T>* __restrict__ twg) {
  using F = FieldShape<T, M, NT, CSPLIT>;
  using Pre = RowPre<T, M, NT, CSPLIT>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);  // M
  cx<T>* twr = tw + M;                             // TWR_N: row stage-2 twiddles by (r, j2)
  cx<T>* Cb = twr + F::TWR_N;                      // CB_ROWS_FWD x LDC (row scratch too)
  cx<T>* Tb = Cb + Pre::CB_FWD;                    // W x MC
  const int tid = threadIdx.x;
  load_roots(tw, twg, M, tid, NT);
  load_row_tw<F>(twr, twg, tid, NT);
  __syncthreads();
  const int c1 = tid % F::CPR, jc1 = tid / F::CPR;  // column stage-1 mapping

  const int n_items = src.count();
  __shared__ int s_next;
  for (int item = blockIdx.x; item < n_items; item = next_item(src.queue, item, &s_next)) {
    const ItemView<T> iv = src.get(item);
    const cx<T>* Xm = iv.X;
    const int Wb = iv.Wb, wrapN = iv.wrapN, W = iv.W, s0 = iv.s0, s1 = iv.s1;
    const cx<T>* H = iv.H;
    const int2* box = iv.box;
    const T* wsys = iv.wsys;
    Pre pre;
    const int rl1 = stage1_row<M>(tid / F::T1), j1 = tid - (tid / F::T1) * F::T1;
    if (Pre::ON && s0 < s1)
      pre.fetch(Xm, Wb, wrapN, H + (long long)s0 * W * W, W, box[s0], rl1, j1);
    using RP = RoundPairs<F, NT>;
    constexpr bool PAIR = Pre::DB && F::NT2 % NT != 0 && (2 * F::NT2) % NT == 0;
    constexpr int NACC = PAIR ? RP::NACC : F::NACC;
    for (int cs = 0; cs < CSPLIT; ++cs) {
      T acc[NACC];
#pragma unroll
      for (int i = 0; i < NACC; ++i) acc[i] = 0;
      for (int j = s0; j < s1; ++j) {
        const T w = wsys[j];
        if constexpr (Pre::ON) {
          row_pass_pre<true>(pre, W, tw, Cb, Tb, cs, tid, twr);
          // the next system's inputs load while this one runs its columns
          const int jn = j + 1 < s1 ? j + 1 : (cs + 1 < CSPLIT ? s0 : -1);
          if (jn >= 0) pre.fetch(Xm, Wb, wrapN, H + (long long)jn * W * W, W, box[jn], rl1, j1);
        } else {
          row_pass<T, M, NT, CSPLIT>(Xm, Wb, wrapN, H + (long long)j * W * W, W, box[j], tw, Cb,
                                     Tb, cs, tid, twr);
        }
        if constexpr (PAIR) {
#pragma unroll
          for (int g = 0; g < RP::G; ++g) {
            const int nr = 2 * g + 1 < F::ROUNDS ? 2 : 1;
            // stage 1 of the group's rounds into the halves of the round buffer
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              if (h < nr) {
                cx<T> v[F::R1];
                const int xl = (2 * g + h) * F::CPR + c1;
#pragma unroll
                for (int r = 0; r < F::R1; ++r) {
                  const int a = jc1 + r * F::T1;
                  v[r] = (r < F::NRW && a < W) ? Tb[a * F::MC + xl] : cx<T>{0, 0};
                }
                dft<F::R1, +1>(v);
#pragma unroll
                for (int r = 0; r < F::R1; ++r)
                  Cb[(h * F::CPR + c1) * F::LDC + jc1 * F::R1 + r] = v[r];
              }
            }
            __syncthreads();
            const int ncol = nr * F::CPR;
#pragma unroll
            for (int q = 0; q < RP::passes(g); ++q) {
              const int t = tid + q * NT;
              if (t < ncol * F::T2) {
                const int c2 = t % ncol, j2 = t / ncol;
                cx<T> v[F::R2];
#pragma unroll
                for (int r = 0; r < F::R2; ++r) {
                  v[r] = Cb[c2 * F::LDC + j2 + r * F::R1];
                  if (r > 0) v[r] = v[r] * tw[r * j2];
                }
                dft<F::R2, +1>(v);
#pragma unroll
                for (int r = 0; r < F::R2; ++r)
                  acc[RP::off(g) + q * F::R2 + r] += w * norm2(v[r]);
              }
            }
            __syncthreads();  // the next group's stage 1 (or row pass) rewrites Cb
          }
        } else {
        // Column rounds.  With Pre::DB round rd uses half (rd & 1) of Cb: a
        // round's stage 1 may overwrite the other half while slower threads
        // still read this one, and every thread has left round rd - 2 when it
        // passes round rd - 1's barrier.
#pragma unroll
        for (int rd = 0; rd < F::ROUNDS; ++rd) {
          cx<T>* Cr = Pre::DB ? Cb + (rd & 1) * F::CPR * F::LDC : Cb;
          {
            cx<T> v[F::R1];
            const int xl = rd * F::CPR + c1;
#pragma unroll
            for (int r = 0; r < F::R1; ++r) {
              const int a = jc1 + r * F::T1;
              v[r] = (r < F::NRW && a < W) ? Tb[a * F::MC + xl] : cx<T>{0, 0};
            }
            dft<F::R1, +1>(v);
#pragma unroll
            for (int r = 0; r < F::R1; ++r) Cr[c1 * F::LDC + jc1 * F::R1 + r] = v[r];
          }
          __syncthreads();
#pragma unroll
          for (int q = 0; q < F::Q2; ++q) {
            const int t = tid + q * NT;
            if (F::NT2 % NT == 0 || t < F::NT2) {
              const int c2 = t % F::CPR, j2 = t / F::CPR;
              cx<T> v[F::R2];
#pragma unroll
              for (int r = 0; r < F::R2; ++r) {
                v[r] = Cr[c2 * F::LDC + j2 + r * F::R1];
                if (r > 0) v[r] = v[r] * tw[r * j2];
              }
              dft<F::R2, +1>(v);
#pragma unroll
              for (int r = 0; r < F::R2; ++r)
                acc[(rd * F::Q2 + q) * F::R2 + r] += w * norm2(v[r]);
            }
          }
          if (!Pre::DB) __syncthreads();
        }
        if (Pre::DB) __syncthreads();  // the next row pass writes over both halves
        }
      }
      T* out = iv.part;
      if constexpr (PAIR) {
#pragma unroll
        for (int g = 0; g < RP::G; ++g) {
          const int ncol = (2 * g + 1 < F::ROUNDS ? 2 : 1) * F::CPR;
#pragma unroll
          for (int q = 0; q < RP::passes(g); ++q) {
            const int t = tid + q * NT;
            if (t < ncol * F::T2) {
              const int c2 = t % ncol, j2 = t / ncol;
              const int x = cs * F::MC + 2 * g * F::CPR + c2;
#pragma unroll
              for (int r = 0; r < F::R2; ++r)
                out[(j2 + r * F::R1) * M + x] = acc[RP::off(g) + q * F::R2 + r];
            }
          }
        }
      } else {
#pragma unroll
        for (int rd = 0; rd < F::ROUNDS; ++rd)
#pragma unroll
          for (int q = 0; q < F::Q2; ++q) {
            const int t = tid + q * NT;
            if (F::NT2 % NT == 0 || t < F::NT2) {
              const int c2 = t % F::CPR, j2 = t / F::CPR;
              const int x = cs * F::MC + rd * F::CPR + c2;
#pragma unroll
              for (int r = 0; r < F::R2; ++r)
                out[(j2 + r * F::R1) * M + x] = acc[(rd * F::Q2 + q) * F::R2 + r];
            }
          }
      }
    }
  }
}

template <typename T, int M, int NT, int CSPLIT, typename Src>
__global__ __launch_bounds__(NT) void k_field_adj(Src src, int Wmax,
                                                  const cx<T>* __restrict__ twg) {
  using F = FieldShape<T, M, NT, CSPLIT>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);  // M
  cx<T>* twr = tw + M;                             // TWR_N: row stage-2 twiddles by (r, j2)
  cx<T>* Cb = twr + F::TWR_N;                      // CPR rows (column or row layout)
  cx<T>* Tb = Cb + RowPre<T, M, NT, CSPLIT>::CB_ADJ;  // W x MC
  // The transposed column stage writes its box rows over the row table: round
  // rd's stage 1 has read Tb at exactly those (row, column) places two
  // barriers earlier, later rounds read other columns, and the transposed
  // row pass ends with a barrier before the next system's row pass.
  cx<T>* Tp = Tb;
  (void)Wmax;  // the table extent is the launch's LDS size
  const int tid = threadIdx.x;
  load_roots(tw, twg, M, tid, NT);
  load_row_tw<F>(twr, twg, tid, NT);
  __syncthreads();
  const int c1 = tid % F::CPR, jc1 = tid / F::CPR;
  const int rl1 = stage1_row<M>(tid / F::T1), j1r = tid - (tid / F::T1) * F::T1;

  const int n_items = src.count();
  __shared__ int s_next;
  for (int item = blockIdx.x; item < n_items; item = next_item(src.queue, item, &s_next)) {
    const ItemView<T> iv = src.get(item);
    const cx<T>* Xm = iv.X;
    const int Wb = iv.Wb, wrapN = iv.wrapN, W = iv.W, s0 = iv.s0, s1 = iv.s1;
    const cx<T>* H = iv.H;
    const int2* box = iv.box;
    const T* wsys = iv.wsys;
    const T* g = iv.gM;
    cx<T>* gout = iv.gX;
    for (int e = tid; e < Wb * Wb; e += NT) gout[e] = cx<T>{0, 0};
    __syncthreads();
    using Pre = RowPre<T, M, NT, CSPLIT>;
    Pre pre;
    if (Pre::ON_ADJ && s0 < s1)
      pre.fetch(Xm, Wb, wrapN, H + (long long)s0 * W * W, W, box[s0], rl1, j1r);
    for (int cs = 0; cs < CSPLIT; ++cs) {
      T gm[F::NACC];
#pragma unroll
      for (int rd = 0; rd < F::ROUNDS; ++rd)
#pragma unroll
        for (int q = 0; q < F::Q2; ++q) {
          const int t = tid + q * NT;
          const bool ok = F::NT2 % NT == 0 || t < F::NT2;
          const int c2 = t % F::CPR, j2 = t / F::CPR;
          const int x = cs * F::MC + rd * F::CPR + c2;
#pragma unroll
          for (int r = 0; r < F::R2; ++r)
            gm[(rd * F::Q2 + q) * F::R2 + r] = ok ? g[(j2 + r * F::R1) * M + x] : (T)0;
        }
      for (int j = s0; j < s1; ++j) {
        const int2 bx = box[j];
        const T w2 = (T)2 * wsys[j];
        const cx<T>* Hj = H + (long long)j * W * W;
        if constexpr (Pre::ON_ADJ) {
          row_pass_pre<false>(pre, W, tw, Cb, Tb, cs, tid, twr);
          const int jn = j + 1 < s1 ? j + 1 : (cs + 1 < CSPLIT ? s0 : -1);
          if (jn >= 0) pre.fetch(Xm, Wb, wrapN, H + (long long)jn * W * W, W, box[jn], rl1, j1r);
        } else {
          row_pass<T, M, NT, CSPLIT>(Xm, Wb, wrapN, Hj, W, bx, tw, Cb, Tb, cs, tid, twr);
        }
#pragma unroll
        for (int rd = 0; rd < F::ROUNDS; ++rd) {
          const int xl = rd * F::CPR + c1;
          {
            cx<T> v[F::R1];
#pragma unroll
            for (int r = 0; r < F::R1; ++r) {
              const int a = jc1 + r * F::T1;
              v[r] = (r < F::NRW && a < W) ? Tb[a * F::MC + xl] : cx<T>{0, 0};
            }
            dft<F::R1, +1>(v);
#pragma unroll
            for (int r = 0; r < F::R1; ++r) Cb[c1 * F::LDC + jc1 * F::R1 + r] = v[r];
          }
          __syncthreads();
#pragma unroll
          for (int q = 0; q < F::Q2; ++q) {
            const int t = tid + q * NT;
            if (!(F::NT2 % NT == 0 || t < F::NT2)) continue;
            const int c2 = t % F::CPR, j2 = t / F::CPR;
            cx<T> v[F::R2];
#pragma unroll
            for (int r = 0; r < F::R2; ++r) {
              v[r] = Cb[c2 * F::LDC + j2 + r * F::R1];
              if (r > 0) v[r] = v[r] * tw[r * j2];
            }
            dft<F::R2, +1>(v);
            // h = gM * E, then the transposed stage 2: forward DFT, conj twiddle
#pragma unroll
            for (int r = 0; r < F::R2; ++r) v[r] = scale(v[r], gm[(rd * F::Q2 + q) * F::R2 + r]);
            dft<F::R2, -1>(v);
#pragma unroll
            for (int r = 0; r < F::R2; ++r) {
              if (r > 0) v[r] = mulc(v[r], tw[r * j2]);
              Cb[c2 * F::LDC + j2 + r * F::R1] = v[r];
            }
          }
          __syncthreads();
          {
            // transposed stage 1: outputs pruned to the W box rows.  A thread
            // reads back exactly the slots its own stage 1 wrote, and the next
            // round's stage 1 writes those same slots, so only the last round
            // needs a barrier (the transposed row pass reads the whole table).
            cx<T> v[F::R1];
#pragma unroll
            for (int r = 0; r < F::R1; ++r) v[r] = Cb[c1 * F::LDC + jc1 * F::R1 + r];
            dft<F::R1, -1>(v);
#pragma unroll
            for (int r = 0; r < F::R1; ++r) {
              const int a = jc1 + r * F::T1;
              if (r < F::NRW && a < W) Tp[a * F::MC + xl] = v[r];
            }
          }
          if (rd + 1 == F::ROUNDS) __syncthreads();
        }
        // transposed row pass over Tp, in blocks of CPR rows: the transposed
        // stage 2 reads its rows straight from Tp (columns outside this split
        // are zero) and leaves its outputs in Cb in the row-pass layout
        for (int a0 = 0; a0 < W; a0 += F::CPR) {
          // This thread's band-gradient entries and transfer values for the
          // stage 1 below.  In FP32 they load before the stage-2 work so their
          // latency overlaps it (the previous system's stores to these entries
          // were made visible by the barriers since); FP64 has no registers to
          // spare and loads them after its transform.
          const int a1 = a0 + rl1;
          const int xr1 = wrap_idx(bx.x + a1, wrapN);
          constexpr bool kEarly = sizeof(T) == 4;
          cx<T> old[F::NRW], hv[F::NRW];
          if constexpr (kEarly) {
#pragma unroll
            for (int r = 0; r < F::NRW; ++r) {
              const int b = j1r + r * F::T1;
              const bool in = a1 < W && b < W;
              old[r] = in ? gout[(long long)xr1 * Wb + wrap_idx(bx.y + b, wrapN)] : cx<T>{0, 0};
              hv[r] = in ? Hj[a1 * W + b] : cx<T>{0, 0};
            }
          }
#pragma unroll
          for (int q = 0; q < F::Q2; ++q) {
            const int t = tid + q * NT;
            if (!(F::NT2 % NT == 0 || t < F::NT2)) continue;
            const int rl = t / F::T2, j2 = t - (t / F::T2) * F::T2;
            const int a = a0 + rl;
            cx<T> v[F::R2];
#pragma unroll
            for (int r = 0; r < F::R2; ++r) {
              const int xl = j2 + r * F::R1 - cs * F::MC;
              v[r] = (a < W && (CSPLIT == 1 || (xl >= 0 && xl < F::MC))) ? Tp[a * F::MC + xl]
                                                                          : cx<T>{0, 0};
            }
            dft<F::R2, -1>(v);
#pragma unroll
            for (int r = 0; r < F::R2; ++r) {
              if (r > 0) v[r] = mulc(v[r], row_tw<F>(tw, twr, r, j2));
              Cb[rl * RowLayout<M>::LDR + r * RowLayout<M>::SP + j2] = v[r];
            }
          }
          __syncthreads();
          {
            cx<T> v[F::R1];
#pragma unroll
            for (int r = 0; r < F::R1; ++r)
              v[r] = Cb[rl1 * RowLayout<M>::LDR + j1r * RowLayout<M>::SP + r];
            dft<F::R1, -1>(v);
            if (a1 < W) {
#pragma unroll
              for (int r = 0; r < F::NRW; ++r) {
                const int b = j1r + r * F::T1;
                if (b < W) {
                  cx<T>* dst = gout + (long long)xr1 * Wb + wrap_idx(bx.y + b, wrapN);
                  if constexpr (kEarly)
                    *dst = old[r] + scale(mulc(v[r], hv[r]), w2);
                  else
                    *dst = *dst + scale(mulc(v[r], Hj[a1 * W + b]), w2);
                }
              }
            }
          }
          __syncthreads();
        }
      }
    }
  }
}

// ------------------------------------------- staged path (row tables in memory)
// On large coarse grids a system's W x M row table does not fit in LDS beside
// the round buffer (W = 117, M = 256: 240 KB in FP32), and column splits
// recompute the row pass once per split.  The staged path instead writes the
// row table of every (image, system) of a pass to device memory once
// (k_rows_staged: one block task = CPR rows of one table), then
// k_cols_staged runs the column transforms of CPR columns of one work item
// over the item's systems, accumulating w |E|^2 in registers; the next
// system's column inputs load while the current one runs its second stage.
// A pass's tables are sized to stay in the last-level cache between the two
// launches.  Same arithmetic as the single-launch kernel, stage by stage.
template <typename T, int M, int NT>
__global__ __launch_bounds__(NT) void k_rows_staged(MapSource<T, M> src, int img0, int n_img,
                                                    cx<T>* __restrict__ G,
                                                    const cx<T>* __restrict__ twg) {
  using F = FieldShape<T, M, NT, 1>;
  using L = RowLayout<M>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);  // M
  cx<T>* S = tw + M;                               // CPR x LDR
  const int tid = threadIdx.x;
  load_roots(tw, twg, M, tid, NT);
  __syncthreads();
  const int W = src.W, spp = src.map.sys_per_plane, n_planes = src.map.n_planes;
  const int nrc = (W + F::CPR - 1) / F::CPR;
  const long long n_tasks = (long long)n_img * spp * nrc;
  const int rl1 = tid / F::T1, j1 = tid - rl1 * F::T1;
  // this thread's stage-1 inputs of a task (mask band and transfer values),
  // loaded one task ahead so their latency overlaps the current task
  cx<T> xv[F::NRW], hv[F::NRW];
  auto fetch = [&](long long task) {
    const int rc = (int)(task % nrc);
    const long long slot = task / nrc;
    const int jl = (int)(slot % spp);
    const int img = img0 + (int)(slot / spp);
    const int plane = img % n_planes, mask = img / n_planes;
    const int j = plane * spp + jl;
    const cx<T>* Xm = src.X + (long long)mask * src.Wb * src.Wb;
    const cx<T>* Hj = src.H + (long long)j * W * W;
    const int2 bx = src.box[j];
    const int a = rc * F::CPR + rl1;
    const int xr = wrap_idx(bx.x + (a < W ? a : 0), src.wrapN);
#pragma unroll
    for (int r = 0; r < F::NRW; ++r) {
      const int b = j1 + r * F::T1;
      const bool in = a < W && b < W;
      xv[r] = in ? Xm[(long long)xr * src.Wb + wrap_idx(bx.y + b, src.wrapN)] : cx<T>{0, 0};
      hv[r] = in ? Hj[a * W + b] : cx<T>{0, 0};
    }
  };
  if (blockIdx.x < n_tasks) fetch(blockIdx.x);
  for (long long task = blockIdx.x; task < n_tasks; task += gridDim.x) {
    const int rc = (int)(task % nrc);
    const long long slot = task / nrc;  // (image in pass) * spp + system in plane
    cx<T>* Gt = G + slot * W * M;
    const int a0 = rc * F::CPR;
    {
      const int a = a0 + rl1;
      cx<T> v[F::R1];
#pragma unroll
      for (int r = 0; r < F::R1; ++r) v[r] = r < F::NRW ? xv[r] * hv[r] : cx<T>{0, 0};
      if (task + gridDim.x < n_tasks) fetch(task + gridDim.x);
      if (a < W) {
        dft<F::R1, +1>(v);
#pragma unroll
        for (int r = 0; r < F::R1; ++r) S[rl1 * L::LDR + j1 * L::SP + r] = v[r];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < F::Q2; ++q) {
      const int t = tid + q * NT;
      const int rl = t / F::T2, j2 = t - (t / F::T2) * F::T2;
      const int a = a0 + rl;
      if (t < F::CPR * F::T2 && a < W) {
        // stage-2 input r is stage-1 output j2 of task r
        cx<T> v[F::R2];
#pragma unroll
        for (int r = 0; r < F::R2; ++r) {
          v[r] = S[rl * L::LDR + r * L::SP + j2];
          if (r > 0) v[r] = v[r] * tw[r * j2];
        }
        dft<F::R2, +1>(v);
#pragma unroll
        for (int r = 0; r < F::R2; ++r) Gt[(long long)a * M + j2 + r * F::R1] = v[r];
      }
    }
    __syncthreads();
  }
}

template <typename T, int M, int NT>
__global__ __launch_bounds__(NT) void k_cols_staged(MapSource<T, M> src, int img0, int n_img,
                                                    const cx<T>* __restrict__ G,
                                                    const cx<T>* __restrict__ twg) {
  using F = FieldShape<T, M, NT, 1>;
  constexpr int NG = M / F::CPR;  // column groups of a coarse grid
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);  // M
  cx<T>* Cb = tw + M;                              // CPR x LDC
  const int tid = threadIdx.x;
  load_roots(tw, twg, M, tid, NT);
  __syncthreads();
  const int W = src.W, spp = src.map.sys_per_plane, n_planes = src.map.n_planes;
  const int n_chunks = src.map.n_chunks;
  const int c1 = tid % F::CPR, jc1 = tid / F::CPR;
  const long long n_tasks = (long long)n_img * n_chunks * NG;
  for (long long task = blockIdx.x; task < n_tasks; task += gridDim.x) {
    const int g = (int)(task % NG);
    const int item = img0 * n_chunks + (int)(task / NG);
    const ItemView<T> iv = src.get(item);
    const int img = item / n_chunks, plane = img % n_planes;
    // row table of system j (of this image's plane): Gi + j * W * M
    const cx<T>* Gi = G + ((long long)(img - img0) * spp - (long long)plane * spp) * W * M;
    const int x = g * F::CPR + c1;
    cx<T> nx[F::NRW];
    auto fetch = [&](int j) {
      const cx<T>* Gt = Gi + (long long)j * W * M;
#pragma unroll
      for (int r = 0; r < F::NRW; ++r) {
        const int a = jc1 + r * F::T1;
        nx[r] = a < W ? Gt[(long long)a * M + x] : cx<T>{0, 0};
      }
    };
    T acc[F::Q2 * F::R2];
#pragma unroll
    for (int i = 0; i < F::Q2 * F::R2; ++i) acc[i] = 0;
    if (iv.s0 < iv.s1) fetch(iv.s0);
    for (int j = iv.s0; j < iv.s1; ++j) {
      const T w = iv.wsys[j];
      {
        cx<T> v[F::R1];
#pragma unroll
        for (int r = 0; r < F::R1; ++r) v[r] = r < F::NRW ? nx[r] : cx<T>{0, 0};
        dft<F::R1, +1>(v);
#pragma unroll
        for (int r = 0; r < F::R1; ++r) Cb[c1 * F::LDC + jc1 * F::R1 + r] = v[r];
      }
      if (j + 1 < iv.s1) fetch(j + 1);
      __syncthreads();
#pragma unroll
      for (int q = 0; q < F::Q2; ++q) {
        const int t = tid + q * NT;
        if (F::NT2 % NT == 0 || t < F::NT2) {
          const int c2 = t % F::CPR, j2 = t / F::CPR;
          cx<T> v[F::R2];
#pragma unroll
          for (int r = 0; r < F::R2; ++r) {
            v[r] = Cb[c2 * F::LDC + j2 + r * F::R1];
            if (r > 0) v[r] = v[r] * tw[r * j2];
          }
          dft<F::R2, +1>(v);
#pragma unroll
          for (int r = 0; r < F::R2; ++r) acc[q * F::R2 + r] += w * norm2(v[r]);
        }
      }
      __syncthreads();
    }
    T* out = iv.part;
#pragma unroll
    for (int q = 0; q < F::Q2; ++q) {
      const int t = tid + q * NT;
      if (F::NT2 % NT == 0 || t < F::NT2) {
        const int c2 = t % F::CPR, j2 = t / F::CPR;
#pragma unroll
        for (int r = 0; r < F::R2; ++r)
          out[(j2 + r * F::R1) * M + g * F::CPR + c2] = acc[q * F::R2 + r];
      }
    }
  }
}

// Staged adjoint.  k_rows_staged recomputes the row tables; k_cols_adj_staged
// runs, per (item, column group), each system's column stages, multiplies
// the field by the coarse sensitivity gM and runs the transposed column
// stages, writing the result over the system's row table (each block owns
// its columns of every table it touches); k_rows_adj_staged runs the
// transposed row pass of one chunk of CPR rows of an item over the item's
// systems and accumulates 2 w G_C conj(H) into its own band-gradient partial
// (item * nrc + chunk), so no two blocks share an accumulator even when the
// systems' boxes differ.
template <typename T, int M, int NT>
__global__ __launch_bounds__(NT) void k_cols_adj_staged(MapSource<T, M> src, int img0, int n_img,
                                                        cx<T>* __restrict__ G,
                                                        const cx<T>* __restrict__ twg) {
  using F = FieldShape<T, M, NT, 1>;
  constexpr int NG = M / F::CPR;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);  // M
  cx<T>* Cb = tw + M;                              // CPR x LDC
  const int tid = threadIdx.x;
  load_roots(tw, twg, M, tid, NT);
  __syncthreads();
  const int W = src.W, spp = src.map.sys_per_plane, n_planes = src.map.n_planes;
  const int n_chunks = src.map.n_chunks;
  const int c1 = tid % F::CPR, jc1 = tid / F::CPR;
  const long long n_tasks = (long long)n_img * n_chunks * NG;
  for (long long task = blockIdx.x; task < n_tasks; task += gridDim.x) {
    const int g = (int)(task % NG);
    const int item = img0 * n_chunks + (int)(task / NG);
    const ItemView<T> iv = src.get(item);
    const int img = item / n_chunks, plane = img % n_planes;
    cx<T>* Gi = G + ((long long)(img - img0) * spp - (long long)plane * spp) * W * M;
    const int x1 = g * F::CPR + c1;
    T gm[F::Q2 * F::R2];
#pragma unroll
    for (int q = 0; q < F::Q2; ++q) {
      const int t = tid + q * NT;
      const bool ok = F::NT2 % NT == 0 || t < F::NT2;
      const int c2 = t % F::CPR, j2 = t / F::CPR;
#pragma unroll
      for (int r = 0; r < F::R2; ++r)
        gm[q * F::R2 + r] = ok ? iv.gM[(j2 + r * F::R1) * M + g * F::CPR + c2] : (T)0;
    }
    cx<T> nx[F::NRW];
    auto fetch = [&](int j) {
      const cx<T>* Gt = Gi + (long long)j * W * M;
#pragma unroll
      for (int r = 0; r < F::NRW; ++r) {
        const int a = jc1 + r * F::T1;
        nx[r] = a < W ? Gt[(long long)a * M + x1] : cx<T>{0, 0};
      }
    };
    if (iv.s0 < iv.s1) fetch(iv.s0);
    for (int j = iv.s0; j < iv.s1; ++j) {
      cx<T>* Gt = Gi + (long long)j * W * M;
      {
        cx<T> v[F::R1];
#pragma unroll
        for (int r = 0; r < F::R1; ++r) v[r] = r < F::NRW ? nx[r] : cx<T>{0, 0};
        dft<F::R1, +1>(v);
#pragma unroll
        for (int r = 0; r < F::R1; ++r) Cb[c1 * F::LDC + jc1 * F::R1 + r] = v[r];
      }
      if (j + 1 < iv.s1) fetch(j + 1);
      __syncthreads();
#pragma unroll
      for (int q = 0; q < F::Q2; ++q) {
        const int t = tid + q * NT;
        if (!(F::NT2 % NT == 0 || t < F::NT2)) continue;
        const int c2 = t % F::CPR, j2 = t / F::CPR;
        cx<T> v[F::R2];
#pragma unroll
        for (int r = 0; r < F::R2; ++r) {
          v[r] = Cb[c2 * F::LDC + j2 + r * F::R1];
          if (r > 0) v[r] = v[r] * tw[r * j2];
        }
        dft<F::R2, +1>(v);
        // h = gM * E, then the transposed stage 2: forward DFT, conj twiddle
#pragma unroll
        for (int r = 0; r < F::R2; ++r) v[r] = scale(v[r], gm[q * F::R2 + r]);
        dft<F::R2, -1>(v);
#pragma unroll
        for (int r = 0; r < F::R2; ++r) {
          if (r > 0) v[r] = mulc(v[r], tw[r * j2]);
          Cb[c2 * F::LDC + j2 + r * F::R1] = v[r];
        }
      }
      __syncthreads();
      {
        // transposed stage 1, outputs pruned to the W box rows; a thread reads
        // back exactly the places it writes in the next system's stage 1
        cx<T> v[F::R1];
#pragma unroll
        for (int r = 0; r < F::R1; ++r) v[r] = Cb[c1 * F::LDC + jc1 * F::R1 + r];
        dft<F::R1, -1>(v);
#pragma unroll
        for (int r = 0; r < F::NRW; ++r) {
          const int a = jc1 + r * F::T1;
          if (a < W) Gt[(long long)a * M + x1] = v[r];
        }
      }
    }
    __syncthreads();  // the next task's first stage writes Cb
  }
}

template <typename T, int M, int NT>
__global__ __launch_bounds__(NT) void k_rows_adj_staged(MapSource<T, M> src, int img0, int n_img,
                                                        const cx<T>* __restrict__ G,
                                                        cx<T>* __restrict__ gXp,
                                                        const cx<T>* __restrict__ twg) {
  using F = FieldShape<T, M, NT, 1>;
  using RL = RowLayout<M>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);  // M
  cx<T>* S = tw + M;                               // CPR x LDR
  const int tid = threadIdx.x;
  load_roots(tw, twg, M, tid, NT);
  __syncthreads();
  const int W = src.W, Wb = src.Wb, wrapN = src.wrapN, spp = src.map.sys_per_plane;
  const int n_planes = src.map.n_planes, n_chunks = src.map.n_chunks;
  const int nrc = (W + F::CPR - 1) / F::CPR;
  const int rl1 = tid / F::T1, j1r = tid - rl1 * F::T1;
  const long long n_tasks = (long long)n_img * n_chunks * nrc;
  for (long long task = blockIdx.x; task < n_tasks; task += gridDim.x) {
    const int rc = (int)(task % nrc);
    const int item = img0 * n_chunks + (int)(task / nrc);
    const ItemView<T> iv = src.get(item);
    const int img = item / n_chunks, plane = img % n_planes;
    const cx<T>* Gi = G + ((long long)(img - img0) * spp - (long long)plane * spp) * W * M;
    cx<T>* gout = gXp + ((long long)item * nrc + rc) * Wb * Wb;
    for (int e = tid; e < Wb * Wb; e += NT) gout[e] = cx<T>{0, 0};
    __syncthreads();
    const int a0 = rc * F::CPR;
    // this thread's share of a system's CPR table rows, loaded one system ahead
    constexpr int NPT = (F::CPR * M + NT - 1) / NT;
    cx<T> nt[NPT];
    auto fetch = [&](int j) {
      const cx<T>* Gt = Gi + (long long)j * W * M;
#pragma unroll
      for (int i = 0; i < NPT; ++i) {
        const int e = tid + i * NT, rl = e / M, x = e - rl * M;
        nt[i] = e < F::CPR * M && a0 + rl < W ? Gt[(long long)(a0 + rl) * M + x] : cx<T>{0, 0};
      }
    };
    if (iv.s0 < iv.s1) fetch(iv.s0);
    for (int j = iv.s0; j < iv.s1; ++j) {
      const int2 bx = src.box[j];
      const T w2 = (T)2 * src.wsys[j];
      const cx<T>* Hj = src.H + (long long)j * W * W;
#pragma unroll
      for (int i = 0; i < NPT; ++i) {
        const int e = tid + i * NT, rl = e / M, x = e - rl * M;
        if (e < F::CPR * M) S[rl * RL::LDR + rpos<M, F::R1>(x)] = nt[i];
      }
      if (j + 1 < iv.s1) fetch(j + 1);
      // this system's transfer values and accumulator entries for the stage 1
      // below, loaded before the stage-2 work so their latency overlaps it (the
      // previous system's stores to the entries are visible after its barrier)
      const int a1 = a0 + rl1;
      const int xr1 = wrap_idx(bx.x + (a1 < W ? a1 : 0), wrapN);
      cx<T> hv[F::NRW], old[F::NRW];
#pragma unroll
      for (int r = 0; r < F::NRW; ++r) {
        const int b = j1r + r * F::T1;
        const bool in = a1 < W && b < W;
        hv[r] = in ? Hj[a1 * W + b] : cx<T>{0, 0};
        old[r] = in ? gout[(long long)xr1 * Wb + wrap_idx(bx.y + b, wrapN)] : cx<T>{0, 0};
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < F::Q2; ++q) {
        const int t = tid + q * NT;
        if (!(t < F::CPR * F::T2)) continue;
        const int rl = t / F::T2, j2 = t - (t / F::T2) * F::T2;
        cx<T> v[F::R2];
#pragma unroll
        for (int r = 0; r < F::R2; ++r) v[r] = S[rl * RL::LDR + r * RL::SP + j2];
        dft<F::R2, -1>(v);
#pragma unroll
        for (int r = 0; r < F::R2; ++r) {
          if (r > 0) v[r] = mulc(v[r], tw[r * j2]);
          S[rl * RL::LDR + r * RL::SP + j2] = v[r];
        }
      }
      __syncthreads();
      {
        cx<T> v[F::R1];
#pragma unroll
        for (int r = 0; r < F::R1; ++r) v[r] = S[rl1 * RL::LDR + j1r * RL::SP + r];
        dft<F::R1, -1>(v);
        if (a1 < W) {
#pragma unroll
          for (int r = 0; r < F::NRW; ++r) {
            const int b = j1r + r * F::T1;
            if (b < W)
              gout[(long long)xr1 * Wb + wrap_idx(bx.y + b, wrapN)] =
                  old[r] + scale(mulc(v[r], hv[r]), w2);
          }
        }
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------- dispatch
#define LB_FIELD_CONFIGS(X)                                                                \
  X(16, 64, 1) X(32, 128, 1) X(48, 64, 1) X(64, 256, 1) X(64, 128, 1) X(64, 256, 2)          \
      X(80, 320, 1) X(80, 320, 2) X(80, 160, 1)                                              \
      X(96, 256, 1) X(96, 128, 1) X(96, 256, 3) X(96, 384, 1) X(128, 512, 1) X(128, 256, 1) X(128, 256, 2) \
      X(128, 256, 4) X(256, 256, 2) X(256, 256, 4) X(256, 256, 8)

// staged shapes (column splits 0): coarse grid, threads per block
#define LB_STAGED_CONFIGS(X) X(128, 256) X(256, 256) X(256, 512)

// Default preference per coarse grid: the list order, except that on M = 256
// (where the row table outgrows LDS and splits recompute the row pass) the
// staged shapes come first.
const std::vector<FieldConfig>& field_configs() {
  static const std::vector<FieldConfig> all = [] {
    std::vector<FieldConfig> v, staged = {
#define LB_CASE(M_, NT_) FieldConfig{M_, NT_, 0, 0},
      LB_STAGED_CONFIGS(LB_CASE)
#undef LB_CASE
    };
    for (const FieldConfig& c : staged)
      if (c.M == 256) v.push_back(c);
    const FieldConfig single[] = {
#define LB_CASE(M_, NT_, CS_) FieldConfig{M_, NT_, CS_, 0},
      LB_FIELD_CONFIGS(LB_CASE)
#undef LB_CASE
    };
    for (const FieldConfig& c : single) v.push_back(c);
    for (const FieldConfig& c : staged)
      if (c.M != 256) v.push_back(c);
    return v;
  }();
  return all;
}

bool field_config_supported(int precision, const FieldConfig& c) {
  (void)precision;
#define LB_CASE(M_, NT_, CS_) \
  if (c.M == M_ && c.nt == NT_ && c.csplit == CS_) return true;
  LB_FIELD_CONFIGS(LB_CASE)
#undef LB_CASE
#define LB_CASE(M_, NT_) \
  if (c.M == M_ && c.nt == NT_ && c.csplit == 0) return true;
  LB_STAGED_CONFIGS(LB_CASE)
#undef LB_CASE
  return false;
}

template <typename T, int M, int NT>
static size_t staged_smem() {  // the larger of the two kernels' scratch
  using F = FieldShape<T, M, NT, 1>;
  const size_t rows = (size_t)F::CPR * RowLayout<M>::LDR, cols = (size_t)F::CPR * F::LDC;
  return (M + (rows > cols ? rows : cols)) * sizeof(cx<T>);
}

template <typename T, int M, int NT, int CS>
static size_t smem_for(int W, bool adjoint) {
  using F = FieldShape<T, M, NT, CS>;
  using P = RowPre<T, M, NT, CS>;
  const size_t cb = (adjoint ? P::CB_ADJ : P::CB_FWD) + F::TWR_N;
  size_t n = M + cb + (size_t)W * F::MC;  // the adjoint's Tp aliases Tb
  return n * sizeof(cx<T>);
}

size_t field_smem_bytes(int precision, const FieldConfig& c, int W, bool adjoint) {
  if (c.csplit == 0) {
    (void)adjoint;  // the staged kernels share one LDS size
#define LB_CASE(M_, NT_)                                                          \
  if (c.M == M_ && c.nt == NT_)                                                   \
    return precision == 64 ? staged_smem<double, M_, NT_>() : staged_smem<float, M_, NT_>();
    LB_STAGED_CONFIGS(LB_CASE)
#undef LB_CASE
    return 0;
  }
#define LB_CASE(M_, NT_, CS_)                                                   \
  if (c.M == M_ && c.nt == NT_ && c.csplit == CS_)                              \
    return precision == 64 ? smem_for<double, M_, NT_, CS_>(W, adjoint)         \
                           : smem_for<float, M_, NT_, CS_>(W, adjoint);
  LB_FIELD_CONFIGS(LB_CASE)
#undef LB_CASE
  return 0;
}

template <typename T, int M>
static MapSource<T, M> map_source(const FieldArgs& a, void* partial, const void* gM, void* gX) {
  MapSource<T, M> m;
  m.X = (const cx<T>*)a.X;
  m.Wb = a.Wb;
  m.wrapN = a.wrapN;
  m.H = (const cx<T>*)a.H;
  m.W = a.W;
  m.box = a.box;
  m.wsys = (const T*)a.wsys;
  m.map = a.map;
  m.partial = (T*)partial;
  m.gM = (const T*)gM;
  m.gX = (cx<T>*)gX;
  m.queue = a.queue;
  return m;
}

template <typename T, int M, int NT, int CS>
static hipError_t launch_fwd(const FieldConfig& c, const FieldArgs& a, void* partial,
                             hipStream_t st) {
  auto k = k_field_fwd<T, M, NT, CS, MapSource<T, M>>;
  const size_t sm = smem_for<T, M, NT, CS>(a.W, false);
  const cx<T>* twg = nullptr;
  hipError_t e = table_for<T>(M, &twg);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(c.grid), dim3(NT), sm, st,
                     map_source<T, M>(a, partial, nullptr, nullptr), twg);
  return hipGetLastError();
}

template <typename T, int M, int NT, int CS>
static hipError_t launch_adj(const FieldConfig& c, const FieldArgs& a, const void* gM, void* gX,
                             hipStream_t st) {
  auto k = k_field_adj<T, M, NT, CS, MapSource<T, M>>;
  const size_t sm = smem_for<T, M, NT, CS>(a.W, true);
  const cx<T>* twg = nullptr;
  hipError_t e = table_for<T>(M, &twg);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(c.grid), dim3(NT), sm, st,
                     map_source<T, M>(a, nullptr, gM, gX), a.W, twg);
  return hipGetLastError();
}

template <typename T, int M, int NT, int CS>
static hipError_t launch_multi(const FieldConfig& c, const ItemDesc* items, int n_items, int Wmax,
                               char* ws, int* queue, bool adjoint, hipStream_t st) {
  const size_t sm = smem_for<T, M, NT, CS>(Wmax, adjoint);
  const cx<T>* twg = nullptr;
  hipError_t e = table_for<T>(M, &twg);
  if (e != hipSuccess) return e;
  TableSource<T> src{items, n_items, ws, queue};
  if (adjoint) {
    auto k = k_field_adj<T, M, NT, CS, TableSource<T>>;
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(c.grid), dim3(NT), sm, st, src, Wmax, twg);
  } else {
    auto k = k_field_fwd<T, M, NT, CS, TableSource<T>>;
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(c.grid), dim3(NT), sm, st, src, twg);
  }
  return hipGetLastError();
}

template <typename T, int M, int NT>
static hipError_t launch_staged(const FieldArgs& a, void* partial, void* G, int img0, int n_img,
                                int cus, hipStream_t st) {
  using F = FieldShape<T, M, NT, 1>;
  const size_t sm = staged_smem<T, M, NT>();
  const cx<T>* twg = nullptr;
  hipError_t e = table_for<T>(M, &twg);
  if (e != hipSuccess) return e;
  const int occ = (int)std::min<size_t>(2048 / NT, (160 * 1024 - 64) / sm);
  const long long max_blocks = (long long)cus * std::max(1, occ);
  auto kr = k_rows_staged<T, M, NT>;
  auto kc = k_cols_staged<T, M, NT>;
  for (const void* k : {reinterpret_cast<const void*>(kr), reinterpret_cast<const void*>(kc)}) {
    e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
  }
  const MapSource<T, M> src = map_source<T, M>(a, partial, nullptr, nullptr);
  const long long row_tasks =
      (long long)n_img * a.map.sys_per_plane * ((a.W + F::CPR - 1) / F::CPR);
  hipLaunchKernelGGL(kr, dim3((unsigned)std::max(1LL, std::min(row_tasks, max_blocks))), dim3(NT),
                     sm, st, src, img0, n_img, (cx<T>*)G, twg);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long long col_tasks = (long long)n_img * a.map.n_chunks * (M / F::CPR);
  hipLaunchKernelGGL(kc, dim3((unsigned)std::max(1LL, std::min(col_tasks, max_blocks))), dim3(NT),
                     sm, st, src, img0, n_img, (const cx<T>*)G, twg);
  return hipGetLastError();
}

template <typename T, int M, int NT>
static hipError_t launch_staged_adj(const FieldArgs& a, const void* gM, void* gXp, void* G,
                                    int img0, int n_img, int cus, hipStream_t st) {
  using F = FieldShape<T, M, NT, 1>;
  const size_t sm = staged_smem<T, M, NT>();
  const cx<T>* twg = nullptr;
  hipError_t e = table_for<T>(M, &twg);
  if (e != hipSuccess) return e;
  const int occ = (int)std::min<size_t>(2048 / NT, (160 * 1024 - 64) / sm);
  const long long max_blocks = (long long)cus * std::max(1, occ);
  auto kr = k_rows_staged<T, M, NT>;
  auto kc = k_cols_adj_staged<T, M, NT>;
  auto ka = k_rows_adj_staged<T, M, NT>;
  for (const void* k : {reinterpret_cast<const void*>(kr), reinterpret_cast<const void*>(kc),
                        reinterpret_cast<const void*>(ka)}) {
    e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
  }
  const MapSource<T, M> src = map_source<T, M>(a, nullptr, gM, nullptr);
  const int nrc = (a.W + F::CPR - 1) / F::CPR;
  const long long row_tasks = (long long)n_img * a.map.sys_per_plane * nrc;
  auto grid = [&](long long tasks) {
    return dim3((unsigned)std::max(1LL, std::min(tasks, max_blocks)));
  };
  hipLaunchKernelGGL(kr, grid(row_tasks), dim3(NT), sm, st, src, img0, n_img, (cx<T>*)G, twg);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long long col_tasks = (long long)n_img * a.map.n_chunks * (M / F::CPR);
  hipLaunchKernelGGL(kc, grid(col_tasks), dim3(NT), sm, st, src, img0, n_img, (cx<T>*)G, twg);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long long adj_tasks = (long long)n_img * a.map.n_chunks * nrc;
  hipLaunchKernelGGL(ka, grid(adj_tasks), dim3(NT), sm, st, src, img0, n_img, (const cx<T>*)G,
                     (cx<T>*)gXp, twg);
  return hipGetLastError();
}

int staged_row_chunks(const FieldConfig& c, int W) {
#define LB_CASE(M_, NT_)                                        \
  if (c.M == M_ && c.nt == NT_) {                               \
    constexpr int cpr = FieldShape<float, M_, NT_, 1>::CPR;     \
    return (W + cpr - 1) / cpr;                                 \
  }
  LB_STAGED_CONFIGS(LB_CASE)
#undef LB_CASE
  return 1;
}

hipError_t field_adjoint_staged(int precision, const FieldConfig& c, const FieldArgs& a,
                                const void* gM, void* gXp, void* G, int img0, int n_img, int cus,
                                hipStream_t st) {
#define LB_CASE(M_, NT_)                                                                       \
  if (c.M == M_ && c.nt == NT_ && c.csplit == 0)                                               \
    return precision == 64                                                                     \
               ? launch_staged_adj<double, M_, NT_>(a, gM, gXp, G, img0, n_img, cus, st)      \
               : launch_staged_adj<float, M_, NT_>(a, gM, gXp, G, img0, n_img, cus, st);
  LB_STAGED_CONFIGS(LB_CASE)
#undef LB_CASE
  return hipErrorInvalidValue;
}

size_t staged_table_bytes(int precision, int M, int W) {
  return (size_t)W * M * (precision == 64 ? 16 : 8);
}

hipError_t field_forward_staged(int precision, const FieldConfig& c, const FieldArgs& a,
                                void* partial, void* G, int img0, int n_img, int cus,
                                hipStream_t st) {
#define LB_CASE(M_, NT_)                                                                     \
  if (c.M == M_ && c.nt == NT_ && c.csplit == 0)                                             \
    return precision == 64                                                                   \
               ? launch_staged<double, M_, NT_>(a, partial, G, img0, n_img, cus, st)        \
               : launch_staged<float, M_, NT_>(a, partial, G, img0, n_img, cus, st);
  LB_STAGED_CONFIGS(LB_CASE)
#undef LB_CASE
  return hipErrorInvalidValue;
}

hipError_t field_multi(int precision, const FieldConfig& c, const ItemDesc* items, int n_items,
                       int Wmax, char* ws, int* queue, bool adjoint, hipStream_t st) {
#define LB_CASE(M_, NT_, CS_)                                                                   \
  if (c.M == M_ && c.nt == NT_ && c.csplit == CS_)                                              \
    return precision == 64                                                                      \
               ? launch_multi<double, M_, NT_, CS_>(c, items, n_items, Wmax, ws, queue, adjoint, \
                                                    st)                                         \
               : launch_multi<float, M_, NT_, CS_>(c, items, n_items, Wmax, ws, queue, adjoint, \
                                                   st);
  LB_FIELD_CONFIGS(LB_CASE)
#undef LB_CASE
  return hipErrorInvalidValue;
}

hipError_t field_forward(int precision, const FieldConfig& c, const FieldArgs& a, void* partial,
                         hipStream_t st) {
#define LB_CASE(M_, NT_, CS_)                                                           \
  if (c.M == M_ && c.nt == NT_ && c.csplit == CS_)                                      \
    return precision == 64 ? launch_fwd<double, M_, NT_, CS_>(c, a, partial, st)        \
                           : launch_fwd<float, M_, NT_, CS_>(c, a, partial, st);
  LB_FIELD_CONFIGS(LB_CASE)
#undef LB_CASE
  return hipErrorInvalidValue;
}

hipError_t field_adjoint(int precision, const FieldConfig& c, const FieldArgs& a,
                         const void* gM, void* gX, hipStream_t st) {
#define LB_CASE(M_, NT_, CS_)                                                           \
  if (c.M == M_ && c.nt == NT_ && c.csplit == CS_)                                      \
    return precision == 64 ? launch_adj<double, M_, NT_, CS_>(c, a, gM, gX, st)         \
                           : launch_adj<float, M_, NT_, CS_>(c, a, gM, gX, st);
  LB_FIELD_CONFIGS(LB_CASE)
#undef LB_CASE
  return hipErrorInvalidValue;
}

}  // namespace lb
