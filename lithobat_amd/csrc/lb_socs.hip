// Hopkins / SOCS decomposition in FP64 on the device.
//
// The reference forms TCC = A^T conj(A) over the band (n_band^2, 1.7 GB at a
// 512^2 / NA 1.35 grid) and calls eigh on it (expose/hopkins.py:155-186).
// TCC = B B^H with B = A^T (n_band x S), so its non-zero eigenpairs are those
// of the S x S Gram matrix G = B^H B (G[s][t] = sum_f conj(A[s][f]) A[t][f]):
// G u = lambda u  <=>  TCC (B u) = lambda (B u), and phi = B u / sqrt(lambda)
// is unit-norm.  The Gram matrix is built here, diagonalised by a parallel
// cyclic two-sided Jacobi (Brent-Luk round-robin pairing, complex Hermitian
// rotations), and the kept eigenvectors are mapped back to kernels.
#include <algorithm>

#include "lb_kernels.h"

namespace lb {

using c64d = cx<double>;

constexpr int kGT = 16;

__global__ __launch_bounds__(kGT* kGT) void k_gram(const c64d* __restrict__ A, int S, long long F,
                                                   c64d* __restrict__ G) {
  __shared__ c64d As[kGT][kGT + 1];
  __shared__ c64d Bs[kGT][kGT + 1];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int s = blockIdx.y * kGT + ty;  // row of G
  const int t = blockIdx.x * kGT + tx;  // col of G
  double re = 0, im = 0;
  for (long long f0 = 0; f0 < F; f0 += kGT) {
    const int sr = blockIdx.y * kGT + ty, tr = blockIdx.x * kGT + ty;
    const long long f = f0 + tx;
    As[ty][tx] = (sr < S && f < F) ? A[(long long)sr * F + f] : c64d{0, 0};
    Bs[ty][tx] = (tr < S && f < F) ? A[(long long)tr * F + f] : c64d{0, 0};
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kGT; ++k) {
      const c64d a = As[ty][k], b = Bs[tx][k];
      // conj(a) * b
      re += a.re * b.re + a.im * b.im;
      im += a.re * b.im - a.im * b.re;
    }
    __syncthreads();
  }
  if (s < S && t < S) G[(long long)s * S + t] = c64d{re, im};
}

hipError_t gram_f64(const c64d* A, int S, long long F, c64d* G, hipStream_t st) {
  dim3 g((S + kGT - 1) / kGT, (S + kGT - 1) / kGT);
  hipLaunchKernelGGL(k_gram, g, dim3(kGT, kGT), 0, st, A, S, F, G);
  return hipGetLastError();
}

constexpr int kJT = 512;
constexpr int kMaxPairs = 1024;  // players (source points) per decomposition: S <= 1023
constexpr int kJU = 5;  // 2 x 2 blocks per thread and pass (all loads issued first)

// One rotation of the log: U = [[c, -s e], [s conj(e), c]] on columns (p, q).
struct JacobiRot {
  double c, s;
  c64d e;
};

// Round-robin pairing (Brent-Luk): pair i of round r over n players (n even).
__device__ __forceinline__ void jacobi_pair(int r, int i, int n, int* p, int* q) {
  // r, i < n - 1: each sum lies in [0, 2 (n - 1)), so one conditional subtraction
  // is the remainder
  const int m = n - 1;
  int a = r + i;
  if (a >= m) a -= m;
  int b = r - i + m;
  if (b >= m) b -= m;
  if (i == 0) b = m;
  *p = a < b ? a : b;
  *q = a < b ? b : a;
}

// Task w of the upper triangle of the half x half block grid, row-major: (a, b), a <= b.
__device__ __forceinline__ void jacobi_task(long long w, int half, int* pa, int* pb) {
  int a = (int)((2.0 * half + 1 - sqrt((2.0 * half + 1) * (2.0 * half + 1) - 8.0 * w)) / 2);
  long long base = (long long)a * half - (long long)a * (a - 1) / 2;
  while (base > w) {
    --a;
    base = (long long)a * half - (long long)a * (a - 1) / 2;
  }
  while (base + (half - a) <= w) {
    base += half - a;
    ++a;
  }
  *pa = a;
  *pb = a + (int)(w - base);
}

// Whether a pair rotates: its coupling must stand above the rounding noise of
// the pair (1e-15 of sqrt(G_pp G_qq)) and of the matrix (1e-17 of its largest
// diagonal entry); rotations below those change nothing a double can hold.
__device__ __forceinline__ bool jacobi_rotates(double ab, double a, double d, double floor_ab) {
  return ab > 1e-300 && ab > floor_ab && ab > 1e-15 * sqrt(fabs(a * d));
}

// The sweep-start test over per-thread partial sums (block of kJT threads):
// done when the off-diagonal Frobenius mass is below 1e-30 of the total or
// the previous sweep applied no rotation.  Returns the noise floor of the
// coming sweep.
template <int NT>
__device__ double jacobi_test(double off, double tot, double dmax, int applied,
                              double (*red)[NT / 64], int* done) {
  const int tid = threadIdx.x;
  for (int o = 32; o > 0; o >>= 1) {
    off += __shfl_xor(off, o, 64);
    tot += __shfl_xor(tot, o, 64);
    dmax = fmax(dmax, __shfl_xor(dmax, o, 64));
  }
  if ((tid & 63) == 0) {
    red[0][tid >> 6] = off;
    red[1][tid >> 6] = tot;
    red[2][tid >> 6] = dmax;
  }
  __syncthreads();
  if (tid == 0) {
    double so = 0, st = 0, sm = 0;
    for (int w = 0; w < NT / 64; ++w) {
      so += red[0][w];
      st += red[1][w];
      sm = fmax(sm, red[2][w]);
    }
    *done = (so <= 1e-30 * st || applied == 0) ? 1 : 0;
    red[2][0] = sm;
  }
  __syncthreads();
  const double floor_ab = 1e-17 * red[2][0];
  __syncthreads();  // red is rewritten by the next test
  return floor_ab;
}

// The cyclic Jacobi sweeps on G (one block; G in device memory, the LDS holds
// a round's rotations and the next round's pivots).  Per round: the pair
// threads form their rotation from the pivots (G_pp, G_qq, G_pq) the previous
// round left in LDS and log it; then every 2 x 2 block (a, b), a <= b, of G is
// rotated on its rows by pair a and on its columns by pair b and written with
// its Hermitian mirror (b, a), and the entries that are the next round's
// pivots are also put in LDS.  The eigenvectors are not updated here: the
// logged rotations are applied to V afterwards (k_jacobi_vectors), row by
// row in parallel.
__global__ __launch_bounds__(kJT) void k_jacobi(c64d* __restrict__ G, int S, int max_sweeps,
                                                JacobiRot* __restrict__ log,
                                                double* __restrict__ evals,
                                                int* __restrict__ sweeps_out) {
  __shared__ JacobiRot rot[kMaxPairs / 2];
  __shared__ int rp[kMaxPairs / 2], rq[kMaxPairs / 2];
  __shared__ double piv_a[kMaxPairs / 2], piv_d[kMaxPairs / 2];
  __shared__ c64d piv_b[kMaxPairs / 2];
  __shared__ short slot_of[kMaxPairs];  // next round: pair index * 2 + (0: p, 1: q)
  __shared__ double red[3][kJT / 64];
  __shared__ int done;
  const int tid = threadIdx.x;
  const int n = S + (S & 1);  // even number of players; index S is a dummy when S is odd
  const int half = n / 2;
  const long long n_blk = (long long)half * (half + 1) / 2;  // blocks (a, b) with a <= b
  int sweep = 0, applied = 1;
  for (; sweep < max_sweeps; ++sweep) {
    // convergence test: off-diagonal vs total Frobenius mass, or a sweep that
    // rotated nothing; the largest diagonal entry sets the noise floor
    double off = 0, tot = 0, dmax = 0;
    for (long long e = tid; e < (long long)S * S; e += kJT) {
      const int r = (int)(e / S), c = (int)(e % S);
      const double m = G[e].re * G[e].re + G[e].im * G[e].im;
      tot += m;
      if (r != c) off += m;
      else dmax = fmax(dmax, fabs(G[e].re));
    }
    const double floor_ab = jacobi_test<kJT>(off, tot, dmax, applied, red, &done);
    if (done) break;
    applied = 0;
    // round 0's pivots from G
    for (int i = tid; i < half; i += kJT) {
      int p, q;
      jacobi_pair(0, i, n, &p, &q);
      piv_a[i] = G[(long long)p * S + p].re;
      piv_d[i] = q < S ? G[(long long)q * S + q].re : 0.0;
      piv_b[i] = q < S ? G[(long long)p * S + q] : c64d{0, 0};
    }
    __syncthreads();
    for (int round = 0; round < n - 1; ++round) {
      for (int i = tid; i < half; i += kJT) {
        int p, q;
        jacobi_pair(round, i, n, &p, &q);
        double c = 1, s = 0;
        c64d ea = c64d{1, 0};
        if (q < S) {
          const double a = piv_a[i], d = piv_d[i];
          const c64d b = piv_b[i];
          const double ab = sqrt(b.re * b.re + b.im * b.im);
          if (jacobi_rotates(ab, a, d, floor_ab)) {
            const double tau = (d - a) / (2.0 * ab);
            const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
            ea = c64d{b.re / ab, b.im / ab};
          }
        }
        rp[i] = p;
        rq[i] = q;
        rot[i] = JacobiRot{c, s, ea};
        log[((long long)sweep * (n - 1) + round) * half + i] = JacobiRot{c, s, ea};
        // where each player sits in the next round (round n - 1 is round 0 again)
        int pn, qn;
        jacobi_pair((round + 1) % (n - 1), i, n, &pn, &qn);
        slot_of[pn] = (short)(2 * i);
        slot_of[qn] = (short)(2 * i + 1);
      }
      applied += __syncthreads_count(tid < half && rot[tid < half ? tid : 0].s != 0.0);
      for (long long w0 = tid; w0 < n_blk; w0 += (long long)kJU * kJT) {
        c64d g[kJU][4];
        int ra[kJU], rb[kJU], ca[kJU], cb[kJU], pa[kJU], pb[kJU];
#pragma unroll
        for (int u = 0; u < kJU; ++u) {
          const long long w = w0 + (long long)u * kJT;
          pa[u] = -1;
          if (w < n_blk) {
            int a, b;
            jacobi_task(w, half, &a, &b);
            pa[u] = a;
            pb[u] = b;
            ra[u] = rp[a];
            rb[u] = rq[a] < S ? rq[a] : -1;
            ca[u] = rp[b];
            cb[u] = rq[b] < S ? rq[b] : -1;
            g[u][0] = G[(long long)ra[u] * S + ca[u]];
            g[u][1] = cb[u] >= 0 ? G[(long long)ra[u] * S + cb[u]] : c64d{0, 0};
            g[u][2] = rb[u] >= 0 ? G[(long long)rb[u] * S + ca[u]] : c64d{0, 0};
            g[u][3] = (rb[u] >= 0 && cb[u] >= 0) ? G[(long long)rb[u] * S + cb[u]] : c64d{0, 0};
          }
        }
#pragma unroll
        for (int u = 0; u < kJU; ++u) {
          if (pa[u] < 0) continue;
          const int a = pa[u], b = pb[u];
          if (rb[u] >= 0 && rot[a].s != 0.0) {  // rows: G <- U_a^H G
            const double c = rot[a].c, s_ = rot[a].s;
            const c64d ea = rot[a].e;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              const c64d gp = g[u][j], gq = g[u][2 + j];
              g[u][j] = scale(gp, c) - scale(ea * gq, s_);
              g[u][2 + j] = scale(mulc(gp, ea), s_) + scale(gq, c);
            }
          }
          if (cb[u] >= 0 && rot[b].s != 0.0) {  // columns: G <- G U_b
            const double c = rot[b].c, s_ = rot[b].s;
            const c64d eb = rot[b].e;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
              const c64d gp = g[u][2 * i], gq = g[u][2 * i + 1];
              g[u][2 * i] = scale(gp, c) - scale(mulc(gq, eb), s_);
              g[u][2 * i + 1] = scale(eb * gp, s_) + scale(gq, c);
            }
          }
          const int rows[2] = {ra[u], rb[u]}, cols[2] = {ca[u], cb[u]};
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              const int r = rows[i], c = cols[j];
              if (r < 0 || c < 0) continue;
              const c64d v = g[u][2 * i + j];
              G[(long long)r * S + c] = v;
              if (a != b) G[(long long)c * S + r] = conj(v);
              // the next round's pivots: entries (r, c) of one next-round pair
              const int sr = slot_of[r], sc = slot_of[c];
              if ((sr >> 1) == (sc >> 1)) {
                const int k = sr >> 1;
                if (r == c) {
                  if (sr & 1) piv_d[k] = v.re; else piv_a[k] = v.re;
                } else if (!(sr & 1)) {  // r is the pair's p, c its q
                  piv_b[k] = v;
                } else if (a != b) {     // the mirror (c, r) is (p, q)
                  piv_b[k] = conj(v);
                }
              }
            }
        }
      }
      __syncthreads();
    }
  }
  for (int i = tid; i < S; i += kJT) evals[i] = G[(long long)i * S + i].re;
  if (tid == 0) *sweeps_out = sweep;
}

// The same sweeps with G held in LDS: the strict upper triangle packed row by
// row (complex), the real diagonal beside it, and a round's rotations as
// (c, s e) with their players -- 16 S (S - 1) / 2 + 8 S + 28 n / 2 bytes, which
// fits a CU's 160 KB for S <= 142.  Every unordered entry {r, c} belongs to
// exactly one 2 x 2 block (a, b), a <= b, so a round reads and writes each slot
// from one thread and needs no second copy; the pivots are read straight from
// the LDS matrix, and a block whose two pairs both stay unrotated is skipped.

__device__ __forceinline__ int tri_idx(int r, int c, int S) {  // r < c < S < 2^12
  return __mul24(r, S) - (__mul24(r, r + 1) >> 1) + (c - r - 1);
}

// the rotation cosines' slots, which the sweep test's partial sums also use
__host__ __device__ inline int jacobi_rc_slots(int half) {
  return half > 3 * 16 ? half : 3 * 16;  // 3 x (1024 / 64) partials
}

size_t jacobi_lds_bytes(int S) {
  const int n = S + (S & 1), half = n / 2;
  return sizeof(c64d) * ((size_t)S * (S - 1) / 2 + (size_t)half) +
         sizeof(double) * ((size_t)S + (size_t)jacobi_rc_slots(half)) +
         2 * sizeof(short) * (size_t)half;
}

constexpr int kJTL = 1024;  // threads of the LDS form (jacobi_rc_slots holds 3 x kJTL / 64)
constexpr int kJUL = 3;     // its blocks per thread: kJUL * kJTL >= 71 * 72 / 2
__global__ __launch_bounds__(kJTL) void k_jacobi_lds(const c64d* __restrict__ G, int S,
                                                    int max_sweeps, JacobiRot* __restrict__ log,
                                                    double* __restrict__ evals,
                                                    int* __restrict__ sweeps_out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int tid = threadIdx.x;
  const int n = S + (S & 1), half = n / 2;
  const int n_off = S * (S - 1) / 2;
  c64d* Go = reinterpret_cast<c64d*>(smem_raw);  // packed strict upper triangle
  c64d* rse = Go + n_off;                          // this round's rotations: s e
  double* Gd = reinterpret_cast<double*>(rse + half);  // diagonal
  double* rc = Gd + S;                                 // this round's rotations: c
  short* rp = reinterpret_cast<short*>(rc + jacobi_rc_slots(half));  // this round's players
  short* rq = rp + half;                               //   (-1: the dummy)
  __shared__ int done;
  for (int e = tid; e < S * S; e += kJTL) {
    const int r = e / S, c = e - (e / S) * S;
    if (r == c) Gd[r] = G[e].re;
    else if (r < c) Go[tri_idx(r, c, S)] = G[e];
  }
  __syncthreads();
  const long long n_blk = (long long)half * (half + 1) / 2;
  // this thread's blocks, the same in every round (-1: none): the half diagonal
  // blocks first, so that they share a few wavefronts and the rest take the
  // off-diagonal path only
  int ta[kJUL], tb[kJUL];
#pragma unroll
  for (int u = 0; u < kJUL; ++u) {
    const long long w = tid + (long long)u * kJTL;
    ta[u] = tb[u] = -1;
    if (w < half) {
      ta[u] = tb[u] = (int)w;
    } else if (w < n_blk) {  // strict upper triangle, row-major
      int a, b;
      jacobi_task(w - half, half - 1, &a, &b);
      ta[u] = a;
      tb[u] = b + 1;
    }
  }
  int sweep = 0, applied = 1;
  for (; sweep < max_sweeps; ++sweep) {
    double off = 0, tot = 0, dmax = 0;
    for (int e = tid; e < n_off; e += kJTL) off += 2.0 * norm2(Go[e]);
    for (int e = tid; e < S; e += kJTL) {
      tot += Gd[e] * Gd[e];
      dmax = fmax(dmax, fabs(Gd[e]));
    }
    tot += off;
    // the test's partial sums borrow the rotation cosines' slots (written after it)
    const double floor_ab = jacobi_test<kJTL>(off, tot, dmax, applied,
                                              reinterpret_cast<double (*)[kJTL / 64]>(rc),
                                              &done);
    if (done) break;
    applied = 0;
    for (int round = 0; round < n - 1; ++round) {
      for (int i = tid; i < half; i += kJTL) {
        int p, q;
        jacobi_pair(round, i, n, &p, &q);
        double c = 1, s = 0;
        c64d ea = c64d{1, 0};
        if (q < S) {
          const double a = Gd[p], d = Gd[q];
          const c64d b = Go[tri_idx(p, q, S)];
          const double ab = sqrt(b.re * b.re + b.im * b.im);
          if (jacobi_rotates(ab, a, d, floor_ab)) {
            const double tau = (d - a) / (2.0 * ab);
            const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
            ea = c64d{b.re / ab, b.im / ab};
          }
        }
        rc[i] = c;
        rse[i] = scale(ea, s);
        rp[i] = (short)p;
        rq[i] = (short)(q < S ? q : -1);
        log[((long long)sweep * (n - 1) + round) * half + i] = JacobiRot{c, s, ea};
      }
      applied += __syncthreads_count(tid < half && (rse[tid < half ? tid : 0].re != 0.0 ||
                                                    rse[tid < half ? tid : 0].im != 0.0));
#pragma unroll
      for (int u = 0; u < kJUL; ++u) {
        const int a = ta[u], b = tb[u];
        if (a < 0) continue;
        const c64d sa = rse[a], sb = rse[b];
        const bool rot_a = sa.re != 0.0 || sa.im != 0.0;
        const bool rot_b = sb.re != 0.0 || sb.im != 0.0;
        if (!rot_a && !rot_b) continue;  // both pairs left as they are: the block is too
        const double ca_ = rc[a], cb_ = rc[b];
        const int ra = rp[a], rb = rq[a];
        if (a == b) {
          // the pair's own 2 x 2 block [[G_pp, G_pq], [conj G_pq, G_qq]]: U^H B U
          if (rb < 0) continue;  // paired with the dummy: nothing to rotate
          const int spq = tri_idx(ra, rb, S);
          c64d g0 = c64d{Gd[ra], 0.0}, g1 = Go[spq], g2 = conj(g1), g3 = c64d{Gd[rb], 0.0};
          {
            const c64d n0 = scale(g0, ca_) - sa * g2, n2 = mulc(g0, sa) + scale(g2, ca_);
            const c64d n1 = scale(g1, ca_) - sa * g3, n3 = mulc(g1, sa) + scale(g3, ca_);
            g0 = n0; g1 = n1; g2 = n2; g3 = n3;
          }
          const c64d m0 = scale(g0, ca_) - mulc(g1, sa), m1 = sa * g0 + scale(g1, ca_);
          const c64d m3 = sa * g2 + scale(g3, ca_);
          Gd[ra] = m0.re;
          Gd[rb] = m3.re;
          Go[spq] = m1;
          continue;
        }
        // off-diagonal block: rows {p_a, q_a} x columns {p_b, q_b}, all players distinct
        const int ca = rp[b], cb = rq[b];
        const int rows[2] = {ra, rb}, cols[2] = {ca, cb};
        int sl[4];
        bool cj[4];
        c64d g[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int r = rows[k >> 1], c = cols[k & 1];
          cj[k] = r > c;
          sl[k] = (r < 0 || c < 0) ? -1 : (r < c ? tri_idx(r, c, S) : tri_idx(c, r, S));
          const c64d v = sl[k] >= 0 ? Go[sl[k]] : c64d{0, 0};
          g[k] = cj[k] ? conj(v) : v;
        }
        if (rb >= 0 && rot_a) {  // rows: G <- U_a^H G
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            const c64d gp = g[j], gq = g[2 + j];
            g[j] = scale(gp, ca_) - sa * gq;
            g[2 + j] = mulc(gp, sa) + scale(gq, ca_);
          }
        }
        if (cb >= 0 && rot_b) {  // columns: G <- G U_b
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            const c64d gp = g[2 * i], gq = g[2 * i + 1];
            g[2 * i] = scale(gp, cb_) - mulc(gq, sb);
            g[2 * i + 1] = sb * gp + scale(gq, cb_);
          }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (sl[k] >= 0) Go[sl[k]] = cj[k] ? conj(g[k]) : g[k];
      }
      __syncthreads();
    }
  }
  for (int i = tid; i < S; i += kJTL) evals[i] = Gd[i];
  if (tid == 0) *sweeps_out = sweep;
}

// V = product of the logged rotations (V starts as the identity): every row
// of V is independent, so one wavefront per row holds it in LDS and applies
// the rounds in order.  The block stages the rotations of kJVR rounds at a
// time in LDS (one cooperative load, double-buffered), which its waves share.
constexpr int kJVW = 8;  // rows (wavefronts) per block, at most
__global__ __launch_bounds__(64 * kJVW) void k_jacobi_vectors(c64d* __restrict__ V, int S,
                                                              const JacobiRot* __restrict__ log,
                                                              const int* __restrict__ sweeps,
                                                              int nrc) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int wv = (int)(blockDim.x >> 6);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = S + (S & 1), half = n / 2;
  JacobiRot* stage = reinterpret_cast<JacobiRot*>(smem_raw);  // 2 x nrc x half
  c64d* rows = reinterpret_cast<c64d*>(stage + 2LL * nrc * half);
  const int k = blockIdx.x * wv + wave;
  c64d* row = rows + (long long)wave * (S + 1);
  const bool live = k < S;
  if (live)
    for (int c = lane; c < S; c += 64) row[c] = c64d{c == k ? 1.0 : 0.0, 0.0};
  const long long rounds = (long long)(*sweeps) * (n - 1);
  const long long chunks = (rounds + nrc - 1) / nrc;
  auto load = [&](long long ch, int buf) {
    const long long r0 = ch * nrc;
    const int nr = (int)min<long long>(nrc, rounds - r0);
    JacobiRot* dst = stage + (long long)buf * nrc * half;
    for (int e = threadIdx.x; e < nr * half; e += (int)blockDim.x) dst[e] = log[r0 * half + e];
  };
  if (chunks > 0) load(0, 0);
  __syncthreads();
  for (long long ch = 0; ch < chunks; ++ch) {
    if (ch + 1 < chunks) load(ch + 1, (int)((ch + 1) & 1));  // in flight during this chunk
    const long long r0 = ch * nrc;
    const int nr = (int)min<long long>(nrc, rounds - r0);
    const JacobiRot* src = stage + (long long)(ch & 1) * nrc * half;
    if (live) {
      for (int t = 0; t < nr; ++t) {
        const int r = (int)((r0 + t) % (n - 1));
        for (int i = lane; i < half; i += 64) {
          int p, q;
          jacobi_pair(r, i, n, &p, &q);
          const JacobiRot u = src[t * half + i];
          if (q >= S || u.s == 0.0) continue;
          const c64d vp = row[p], vq = row[q];
          row[p] = scale(vp, u.c) - scale(mulc(vq, u.e), u.s);
          row[q] = scale(u.e * vp, u.s) + scale(vq, u.c);
        }
      }
    }
    // every wave is done with this buffer (reloaded next chunk) and the next one has landed
    __syncthreads();
  }
  if (live)
    for (int c = lane; c < S; c += 64) V[(long long)k * S + c] = row[c];
}

size_t jacobi_log_bytes(int S, int max_sweeps) {
  const int n = S + (S & 1);
  return sizeof(JacobiRot) * (size_t)max_sweeps * (size_t)(n - 1) * (size_t)(n / 2);
}

constexpr size_t kJacobiLdsMax = 160 * 1024 - 256;  // dynamic LDS beside the static slots

hipError_t jacobi_eigh_f64(c64d* G, c64d* V, int S, int max_sweeps, void* log, double* evals,
                           int* sweeps_out, hipStream_t st) {
  if (S > kMaxPairs - 1) return hipErrorInvalidValue;
  JacobiRot* lg = (JacobiRot*)log;
  hipError_t e;
  const size_t lds = jacobi_lds_bytes(S);
  const long long half = (S + (S & 1)) / 2;
  bool in_lds = S >= 2 && lds <= kJacobiLdsMax && half * (half + 1) / 2 <= (long long)kJUL * kJTL;
  if (in_lds) {
    // a device that grants less LDS per block takes the device-memory form
    in_lds = hipFuncSetAttribute(reinterpret_cast<const void*>(k_jacobi_lds),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds) == hipSuccess;
    (void)hipGetLastError();
  }
  if (in_lds) {
    hipLaunchKernelGGL(k_jacobi_lds, dim3(1), dim3(kJTL), lds, st, G, S, max_sweeps, lg, evals,
                       sweeps_out);
  } else {
    hipLaunchKernelGGL(k_jacobi, dim3(1), dim3(kJT), 0, st, G, S, max_sweeps, lg, evals,
                       sweeps_out);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // eigenvector pass: up to kJVW rows per block and 16 staged rounds, fewer where the
  // rows and the two staging buffers would not fit in LDS
  const int n = S + (S & 1);
  const size_t per_round = sizeof(JacobiRot) * 2 * (size_t)(n / 2);
  int wv = kJVW, nrc = 16;
  while (wv > 1 && sizeof(c64d) * wv * (size_t)(S + 1) + per_round > kJacobiLdsMax) wv /= 2;
  const size_t rows_b = sizeof(c64d) * wv * (size_t)(S + 1);
  if (rows_b + per_round > kJacobiLdsMax) return hipErrorInvalidValue;
  nrc = (int)std::min<size_t>(16, (kJacobiLdsMax - rows_b) / per_round);
  const size_t sm = rows_b + per_round * nrc;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_jacobi_vectors),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_jacobi_vectors, dim3((S + wv - 1) / wv), dim3(64 * wv), sm, st, V, S, lg,
                     sweeps_out, nrc);
  return hipGetLastError();
}

__global__ void k_phi(const c64d* __restrict__ A, const c64d* __restrict__ V, int S, long long F,
                      const int* __restrict__ idx, const double* __restrict__ inv_sqrt,
                      c64d* __restrict__ phi) {
  const int k = blockIdx.y;
  const int col = idx[k];
  const double sc = inv_sqrt[k];
  for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < F;
       f += (long long)gridDim.x * blockDim.x) {
    double re = 0, im = 0;
    for (int s = 0; s < S; ++s) {
      const c64d a = A[(long long)s * F + f], u = V[(long long)s * S + col];
      re += a.re * u.re - a.im * u.im;
      im += a.re * u.im + a.im * u.re;
    }
    phi[(long long)k * F + f] = c64d{re * sc, im * sc};
  }
}

hipError_t socs_phi_f64(const c64d* A, const c64d* V, int S, long long F, const int* idx,
                        const double* inv_sqrt, int K, c64d* phi, hipStream_t st) {
  if (K <= 0) return hipSuccess;
  long long bx = (F + 255) / 256;
  if (bx > 1024) bx = 1024;
  hipLaunchKernelGGL(k_phi, dim3((unsigned)bx, K), dim3(256), 0, st, A, V, S, F, idx, inv_sqrt,
                     phi);
  return hipGetLastError();
}

__global__ void k_cast(const c64d* __restrict__ in, cx<float>* __restrict__ out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = cx<float>{(float)in[i].re, (float)in[i].im};
}

hipError_t cast_c128_c64(const c64d* in, cx<float>* out, long long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long long b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  hipLaunchKernelGGL(k_cast, dim3((unsigned)b), dim3(256), 0, st, in, out, n);
  return hipGetLastError();
}

}  // namespace lb
