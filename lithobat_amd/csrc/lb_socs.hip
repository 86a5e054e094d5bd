// Hopkins / SOCS decomposition in FP64 on the device.
//
// The reference forms TCC = A^T conj(A) over the band (n_band^2, 1.7 GB at a
// 512^2 / NA 1.35 grid) and calls eigh on it (expose/hopkins.py:155-186).
// TCC = B B^H with B = A^T (n_band x S), so its non-zero eigenpairs are those
// of the S x S Gram matrix G = B^H B (G[s][t] = sum_f conj(A[s][f]) A[t][f]):
// G u = lambda u  <=>  TCC (B u) = lambda (B u), and phi = B u / sqrt(lambda)
// is unit-norm.  The Gram matrix is built here, diagonalised by a
// Householder tridiagonalisation followed by the implicit QL iteration, and
// the kept eigenvectors are mapped back to kernels.
#include <algorithm>

#include "lb_kernels.h"

namespace lb {

using c64d = cx<double>;

constexpr int kGT = 16;

// One 16 x 16 tile of G over one slice [z fs, z fs + fs) of the band (blockIdx.z):
// partial Gram matrices, summed in slice order by k_gram_sum (deterministic).
__global__ __launch_bounds__(kGT* kGT) void k_gram(const c64d* __restrict__ A, int S, long long F,
                                                   long long fs, c64d* __restrict__ Gp) {
  __shared__ c64d As[kGT][kGT + 1];
  __shared__ c64d Bs[kGT][kGT + 1];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int s = blockIdx.y * kGT + ty;  // row of G
  const int t = blockIdx.x * kGT + tx;  // col of G
  const long long fa = (long long)blockIdx.z * fs;
  const long long fb = fa + fs < F ? fa + fs : F;
  double re = 0, im = 0;
  for (long long f0 = fa; f0 < fb; f0 += kGT) {
    const int sr = blockIdx.y * kGT + ty, tr = blockIdx.x * kGT + ty;
    const long long f = f0 + tx;
    As[ty][tx] = (sr < S && f < fb) ? A[(long long)sr * F + f] : c64d{0, 0};
    Bs[ty][tx] = (tr < S && f < fb) ? A[(long long)tr * F + f] : c64d{0, 0};
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
  if (s < S && t < S) Gp[(long long)blockIdx.z * S * S + (long long)s * S + t] = c64d{re, im};
}

__global__ void k_gram_sum(const c64d* __restrict__ Gp, int ks, long long ss,
                           c64d* __restrict__ G) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < ss;
       e += (long long)gridDim.x * blockDim.x) {
    c64d acc = Gp[e];
    for (int z = 1; z < ks; ++z) acc = acc + Gp[(long long)z * ss + e];
    G[e] = acc;
  }
}

// Slices of the band: enough tiles x slices for about 1024 blocks, at least
// 64 band entries per slice.
static int gram_slices(int S, long long F) {
  const long long nb = (S + kGT - 1) / kGT;
  long long ks = (1024 + nb * nb - 1) / (nb * nb);
  const long long cap = (F + 63) / 64;
  if (ks > cap) ks = cap;
  return (int)(ks < 1 ? 1 : ks);
}

size_t gram_work_bytes(int S, long long F) {
  return sizeof(c64d) * (size_t)gram_slices(S, F) * (size_t)S * (size_t)S;
}

hipError_t gram_f64(const c64d* A, int S, long long F, c64d* G, void* work, hipStream_t st) {
  if (S <= 0) return hipSuccess;
  const int ks = gram_slices(S, F);
  long long fs = (F + ks - 1) / ks;
  fs = (fs + kGT - 1) / kGT * kGT;
  c64d* Gp = static_cast<c64d*>(work);
  dim3 g((S + kGT - 1) / kGT, (S + kGT - 1) / kGT, ks);
  hipLaunchKernelGGL(k_gram, g, dim3(kGT, kGT), 0, st, A, S, F, fs, Gp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long long ss = (long long)S * S;
  const long long b = std::min<long long>((ss + 255) / 256, 1024);
  hipLaunchKernelGGL(k_gram_sum, dim3((unsigned)b), dim3(256), 0, st, Gp, ks, ss, G);
  return hipGetLastError();
}

// ---------------------------------------------------------------- eigensolve
// G = Q T Q^H (Householder tridiagonalisation, T real symmetric), T = Z L Z^T
// (implicit QL with Wilkinson shifts), so the eigenvectors are the columns of
// V = Q Z.  The QL runs on one thread (its rotations form one dependent chain)
// and logs every rotation; the rows of Z, which the rotations treat
// independently, replay the log in parallel afterwards, and the reflectors are
// applied to the columns of Z last.  Householder, "Unitary triangularization
// of a nonsymmetric matrix" (J. ACM 5, 1958) in the Hermitian reduction form
// of LAPACK's zhetd2 (lower); the QL iteration of Bowdler, Martin, Reinsch and
// Wilkinson, "The QR and QL algorithms for symmetric matrices" (Numer. Math.
// 11, 1968).

__device__ __forceinline__ double wave_sum(double x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

constexpr int kHT = 1024;  // threads of the tridiagonalisation block

// One block.  A (n x n, full Hermitian storage, row-major) is reduced in
// place: step i forms the reflector H(i) = I - tau[i] v v^H (v[0] = 1 at
// index i + 1, stored in row i of R) with H(i)^H A H(i) zero in column i below
// the subdiagonal, whose entry becomes e[i] (real).  Row i of A holds the
// conjugate of column i, so that column is read from the row (contiguous).
// Every wavefront forms the reflector itself, its lanes holding v at
// c = lane + 64 q (NV slots, n <= 64 NV), so a step needs two barriers: after
// p = tau A22 v, and after the rank-2 update.  From step i_lds on, the
// trailing matrix A[i_lds:, i_lds:] is held in LDS.
template <int NV>
__global__ __launch_bounds__(kHT) void k_hetrd(c64d* __restrict__ A, int n, int i_lds,
                                               double* __restrict__ d, double* __restrict__ e,
                                               c64d* __restrict__ tau, c64d* __restrict__ R) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  c64d* p = reinterpret_cast<c64d*>(smem_raw);  // n: tau A22 v
  c64d* L = p + n;                              // (n - i_lds)^2 from step i_lds on
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NW = kHT / 64;
  const int nt = n - i_lds;
  for (int i = 0; i + 1 < n; ++i) {
    if (i == i_lds) {
      for (int q = tid; q < nt * nt; q += kHT) {
        const int a = q / nt, b = q - a * nt;
        L[q] = A[(long long)(i_lds + a) * n + i_lds + b];
      }
      __syncthreads();
    }
    const bool in_lds = i >= i_lds;
    // Ai[j] = A[i][j] (j >= i); A22[r * ld + c] = A[i+1+r][i+1+c]
    const c64d* Ai = in_lds ? L + (long long)(i - i_lds) * nt - i_lds : A + (long long)i * n;
    c64d* A22 = in_lds ? L + (long long)(i + 1 - i_lds) * nt + (i + 1 - i_lds)
                       : A + (long long)(i + 1) * n + (i + 1);
    const int ld = in_lds ? nt : n;
    const int m = n - i - 1;  // order of A22
    c64d x[NV];  // column i below the subdiagonal, at c = lane + 64 q (c >= 1)
    double part = 0;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const int c = lane + 64 * q;
      x[q] = (c >= 1 && c < m) ? conj(Ai[i + 1 + c]) : c64d{0, 0};
      part += norm2(x[q]);
    }
    const double xn2 = wave_sum(part);
    const c64d alpha = conj(Ai[i + 1]);
    c64d t = c64d{0, 0}, scal = c64d{0, 0};
    double beta = alpha.re;
    if (xn2 != 0.0 || alpha.im != 0.0) {
      const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
      beta = alpha.re >= 0.0 ? -nrm : nrm;
      t = c64d{(beta - alpha.re) / beta, -alpha.im / beta};
      const c64d den = c64d{alpha.re - beta, alpha.im};  // scal = 1 / (alpha - beta)
      const double dd = den.re * den.re + den.im * den.im;
      scal = c64d{den.re / dd, -den.im / dd};
    }
    c64d v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const int c = lane + 64 * q;
      v[q] = c == 0 ? c64d{1, 0} : x[q] * scal;
      if (wave == 0 && c < m) R[(long long)i * n + i + 1 + c] = v[q];
    }
    if (tid == 0) {
      d[i] = Ai[i].re;
      e[i] = beta;
      tau[i] = t;
    }
    // every wave formed the same t from the same values in the same order
    if (t.re == 0.0 && t.im == 0.0) continue;
    for (int r = wave; r < m; r += NW) {
      const c64d* row = A22 + (long long)r * ld;
      double re = 0, im = 0;
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const int c = lane + 64 * q;
        if (c < m) {
          const c64d a = row[c];
          re += a.re * v[q].re - a.im * v[q].im;
          im += a.re * v[q].im + a.im * v[q].re;
        }
      }
      re = wave_sum(re);
      im = wave_sum(im);
      if (lane == 0) p[r] = t * c64d{re, im};
    }
    __syncthreads();
    // a2 = -tau/2 (p^H v), formed by every wave; w = p + a2 v
    double re = 0, im = 0;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const int c = lane + 64 * q;
      if (c < m) {
        const c64d a = p[c];
        re += a.re * v[q].re + a.im * v[q].im;
        im += a.re * v[q].im - a.im * v[q].re;
      }
    }
    re = wave_sum(re);
    im = wave_sum(im);
    const c64d a2 = scale(t * c64d{re, im}, -0.5);
    c64d w[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const int c = lane + 64 * q;
      w[q] = c < m ? p[c] + a2 * v[q] : c64d{0, 0};
    }
    // A22 <- A22 - v w^H - w v^H (the diagonal kept real)
    for (int r = wave; r < m; r += NW) {
      // v_r, w_r from lane r % 64, slot r / 64 (uniform over the wave)
      const int qr = r >> 6, lr = r & 63;
      c64d vr = c64d{0, 0}, wr = c64d{0, 0};
#pragma unroll
      for (int q = 0; q < NV; ++q)
        if (q == qr) {
          vr = v[q];
          wr = w[q];
        }
      vr = c64d{__shfl(vr.re, lr, 64), __shfl(vr.im, lr, 64)};
      wr = c64d{__shfl(wr.re, lr, 64), __shfl(wr.im, lr, 64)};
      c64d* row = A22 + (long long)r * ld;
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const int c = lane + 64 * q;
        if (c < m) {
          c64d val = row[c] - (mulc(vr, w[q]) + mulc(wr, v[q]));
          if (c == r) val.im = 0.0;
          row[c] = val;
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0 && n > 0)
    d[n - 1] = (i_lds <= n - 2 ? L[(long long)(nt - 1) * nt + (nt - 1)]
                               : A[(long long)(n - 1) * n + (n - 1)]).re;
}

// One QL sweep of the log: rotations at i = m - 1, m - 2, ..., m - n_rot,
// their (c, s) at rots[off ...].
struct QLSweep {
  int l, m, n_rot, off;
};

// 1 / sqrt(x) for x > 0: the hardware estimate and two Newton steps.
__device__ __forceinline__ double rsqrt_refined(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double h = x * y;
    y = fma(0.5 * y, fma(-h, y, 1.0), y);
  }
  return y;
}

// Implicit QL on the real symmetric tridiagonal (d, e), e[i] coupling i and
// i + 1, on one thread.  d becomes the eigenvalues; every rotation is logged.
// counts = {sweeps, rotations, status}: status 1 when an eigenvalue takes more
// than 60 iterations, 2 when the log is full (both reported as errors).  The
// rotations form one dependent chain, so each is kept short: the next
// rotation's d and e are loaded ahead and 1 / r comes from a refined
// reciprocal square root instead of a square root and a division.
//
// With a split point h > 0 (divide and conquer), block b solves part b of
// the torn matrix, T = diag(T1', T2') + beta u u^T with beta = e[h-1] and
// u = e_{h-1} + e_h: part 0 is rows [0, h) with d[h-1] - beta, part 1 rows
// [h, n) with d[h] - beta.  Each part has its own log (local indices) and
// counts.
__global__ void k_tql(double* __restrict__ d_io, const double* __restrict__ e_in, int n_all,
                      int h, QLSweep* __restrict__ sweeps_all, double2* __restrict__ rots_all,
                      int max_sweeps, int max_rots, int* __restrict__ counts_all) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int part = blockIdx.x;
  const int off = part == 0 ? 0 : h;
  const int n = h == 0 ? n_all : (part == 0 ? h : n_all - h);
  QLSweep* sweeps = sweeps_all + (long long)part * max_sweeps;
  double2* rots = rots_all + (long long)part * max_rots;
  int* counts = counts_all + 4 * part;
  double* d = reinterpret_cast<double*>(smem_raw);  // n
  double* e = d + n;                                // n + 3 (e[n-1 ...] = 0)
  const double beta = h > 0 ? e_in[h - 1] : 0.0;
  for (int i = threadIdx.x; i < n + 3; i += blockDim.x) {
    if (i < n) {
      double di = d_io[off + i];
      if (h > 0 && ((part == 0 && i == n - 1) || (part == 1 && i == 0))) di -= beta;
      d[i] = di;
    }
    e[i] = i + 1 < n ? e_in[off + i] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  // A coupling deflates once it is below the rounding level of the whole
  // matrix (eps times its largest row sum).  A test relative to the two
  // neighbouring diagonal entries never passes in a block of rounding-level
  // entries, which a rank-deficient Gram matrix has (more source points than
  // band frequencies): its near-zero eigenvalues would never converge.
  const double eps = 2.220446049250313e-16;
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i)
    tnorm = fmax(tnorm, fabs(d[i]) + fabs(e[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0));
  const double tol = eps * tnorm;
  int n_sw = 0, n_rot = 0, status = 0;
  for (int l = 0; l < n && status == 0; ++l) {
    int iter = 0;
    for (;;) {
      // first m >= l with a negligible coupling (e[n-1] = 0 ends the search);
      // four couplings per step, loaded together
      int m = l;
      for (;;) {
        const double e0 = e[m], e1 = e[m + 1], e2 = e[m + 2], e3 = e[m + 3];
        if (fabs(e0) <= tol) break;
        if (fabs(e1) <= tol) { m += 1; break; }
        if (fabs(e2) <= tol) { m += 2; break; }
        if (fabs(e3) <= tol) { m += 3; break; }
        m += 4;
      }
      if (m > n - 1) m = n - 1;
      if (m == l) break;
      if (++iter > 60) {
        status = 1;
        break;
      }
      if (n_sw >= max_sweeps || n_rot + (m - l) > max_rots) {
        status = 2;
        break;
      }
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = sqrt(g * g + 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      const int off = n_rot;
      bool under = false;
      double di1 = d[m];
      // rotations i = m-1 down to l in chunks of up to 8: the chunk's d and e
      // are loaded together and its outputs stored together, so the chain
      // waits on LDS once per chunk
      int i = m - 1;
      while (i >= l && !under) {
        const int nk = i - l + 1 < 8 ? i - l + 1 : 8;
        double dv[8], ev[8], dout[8], eout[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int j = i - k >= 0 ? i - k : 0;
          dv[k] = d[j];
          ev[k] = e[j];
        }
        int done = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          if (k < nk && !under) {
            const double f = s * ev[k], b = c * ev[k];
            const double rr = fma(f, f, g * g);
            if (rr == 0.0) {  // the coupling underflowed: split here and start again
              eout[k] = 0.0;
              dout[k] = di1 - p;
              under = true;
            } else {
              const double inv = rsqrt_refined(rr);
              eout[k] = rr * inv;
              s = f * inv;
              c = g * inv;
              g = di1 - p;
              r = fma(dv[k] - g, s, 2.0 * c * b);
              p = s * r;
              dout[k] = g + p;
              g = fma(c, r, -b);
              rots[n_rot + k] = make_double2(c, s);
              di1 = dv[k];
              ++done;
            }
          }
        }
        // outputs of rotation k: e[i-k+1], d[i-k+1] (the underflowed one too)
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (k < done + (under ? 1 : 0)) {
            e[i - k + 1] = eout[k];
            d[i - k + 1] = dout[k];
          }
        n_rot += done;
        i -= done;
      }
      if (under) e[m] = 0.0;
      sweeps[n_sw++] = QLSweep{l, m, n_rot - off, off};
      if (under) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  for (int i = 0; i < n; ++i) d_io[off + i] = d[i];
  counts[0] = n_sw;
  counts[1] = n_rot;
  counts[2] = status;
}

// Z = product of the logged rotations (Z starts as the identity), rows held
// in LDS, two rows per lane when a block's 128 rows fit (two independent
// chains per lane: the rotations are latency-bound).  A sweep's rotation at i
// turns (z[i], z[i+1]); going down a sweep, the updated z[i] is the next
// rotation's partner, so it stays in a register and each rotation reads one
// element and writes one.
// The log is read with vector loads (an offset the compiler cannot prove
// uniform, zero at run time): scalar loads share the LDS wait counter and
// return out of order, so every wait for a row element would also wait for
// them.
// With a split h > 0, blockIdx.y is the part: its rows and columns
// [off, off + n) of the n_all x n_all Z (zero outside the part's block).
template <int RPL>
__global__ __launch_bounds__(64) void k_ql_rows(int n_all, int h, int max_sweeps, int max_rots,
                                                const QLSweep* __restrict__ sweeps_all,
                                                const double2* __restrict__ rots_all,
                                                const int* __restrict__ counts_all,
                                                double* __restrict__ Z, int zero) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int part = blockIdx.y;
  const int off = part == 0 ? 0 : h;
  const int n = h == 0 ? n_all : (part == 0 ? h : n_all - h);
  if ((int)blockIdx.x * 64 * RPL >= n) return;  // uniform: this part has fewer rows
  const QLSweep* sweeps = sweeps_all + (long long)part * max_sweeps;
  const double2* rots = rots_all + (long long)part * max_rots;
  const int* counts = counts_all + 4 * part;
  const int ld = n + 1;
  double* z[RPL];
  int row[RPL];
#pragma unroll
  for (int u = 0; u < RPL; ++u) {
    z[u] = reinterpret_cast<double*>(smem_raw) + (long long)(threadIdx.x + 64 * u) * ld;
    row[u] = blockIdx.x * 64 * RPL + threadIdx.x + 64 * u;  // >= n: a spare slot
#pragma unroll 1
    for (int c = 0; c < n; ++c) z[u][c] = c == row[u] ? 1.0 : 0.0;
  }
  const int n_sw = counts[0];
  for (int w = 0; w < n_sw; ++w) {
    const QLSweep h = sweeps[w];
    const double2* rr = rots + h.off + zero * (int)threadIdx.x;
    double carry[RPL];
#pragma unroll
    for (int u = 0; u < RPL; ++u) carry[u] = z[u][h.m];
    int i = h.m - 1;
    int k = 0;
    for (; k + 4 <= h.n_rot; k += 4, i -= 4) {
      const double2 r0 = rr[k], r1 = rr[k + 1], r2 = rr[k + 2], r3 = rr[k + 3];
      double a[RPL][4];
#pragma unroll
      for (int u = 0; u < RPL; ++u)
#pragma unroll
        for (int t = 0; t < 4; ++t) a[u][t] = z[u][i - t];
#pragma unroll
      for (int u = 0; u < RPL; ++u) {
        double cy = carry[u];
        z[u][i + 1] = r0.y * a[u][0] + r0.x * cy;
        cy = r0.x * a[u][0] - r0.y * cy;
        z[u][i] = r1.y * a[u][1] + r1.x * cy;
        cy = r1.x * a[u][1] - r1.y * cy;
        z[u][i - 1] = r2.y * a[u][2] + r2.x * cy;
        cy = r2.x * a[u][2] - r2.y * cy;
        z[u][i - 2] = r3.y * a[u][3] + r3.x * cy;
        carry[u] = r3.x * a[u][3] - r3.y * cy;
      }
    }
    for (; k < h.n_rot; ++k, --i) {
      const double2 rk = rr[k];
#pragma unroll
      for (int u = 0; u < RPL; ++u) {
        const double a = z[u][i];
        z[u][i + 1] = rk.y * a + rk.x * carry[u];
        carry[u] = rk.x * a - rk.y * carry[u];
      }
    }
#pragma unroll
    for (int u = 0; u < RPL; ++u) z[u][h.m - h.n_rot] = carry[u];
  }
#pragma unroll
  for (int u = 0; u < RPL; ++u)
    if (row[u] < n) {
      double* zr = Z + (long long)(off + row[u]) * n_all;
      for (int c = 0; c < n_all; ++c)
        zr[c] = (c >= off && c < off + n) ? z[u][c - off] : 0.0;
    }
}

// ------------------------------------------------ divide and conquer merge
// The two parts' eigenpairs (T1' = Z1 D1 Z1^T, T2' = Z2 D2 Z2^T) give
// T = Q (D + beta z z^T) Q^T with Q = diag(Z1, Z2) and z = Q^T u (row h-1 of
// Z1 beside row 0 of Z2).  The rank-one problem is solved as in Cuppen, "A
// divide and conquer method for the symmetric tridiagonal eigenproblem"
// (Numer. Math. 36, 1981), with the eigenvectors from the recomputed z of Gu
// and Eisenstat, "A divide-and-conquer algorithm for the symmetric tridiagonal
// eigenproblem" (SIAM J. Matrix Anal. Appl. 16, 1995), which keeps them
// orthogonal: beta < 0 is handled by negating the problem, z is normalised
// (rho = |beta| |z|^2), D sorted ascending, entries with |rho z_i| below the
// tolerance deflate (eigenvalue d_i, unit eigenvector), and of two kept
// entries closer than the tolerance a rotation zeroes one z and deflates it.
// The K kept eigenvalues are the roots of 1 + rho sum z_i^2 / (d_i - lambda),
// one in each gap and one above the last d, found by bisection on
// tau = lambda - d_o from the nearer pole d_o (so lambda - d_i is formed from
// exact pole differences).
struct DCInfo {
  int K, n_giv;
  double rho, sgn;
};
struct DCGivens {
  int a, b;
  double c, s;
};

constexpr int kDCT = 256;

__global__ __launch_bounds__(kDCT) void k_dc_prep(const double* __restrict__ lam0,
                                                  const double* __restrict__ Z,
                                                  const double* __restrict__ e, int n, int h,
                                                  double* __restrict__ delta,
                                                  double* __restrict__ zz, int* __restrict__ perm,
                                                  int* __restrict__ kidx, int* __restrict__ kpos,
                                                  DCGivens* __restrict__ giv,
                                                  DCInfo* __restrict__ info) {
  __shared__ double red[kDCT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double beta = e[h - 1];
  const double sgn = beta < 0.0 ? -1.0 : 1.0;
  double part = 0;
  for (int i = tid; i < n; i += kDCT) {
    const double zi = Z[(long long)(h - 1) * n + i] + Z[(long long)h * n + i];
    part += zi * zi;
  }
  part = wave_sum(part);
  if (lane == 0) red[wave] = part;
  __syncthreads();
  double nz2 = 0;
  for (int w = 0; w < kDCT / 64; ++w) nz2 += red[w];
  const double inz = nz2 > 0.0 ? 1.0 / sqrt(nz2) : 0.0;
  // rank sort, ascending (ties by index)
  for (int i = tid; i < n; i += kDCT) {
    const double di = sgn * lam0[i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = sgn * lam0[j];
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    perm[rank] = i;
    delta[rank] = di;
    zz[rank] = (Z[(long long)(h - 1) * n + i] + Z[(long long)h * n + i]) * inz;
  }
  __syncthreads();
  if (tid != 0) return;
  const double rho = fabs(beta) * nz2;
  double dmax = 0;
  for (int i = 0; i < n; ++i) dmax = fmax(dmax, fabs(delta[i]));
  const double tol = 8.0 * 2.220446049250313e-16 * fmax(dmax, rho);
  int K = 0, ng = 0, last = -1;
  for (int i = 0; i < n; ++i) {
    kpos[i] = -1;
    if (!(fabs(rho * zz[i]) > tol)) {
      zz[i] = 0.0;
      continue;
    }
    if (last >= 0 && delta[i] - delta[last] <= tol) {
      // rotate (last, i) so that z[last] becomes zero; last deflates
      const double zl = zz[last], zi = zz[i];
      const double r = sqrt(zl * zl + zi * zi);
      const double c = zi / r, sn = zl / r;
      zz[i] = r;
      zz[last] = 0.0;
      const double dl = delta[last], di = delta[i];
      delta[last] = c * c * dl + sn * sn * di;
      delta[i] = sn * sn * dl + c * c * di;
      giv[ng++] = DCGivens{last, i, c, sn};
      kpos[last] = -1;
      --K;
    }
    kpos[i] = K;
    last = i;
    ++K;
  }
  // kept indices in order (kpos was assigned before later deflations: renumber)
  int k = 0;
  for (int i = 0; i < n; ++i)
    if (kpos[i] >= 0) {
      kpos[i] = k;
      kidx[k++] = i;
    }
  info->K = k;
  info->n_giv = ng;
  info->rho = rho;
  info->sgn = sgn;
}

// One wavefront per kept root j; the secular function's sum split over lanes.
__global__ __launch_bounds__(64) void k_dc_roots(const double* __restrict__ delta,
                                                 const double* __restrict__ zz,
                                                 const int* __restrict__ kidx,
                                                 const DCInfo* __restrict__ info,
                                                 int* __restrict__ ro, double* __restrict__ rt) {
  const int j = blockIdx.x, lane = threadIdx.x;
  const int K = info->K;
  if (j >= K) return;
  const double rho = info->rho;
  auto dk = [&](int i) { return delta[kidx[i]]; };
  // f(o, t) = 1 + rho sum_i z_i^2 / ((d_i - d_o) - t)
  auto f = [&](int o, double t) {
    const double dko = dk(o);
    double acc = 0;
    for (int i = lane; i < K; i += 64) {
      const double zi = zz[kidx[i]];
      acc += zi * zi / ((dk(i) - dko) - t);
    }
    return 1.0 + rho * wave_sum(acc);
  };
  double gap;
  if (j < K - 1) {
    gap = dk(j + 1) - dk(j);
  } else {
    double acc = 0;
    for (int i = lane; i < K; i += 64) acc += zz[kidx[i]] * zz[kidx[i]];
    gap = rho * wave_sum(acc);
  }
  int o = j;
  double a = 0.0, b = gap;
  if (j < K - 1) {
    const double mid = 0.5 * gap;
    if (f(j, mid) > 0.0) {
      b = mid;  // the root lies in (d_j, midpoint): origin d_j
    } else {
      o = j + 1;  // in [midpoint, d_{j+1}): origin d_{j+1}
      a = -mid;
      b = 0.0;
    }
  }
  for (int it = 0; it < 400; ++it) {
    const double t = 0.5 * (a + b);
    if (t == a || t == b) break;
    if (f(o, t) > 0.0)
      b = t;
    else
      a = t;
  }
  if (lane == 0) {
    ro[j] = o;
    rt[j] = 0.5 * (a + b);
  }
}

// lambda_j - d_i from the root's origin: (d_o - d_i) + tau
__device__ __forceinline__ double dc_gap(const double* delta, const int* kidx, const int* ro,
                                         const double* rt, int i, int j) {
  return (delta[kidx[ro[j]]] - delta[kidx[i]]) + rt[j];
}

// Recomputed z (Gu-Eisenstat): zhat_i^2 = (lam_K - d_i) prod_{j<i} (lam_j - d_i) /
// (d_j - d_i) prod_{i<=j<K-1} (lam_j - d_i) / (d_{j+1} - d_i) / rho, signed as z_i.
__global__ void k_dc_zhat(const double* __restrict__ delta, const double* __restrict__ zz,
                          const int* __restrict__ kidx, const DCInfo* __restrict__ info,
                          const int* __restrict__ ro, const double* __restrict__ rt,
                          double* __restrict__ zh) {
  const int K = info->K;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K) return;
  const double di = delta[kidx[i]];
  double v = dc_gap(delta, kidx, ro, rt, i, K - 1) / info->rho;
  for (int j = 0; j < i; ++j) v *= dc_gap(delta, kidx, ro, rt, i, j) / (delta[kidx[j]] - di);
  for (int j = i; j < K - 1; ++j)
    v *= dc_gap(delta, kidx, ro, rt, i, j) / (delta[kidx[j + 1]] - di);
  zh[i] = copysign(sqrt(fmax(v, 0.0)), zz[kidx[i]]);
}

// Column c of U (rank-one eigenvectors in the sorted basis; unit for a
// deflated c) and its eigenvalue, one wavefront per column.
__global__ __launch_bounds__(64) void k_dc_u(int n, const double* __restrict__ delta,
                                             const int* __restrict__ kidx,
                                             const int* __restrict__ kpos,
                                             const DCInfo* __restrict__ info,
                                             const int* __restrict__ ro,
                                             const double* __restrict__ rt,
                                             const double* __restrict__ zh,
                                             double* __restrict__ U, double* __restrict__ lam) {
  const int c = blockIdx.x, lane = threadIdx.x;
  const int j = kpos[c];
  const double sgn = info->sgn;
  if (j < 0) {
    for (int r = lane; r < n; r += 64) U[(long long)r * n + c] = r == c ? 1.0 : 0.0;
    if (lane == 0) lam[c] = sgn * delta[c];
    return;
  }
  const int K = info->K;
  double nrm = 0;
  for (int i = lane; i < K; i += 64) {
    const double u = zh[i] / dc_gap(delta, kidx, ro, rt, i, j);
    nrm += u * u;
  }
  nrm = wave_sum(nrm);
  const double inv = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
  for (int r = lane; r < n; r += 64) U[(long long)r * n + c] = 0.0;
  for (int i = lane; i < K; i += 64)
    U[(long long)kidx[i] * n + c] = -zh[i] / dc_gap(delta, kidx, ro, rt, i, j) * inv;
  if (lane == 0) lam[c] = sgn * (delta[kidx[ro[j]]] + rt[j]);
}

// Row r of Q in the sorted basis with the deflation rotations applied.
__global__ void k_dc_qp(const double* __restrict__ Z, int n, const int* __restrict__ perm,
                        const DCGivens* __restrict__ giv, const DCInfo* __restrict__ info,
                        double* __restrict__ Qp) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double* q = Qp + (long long)r * n;
  for (int c = 0; c < n; ++c) q[c] = Z[(long long)r * n + perm[c]];
  const int ng = info->n_giv;
  for (int g = 0; g < ng; ++g) {
    const DCGivens G = giv[g];
    const double qa = q[G.a], qb = q[G.b];
    q[G.a] = G.c * qa - G.s * qb;
    q[G.b] = G.s * qa + G.c * qb;
  }
}

// C = A B (n x n, FP64), 16 x 16 tiles.
__global__ __launch_bounds__(256) void k_dgemm_nn(const double* __restrict__ A,
                                                  const double* __restrict__ B, int n,
                                                  double* __restrict__ C) {
  __shared__ double As[16][17], Bs[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int r = blockIdx.y * 16 + ty, c = blockIdx.x * 16 + tx;
  double acc = 0;
  for (int k0 = 0; k0 < n; k0 += 16) {
    As[ty][tx] = (r < n && k0 + tx < n) ? A[(long long)r * n + k0 + tx] : 0.0;
    Bs[ty][tx] = (k0 + ty < n && c < n) ? B[(long long)(k0 + ty) * n + c] : 0.0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) acc += As[ty][k] * Bs[k][tx];
    __syncthreads();
  }
  if (r < n && c < n) C[(long long)r * n + c] = acc;
}

// V = H(0) H(1) ... H(n-2) Z: one wavefront per column of Z (held in LDS),
// the reflectors applied from the last to the first.
constexpr int kQBW = 4;  // columns (wavefronts) per block
__global__ __launch_bounds__(64 * kQBW) void k_ql_back(int n, const c64d* __restrict__ R,
                                                       const c64d* __restrict__ tau,
                                                       const double* __restrict__ Z,
                                                       c64d* __restrict__ V) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  c64d* z = reinterpret_cast<c64d*>(smem_raw) + (long long)wave * n;
  const int j = blockIdx.x * kQBW + wave;
  if (j >= n) return;
  for (int r = lane; r < n; r += 64) z[r] = c64d{Z[(long long)r * n + j], 0.0};
  for (int i = n - 2; i >= 0; --i) {
    const c64d t = tau[i];
    if (t.re == 0.0 && t.im == 0.0) continue;
    const c64d* vi = R + (long long)i * n;
    double re = 0, im = 0;
    for (int r = i + 1 + lane; r < n; r += 64) {
      const c64d a = vi[r], b = z[r];  // conj(a) b
      re += a.re * b.re + a.im * b.im;
      im += a.re * b.im - a.im * b.re;
    }
    re = wave_sum(re);
    im = wave_sum(im);
    const c64d coef = t * c64d{re, im};
    for (int r = i + 1 + lane; r < n; r += 64) z[r] = z[r] - coef * vi[r];
  }
  for (int r = lane; r < n; r += 64) V[(long long)r * n + j] = z[r];
}

// Log capacity: the QL takes about m^2 rotations on an m-row part in
// m + m/2 sweeps; four times that is a generous bound.
static size_t eigh_max_sweeps(size_t n) { return 8 * n + 64; }
static size_t eigh_max_rots(size_t n) { return 4 * n * n + 1024; }

// Below this order the tridiagonal goes to one QL; from it on, to two halves
// and a merge.
constexpr int kDCMin = 32;

namespace {
struct EighWork {
  c64d* R;
  c64d* tau;
  double *e, *Z, *Qp, *U, *Zt, *delta, *zz, *rt, *zh;
  int *perm, *kidx, *kpos, *ro;
  DCGivens* giv;
  DCInfo* info;
  QLSweep* sw;
  double2* rots;
  size_t max_sw, max_rot;  // per part
  size_t bytes;
};

EighWork eigh_layout(char* base, int S) {
  const size_t n = (size_t)std::max(S, 1);
  EighWork w{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* q = base ? base + off : nullptr;
    off += (bytes + 255) & ~(size_t)255;
    return q;
  };
  w.max_sw = eigh_max_sweeps(n);
  w.max_rot = eigh_max_rots(n);
  w.R = reinterpret_cast<c64d*>(take(sizeof(c64d) * n * n));
  w.tau = reinterpret_cast<c64d*>(take(sizeof(c64d) * n));
  w.e = reinterpret_cast<double*>(take(sizeof(double) * n));
  w.Z = reinterpret_cast<double*>(take(sizeof(double) * n * n));
  w.Qp = reinterpret_cast<double*>(take(sizeof(double) * n * n));
  w.U = reinterpret_cast<double*>(take(sizeof(double) * n * n));
  w.Zt = reinterpret_cast<double*>(take(sizeof(double) * n * n));
  w.delta = reinterpret_cast<double*>(take(sizeof(double) * n));
  w.zz = reinterpret_cast<double*>(take(sizeof(double) * n));
  w.rt = reinterpret_cast<double*>(take(sizeof(double) * n));
  w.zh = reinterpret_cast<double*>(take(sizeof(double) * n));
  w.perm = reinterpret_cast<int*>(take(sizeof(int) * n));
  w.kidx = reinterpret_cast<int*>(take(sizeof(int) * n));
  w.kpos = reinterpret_cast<int*>(take(sizeof(int) * n));
  w.ro = reinterpret_cast<int*>(take(sizeof(int) * n));
  w.giv = reinterpret_cast<DCGivens*>(take(sizeof(DCGivens) * n));
  w.info = reinterpret_cast<DCInfo*>(take(sizeof(DCInfo)));
  w.sw = reinterpret_cast<QLSweep*>(take(sizeof(QLSweep) * w.max_sw * 2));
  w.rots = reinterpret_cast<double2*>(take(sizeof(double2) * w.max_rot * 2));
  w.bytes = off;
  return w;
}
}  // namespace

size_t eigh_work_bytes(int S) { return eigh_layout(nullptr, S).bytes; }

constexpr size_t kEighLdsMax = 160 * 1024 - 1024;

hipError_t hermitian_eigh_f64(c64d* G, c64d* V, int S, void* work, double* evals, int* counts,
                              hipStream_t st) {
  if (S <= 0) return hipSuccess;
  const int n = S;
  const EighWork W = eigh_layout(static_cast<char*>(work), S);
  if (W.max_rot > (size_t)INT32_MAX) return hipErrorInvalidValue;
  hipError_t err;
  // the trailing matrix moves into LDS at the first step where it fits there
  int i_lds = n;
  for (int i = 0; i + 1 < n; ++i)
    if (sizeof(c64d) * ((size_t)n + (size_t)(n - i) * (n - i)) <= kEighLdsMax) {
      i_lds = i;
      break;
    }
  const size_t lds_h =
      sizeof(c64d) * ((size_t)n + (i_lds < n ? (size_t)(n - i_lds) * (n - i_lds) : 0));
  if (lds_h > kEighLdsMax) return hipErrorInvalidValue;
  auto hetrd = [&](auto kern) -> hipError_t {
    hipError_t e2 = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_h);
    if (e2 != hipSuccess) return e2;
    hipLaunchKernelGGL(kern, dim3(1), dim3(kHT), lds_h, st, G, n, i_lds, evals, W.e, W.tau, W.R);
    return hipGetLastError();
  };
  const int nv = (n + 63) / 64;
  if (nv <= 1) err = hetrd(k_hetrd<1>);
  else if (nv <= 2) err = hetrd(k_hetrd<2>);
  else if (nv <= 3) err = hetrd(k_hetrd<3>);
  else if (nv <= 4) err = hetrd(k_hetrd<4>);
  else if (nv <= 8) err = hetrd(k_hetrd<8>);
  else if (nv <= 16) err = hetrd(k_hetrd<16>);
  else return hipErrorInvalidValue;  // more than 1024 source points
  if (err != hipSuccess) return err;
  // QL on the whole tridiagonal, or on its two halves (one block each)
  const int h = n >= kDCMin ? n / 2 : 0;
  const int parts = h > 0 ? 2 : 1;
  const int len = h > 0 ? n - h : n;  // the longer part
  // both parts' counts start at zero (one part leaves the second's unwritten);
  // plan set-up is never graph-captured
  if ((err = hipMemsetAsync(counts, 0, 8 * sizeof(int), st)) != hipSuccess) return err;
  hipLaunchKernelGGL(k_tql, dim3(parts), dim3(64), sizeof(double) * (2 * (size_t)len + 3), st,
                     evals, W.e, n, h, W.sw, W.rots, (int)W.max_sw, (int)W.max_rot, counts);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  // two rows per lane (128 per block) where they fit in LDS, else one
  const bool two = sizeof(double) * 128 * (size_t)(len + 1) <= kEighLdsMax;
  const int rpb = two ? 128 : 64;
  const size_t lds_r = sizeof(double) * (size_t)rpb * (len + 1);
  if (lds_r > kEighLdsMax) return hipErrorInvalidValue;
  const void* kr = two ? reinterpret_cast<const void*>(k_ql_rows<2>)
                       : reinterpret_cast<const void*>(k_ql_rows<1>);
  if ((err = hipFuncSetAttribute(kr, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_r)) !=
      hipSuccess)
    return err;
  const dim3 gr((len + rpb - 1) / rpb, parts);
  if (two)
    hipLaunchKernelGGL(k_ql_rows<2>, gr, dim3(64), lds_r, st, n, h, (int)W.max_sw,
                       (int)W.max_rot, W.sw, W.rots, counts, W.Z, 0);
  else
    hipLaunchKernelGGL(k_ql_rows<1>, gr, dim3(64), lds_r, st, n, h, (int)W.max_sw,
                       (int)W.max_rot, W.sw, W.rots, counts, W.Z, 0);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  const double* Zt = W.Z;
  if (h > 0) {
    hipLaunchKernelGGL(k_dc_prep, dim3(1), dim3(kDCT), 0, st, evals, W.Z, W.e, n, h, W.delta,
                       W.zz, W.perm, W.kidx, W.kpos, W.giv, W.info);
    hipLaunchKernelGGL(k_dc_roots, dim3(n), dim3(64), 0, st, W.delta, W.zz, W.kidx, W.info, W.ro,
                       W.rt);
    hipLaunchKernelGGL(k_dc_zhat, dim3((n + 255) / 256), dim3(256), 0, st, W.delta, W.zz, W.kidx,
                       W.info, W.ro, W.rt, W.zh);
    hipLaunchKernelGGL(k_dc_u, dim3(n), dim3(64), 0, st, n, W.delta, W.kidx, W.kpos, W.info,
                       W.ro, W.rt, W.zh, W.U, evals);
    hipLaunchKernelGGL(k_dc_qp, dim3((n + 255) / 256), dim3(256), 0, st, W.Z, n, W.perm, W.giv,
                       W.info, W.Qp);
    hipLaunchKernelGGL(k_dgemm_nn, dim3((n + 15) / 16, (n + 15) / 16), dim3(256), 0, st, W.Qp,
                       W.U, n, W.Zt);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    Zt = W.Zt;
  }
  const size_t lds_b = sizeof(c64d) * kQBW * (size_t)n;
  if (lds_b > kEighLdsMax) return hipErrorInvalidValue;
  if ((err = hipFuncSetAttribute(reinterpret_cast<const void*>(k_ql_back),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_b)) !=
      hipSuccess)
    return err;
  hipLaunchKernelGGL(k_ql_back, dim3((n + kQBW - 1) / kQBW), dim3(64 * kQBW), lds_b, st, n, W.R,
                     W.tau, Zt, V);
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
