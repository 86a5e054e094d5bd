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

constexpr int kJT = 1024;
constexpr int kMaxPairs = 2048;

__global__ __launch_bounds__(kJT) void k_jacobi(c64d* __restrict__ G, c64d* __restrict__ V, int S,
                                                int max_sweeps, double* __restrict__ evals,
                                                int* __restrict__ sweeps_out) {
  __shared__ double rc[kMaxPairs / 2], rs[kMaxPairs / 2];
  __shared__ c64d re_[kMaxPairs / 2];
  __shared__ int rp[kMaxPairs / 2], rq[kMaxPairs / 2];
  __shared__ double red[kJT / 64];
  __shared__ int done;
  const int tid = threadIdx.x;
  const int n = S + (S & 1);  // even number of players; index S is a dummy when S is odd
  const int half = n / 2;
  for (long long e = tid; e < (long long)S * S; e += kJT) {
    const int r = (int)(e / S), c = (int)(e % S);
    V[e] = c64d{r == c ? 1.0 : 0.0, 0.0};
  }
  __syncthreads();
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    // convergence test: off-diagonal vs total Frobenius mass
    double off = 0, tot = 0;
    for (long long e = tid; e < (long long)S * S; e += kJT) {
      const int r = (int)(e / S), c = (int)(e % S);
      const double m = G[e].re * G[e].re + G[e].im * G[e].im;
      tot += m;
      if (r != c) off += m;
    }
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_xor(off, o, 64);
      tot += __shfl_xor(tot, o, 64);
    }
    if ((tid & 63) == 0) red[tid >> 6] = off;
    __syncthreads();
    if (tid == 0) {
      double so = 0;
      for (int w = 0; w < kJT / 64; ++w) so += red[w];
      red[0] = so;
    }
    __syncthreads();
    const double off_all = red[0];
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = tot;
    __syncthreads();
    if (tid == 0) {
      double st = 0;
      for (int w = 0; w < kJT / 64; ++w) st += red[w];
      done = (off_all <= 1e-30 * st) ? 1 : 0;
    }
    __syncthreads();
    if (done) break;
    for (int round = 0; round < n - 1; ++round) {
      for (int i = tid; i < half; i += kJT) {
        int p = (round + i) % (n - 1);
        int q = i == 0 ? n - 1 : (round - i + (n - 1)) % (n - 1);
        if (p > q) {
          const int x = p;
          p = q;
          q = x;
        }
        double c = 1, s = 0;
        c64d ea = c64d{1, 0};
        if (q < S) {
          const double a = G[(long long)p * S + p].re, d = G[(long long)q * S + q].re;
          const c64d b = G[(long long)p * S + q];
          const double ab = sqrt(b.re * b.re + b.im * b.im);
          if (ab > 1e-300 && ab > 1e-18 * sqrt(fabs(a * d))) {
            const double tau = (d - a) / (2.0 * ab);
            const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
            ea = c64d{b.re / ab, b.im / ab};
          }
        }
        rp[i] = p;
        rq[i] = q;
        rc[i] = c;
        rs[i] = s;
        re_[i] = ea;
      }
      __syncthreads();
      // G <- U^H G U and V <- V U in one pass: the rotation of pair a acts on
      // rows {p_a, q_a} and that of pair b on columns {p_b, q_b}, so every 2 x 2
      // block (a, b) of G is updated on its own (rows first, then columns, the
      // same arithmetic as two separate passes), and so is every row segment
      // (k, b) of V.  Loads of a group of tasks are issued before any store.
      {
        const long long nG = (long long)half * half, nV = (long long)S * half;
        constexpr int U = 2;
        for (long long w0 = tid; w0 < nG + nV; w0 += U * kJT) {
          c64d g[U][4];
          int ra[U], rb[U], ca[U], cb[U];  // rows p_a, q_a; columns p_b, q_b (-1: absent)
          int pa[U], pb[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const long long w = w0 + (long long)u * kJT;
            ra[u] = rb[u] = ca[u] = cb[u] = -1;
            pa[u] = pb[u] = -1;
            if (w < nG) {
              const int a = (int)(w / half), b = (int)(w - (long long)(w / half) * half);
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
            } else if (w < nG + nV) {
              const long long wv = w - nG;
              const int k = (int)(wv / half), b = (int)(wv - (long long)(wv / half) * half);
              pb[u] = b;
              ra[u] = k;
              ca[u] = rp[b];
              cb[u] = rq[b] < S ? rq[b] : -1;
              g[u][0] = V[(long long)k * S + ca[u]];
              g[u][1] = cb[u] >= 0 ? V[(long long)k * S + cb[u]] : c64d{0, 0};
            }
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const long long w = w0 + (long long)u * kJT;
            if (w < nG) {
              const int a = pa[u], b = pb[u];
              if (rb[u] >= 0 && rs[a] != 0.0) {  // rows: G <- U_a^H G
                const double c = rc[a], s_ = rs[a];
                const c64d ea = re_[a];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                  const c64d gp = g[u][j], gq = g[u][2 + j];
                  g[u][j] = scale(gp, c) - scale(ea * gq, s_);
                  g[u][2 + j] = scale(mulc(gp, ea), s_) + scale(gq, c);
                }
              }
              if (cb[u] >= 0 && rs[b] != 0.0) {  // columns: G <- G U_b
                const double c = rc[b], s_ = rs[b];
                const c64d eb = re_[b];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                  const c64d gp = g[u][2 * i], gq = g[u][2 * i + 1];
                  g[u][2 * i] = scale(gp, c) - scale(mulc(gq, eb), s_);
                  g[u][2 * i + 1] = scale(eb * gp, s_) + scale(gq, c);
                }
              }
              G[(long long)ra[u] * S + ca[u]] = g[u][0];
              if (cb[u] >= 0) G[(long long)ra[u] * S + cb[u]] = g[u][1];
              if (rb[u] >= 0) G[(long long)rb[u] * S + ca[u]] = g[u][2];
              if (rb[u] >= 0 && cb[u] >= 0) G[(long long)rb[u] * S + cb[u]] = g[u][3];
            } else if (w < nG + nV) {
              const int b = pb[u];
              if (cb[u] >= 0 && rs[b] != 0.0) {  // V <- V U_b
                const double c = rc[b], s_ = rs[b];
                const c64d eb = re_[b];
                const c64d vp = g[u][0], vq = g[u][1];
                V[(long long)ra[u] * S + ca[u]] = scale(vp, c) - scale(mulc(vq, eb), s_);
                V[(long long)ra[u] * S + cb[u]] = scale(eb * vp, s_) + scale(vq, c);
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

hipError_t jacobi_eigh_f64(c64d* G, c64d* V, int S, int max_sweeps, double* evals,
                           int* sweeps_out, hipStream_t st) {
  if (S > kMaxPairs - 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_jacobi, dim3(1), dim3(kJT), 0, st, G, V, S, max_sweeps, evals, sweeps_out);
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
