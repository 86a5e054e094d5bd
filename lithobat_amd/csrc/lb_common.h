// Device building blocks shared by every kernel of the imaging path:
// a complex type, in-register radix-2/3/4/5/8/10/12/16 DFTs, and one Stockham
// autosort stage over LDS (Govindaraju et al., "High performance discrete
// Fourier transforms on graphics processors", SC 2008).
//
// Sign convention: DIR = +1 is the inverse transform (e^{+2 pi i n k / L},
// unnormalised), DIR = -1 the forward one (e^{-2 pi i n k / L}).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace lb {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T r, T i) { return cx<T>{r, i}; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cx<T> scale(cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T>
__device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.re, -a.im}; }
// a * conj(b)
template <typename T>
__device__ __forceinline__ cx<T> mulc(cx<T> a, cx<T> b) {
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <typename T>
__device__ __forceinline__ T norm2(cx<T> a) { return a.re * a.re + a.im * a.im; }

// The ILT mask and relaxed-resist sigmoid, 1 / (1 + e^{-v}) with v clipped to +-60
// (opc/ilt.py:311-378 clips the same way).
template <typename T>
__device__ __forceinline__ T sigmoid_clipped(T v) {
  v = v < (T)-60 ? (T)-60 : (v > (T)60 ? (T)60 : v);
  return (T)1 / ((T)1 + exp(-v));
}
// Single precision from the hardware base-2 exponential and reciprocal (about
// 1 ulp each): e^{-v} = 2^{-v log2 e}, |v log2 e| <= 86.6 stays inside the
// exponent range.  The rounding of -v log2 e puts a relative error of up to
// about |v| 2^-24 on e^{-v}; modelled in double on the CPU over |v| <= 60 the
// result is within 1e-7 of the exact sigmoid (absolute), within 1e-6 relative
// for |v| <= 16, and within 1e-13 wherever the sigmoid is below 1e-7.
template <>
__device__ __forceinline__ float sigmoid_clipped<float>(float v) {
  v = v < -60.f ? -60.f : (v > 60.f ? 60.f : v);
  const float e = __builtin_amdgcn_exp2f(v * -1.44269504088896341f);
  return __builtin_amdgcn_rcpf(1.f + e);
}

// Wait until every vector-memory access of this wave has completed
// (s_waitcnt vmcnt(0), other counters unconstrained).  Placed after the
// prologue prefetch of a persistent loop: the compiler's wait analysis merges
// the loop-entry state (prefetch loads still in flight) with the back edge
// (the previous iteration's stores in flight) and would otherwise wait for
// all of those stores at the loop head of every iteration.
// The scheduling barriers keep the compiler from moving the prefetch loads
// past the wait.
__device__ __forceinline__ void wait_vmem_all() {
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __builtin_amdgcn_sched_barrier(0);
}
// multiply by (DIR * i)
template <int DIR, typename T>
__device__ __forceinline__ cx<T> mul_i(cx<T> a) {
  return DIR > 0 ? cx<T>{-a.im, a.re} : cx<T>{a.im, -a.re};
}

// cos / sin (2 pi k / 16), k = 0..4; the rest follow by symmetry.
#define LB_C16_1 0.92387953251128674
#define LB_S16_1 0.38268343236508978
#define LB_R2_2 0.70710678118654752

// e^{DIR * 2 pi i k / R} for compile-time R | 16 and k < R.
template <int R, int DIR, typename T>
__device__ __forceinline__ cx<T> root(int k) {
  const int q = (k * (16 / R)) & 15;  // in units of 2 pi / 16
  T c, s;
  switch (q) {
    case 0: c = 1; s = 0; break;
    case 1: c = (T)LB_C16_1; s = (T)LB_S16_1; break;
    case 2: c = (T)LB_R2_2; s = (T)LB_R2_2; break;
    case 3: c = (T)LB_S16_1; s = (T)LB_C16_1; break;
    case 4: c = 0; s = 1; break;
    case 5: c = -(T)LB_S16_1; s = (T)LB_C16_1; break;
    case 6: c = -(T)LB_R2_2; s = (T)LB_R2_2; break;
    case 7: c = -(T)LB_C16_1; s = (T)LB_S16_1; break;
    case 8: c = -1; s = 0; break;
    case 9: c = -(T)LB_C16_1; s = -(T)LB_S16_1; break;
    case 10: c = -(T)LB_R2_2; s = -(T)LB_R2_2; break;
    case 11: c = -(T)LB_S16_1; s = -(T)LB_C16_1; break;
    case 12: c = 0; s = -1; break;
    case 13: c = (T)LB_S16_1; s = -(T)LB_C16_1; break;
    case 14: c = (T)LB_R2_2; s = -(T)LB_R2_2; break;
    default: c = (T)LB_C16_1; s = -(T)LB_S16_1; break;
  }
  return cx<T>{c, DIR > 0 ? s : -s};
}

// Multiply by a compile-time root, skipping the trivial cases.
template <int R, int DIR, typename T>
__device__ __forceinline__ cx<T> mul_root(cx<T> a, int k) {
  const int q = (k * (16 / R)) & 15;
  if (q == 0) return a;
  if (q == 4) return mul_i<DIR>(a);
  if (q == 8) return cx<T>{-a.re, -a.im};
  if (q == 12) return mul_i<-DIR>(a);
  return a * root<R, DIR, T>(k);
}

// In-register DFT of size R (power of two <= 16), natural order in and out.
// Radix-2 decimation in time, fully unrolled at compile time.
template <int R, int DIR, typename T>
struct DFT2k {
  __device__ __forceinline__ static void run(cx<T>* v) {
    cx<T> e[R / 2], o[R / 2];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    DFT2k<R / 2, DIR, T>::run(e);
    DFT2k<R / 2, DIR, T>::run(o);
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      cx<T> t = mul_root<R, DIR, T>(o[k], k);
      v[k] = e[k] + t;
      v[k + R / 2] = e[k] - t;
    }
  }
};
template <int DIR, typename T>
struct DFT2k<1, DIR, T> {
  __device__ __forceinline__ static void run(cx<T>*) {}
};
template <int DIR, typename T>
struct DFT2k<2, DIR, T> {
  __device__ __forceinline__ static void run(cx<T>* v) {
    cx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
};
template <int DIR, typename T>
struct DFT2k<4, DIR, T> {
  __device__ __forceinline__ static void run(cx<T>* v) {
    cx<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
    cx<T> t2 = v[1] + v[3], t3 = mul_i<DIR>(v[1] - v[3]);
    v[0] = t0 + t2;
    v[2] = t0 - t2;
    v[1] = t1 + t3;
    v[3] = t1 - t3;
  }
};

// e^{DIR * 2 pi i k / R} for compile-time R | 12 (multiples of 30 degrees).
template <int R, int DIR, typename T>
__device__ __forceinline__ cx<T> root12(int k) {
  const int q = (k * (12 / R)) % 12;
  const T h = (T)0.86602540378443865;  // sqrt(3)/2
  T c, s;
  switch (q) {
    case 0: c = 1; s = 0; break;
    case 1: c = h; s = (T)0.5; break;
    case 2: c = (T)0.5; s = h; break;
    case 3: c = 0; s = 1; break;
    case 4: c = (T)-0.5; s = h; break;
    case 5: c = -h; s = (T)0.5; break;
    case 6: c = -1; s = 0; break;
    case 7: c = -h; s = (T)-0.5; break;
    case 8: c = (T)-0.5; s = -h; break;
    case 9: c = 0; s = -1; break;
    case 10: c = (T)0.5; s = -h; break;
    default: c = h; s = (T)-0.5; break;
  }
  return cx<T>{c, DIR > 0 ? s : -s};
}

// e^{DIR * 2 pi i k / R} for compile-time R | 10 (multiples of 36 degrees).
template <int R, int DIR, typename T>
__device__ __forceinline__ cx<T> root10(int k) {
  const int q = (k * (10 / R)) % 10;
  const T c1 = (T)0.80901699437494742, s1 = (T)0.58778525229247313;  // 36 deg
  const T c2 = (T)0.30901699437494742, s2 = (T)0.95105651629515357;  // 72 deg
  T c, s;
  switch (q) {
    case 0: c = 1; s = 0; break;
    case 1: c = c1; s = s1; break;
    case 2: c = c2; s = s2; break;
    case 3: c = -c2; s = s2; break;
    case 4: c = -c1; s = s1; break;
    case 5: c = -1; s = 0; break;
    case 6: c = -c1; s = -s1; break;
    case 7: c = -c2; s = -s2; break;
    case 8: c = c2; s = -s2; break;
    default: c = c1; s = -s1; break;
  }
  return cx<T>{c, DIR > 0 ? s : -s};
}

// Twiddle of a composite in-register DFT of size R (R | 12 or R | 10).
template <int R, int DIR, typename T>
__device__ __forceinline__ cx<T> root_of(int k) {
  if constexpr (12 % R == 0)
    return root12<R, DIR, T>(k);
  else
    return root10<R, DIR, T>(k);
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void dft(cx<T>* v);

// Composite R = A * B in registers (Cooley-Tukey): input n = B n1 + n2,
// output k = k1 + A k2; A-point DFTs, twiddles e^{DIR 2 pi i n2 k1 / R},
// then B-point DFTs.  Used for R = 6, 12 (3 * 2^k) and R = 10.
template <int A, int B, int DIR, typename T>
__device__ __forceinline__ void dft_mixed(cx<T>* v) {
  constexpr int R = A * B;
  cx<T> y[B][A];
#pragma unroll
  for (int n2 = 0; n2 < B; ++n2) {
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) y[n2][n1] = v[B * n1 + n2];
    dft<A, DIR>(y[n2]);
#pragma unroll
    for (int k1 = 1; k1 < A; ++k1)
      if (n2 > 0) y[n2][k1] = y[n2][k1] * root_of<R, DIR, T>(n2 * k1);
  }
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) {
    cx<T> z[B];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) z[n2] = y[n2][k1];
    dft<B, DIR>(z);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = z[k2];
  }
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void dft(cx<T>* v) {
  if constexpr (R == 6) {
    dft_mixed<2, 3, DIR, T>(v);
  } else if constexpr (R == 12) {
    dft_mixed<4, 3, DIR, T>(v);
  } else if constexpr (R == 10) {
    dft_mixed<2, 5, DIR, T>(v);
  } else if constexpr (R == 5) {
    // X_k = sum_n v_n w^{nk}, w = e^{DIR 2 pi i / 5}; pairs (1,4), (2,3)
    const T c1 = (T)0.30901699437494742, s1 = (T)0.95105651629515357;   // 72 deg
    const T c2 = (T)-0.80901699437494742, s2 = (T)0.58778525229247313;  // 144 deg
    const cx<T> p14 = v[1] + v[4], m14 = v[1] - v[4];
    const cx<T> p23 = v[2] + v[3], m23 = v[2] - v[3];
    const cx<T> a1 = v[0] + scale(p14, c1) + scale(p23, c2);
    const cx<T> a2 = v[0] + scale(p14, c2) + scale(p23, c1);
    const cx<T> b1 = mul_i<DIR>(scale(m14, s1) + scale(m23, s2));
    const cx<T> b2 = mul_i<DIR>(scale(m14, s2) - scale(m23, s1));
    v[0] = v[0] + p14 + p23;
    v[1] = a1 + b1;
    v[4] = a1 - b1;
    v[2] = a2 + b2;
    v[3] = a2 - b2;
  } else if constexpr (R == 3) {
    const T h = (T)0.86602540378443865;  // sqrt(3)/2
    cx<T> s = v[1] + v[2];
    cx<T> d = mul_i<DIR>(v[1] - v[2]);
    cx<T> m = v[0] - scale(s, (T)0.5);
    v[0] = v[0] + s;
    v[1] = m + scale(d, h);
    v[2] = m - scale(d, h);
  } else {
    DFT2k<R, DIR, T>::run(v);
  }
}

// e^{DIR 2 pi i t / L} from a table of the L-th roots of unity (tw[t] holds
// the e^{+2 pi i t / L} values).
template <int DIR, typename T>
__device__ __forceinline__ cx<T> tw_get(const cx<T>* tw, int t) {
  cx<T> w = tw[t];
  return DIR > 0 ? w : conj(w);
}

// Fill tw[0..L) with e^{+2 pi i t / L}, computed in double.
template <typename T>
__device__ __forceinline__ void fill_roots(cx<T>* tw, int L, int tid, int nt) {
  for (int t = tid; t < L; t += nt) {
    double s, c;
    sincospi(2.0 * (double)t / (double)L, &s, &c);
    tw[t] = cx<T>{(T)c, (T)s};
  }
}

// One Stockham stage of radix R with span Ns on a length-L transform held
// contiguously at in / out (out of place).  Task j in [0, L/R).
template <int R, int DIR, typename T>
__device__ __forceinline__ void stockham_stage(const cx<T>* in, cx<T>* out, int L, int Ns, int j,
                                               const cx<T>* tw) {
  cx<T> v[R];
  const int stride = L / R;
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = in[j + r * stride];
  const int k = j % Ns;
  if (Ns > 1) {
    const int step = L / (Ns * R);
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = v[r] * tw_get<DIR>(tw, r * k * step);
  }
  dft<R, DIR>(v);
  const int base = (j / Ns) * Ns * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r) out[base + r * Ns] = v[r];
}

// A full length-L transform of `ntr` independent rows laid out at
// buf + row * ld (row < ntr), by all threads of the block, radices R0..R3
// (1 = unused, product must equal L).  Ping-pongs between a and b; returns
// the buffer that holds the result.  Synchronises the block between stages
// and at the end.
template <typename T, int DIR, int L, int R0, int R1, int R2, int R3>
__device__ cx<T>* fft_rows_lds(cx<T>* a, cx<T>* b, int ntr, int ld, const cx<T>* tw, int tid,
                               int nt) {
  static_assert(R0 * R1 * R2 * R3 == L, "radix plan must multiply to L");
  cx<T>* src = a;
  cx<T>* dst = b;
  int Ns = 1;
  auto stage = [&](auto rc) {
    constexpr int R = decltype(rc)::value;
    if constexpr (R > 1) {
      const int tasks = L / R;
      for (int t = tid; t < ntr * tasks; t += nt) {
        const int row = t / tasks, j = t - row * tasks;
        stockham_stage<R, DIR, T>(src + row * ld, dst + row * ld, L, Ns, j, tw);
      }
      __syncthreads();
      cx<T>* x = src;
      src = dst;
      dst = x;
      Ns *= R;
    }
  };
  stage(std::integral_constant<int, R0>{});
  stage(std::integral_constant<int, R1>{});
  stage(std::integral_constant<int, R2>{});
  stage(std::integral_constant<int, R3>{});
  return src;
}

// The stages R_a, R_b, R_c (1 = unused) of a length-L Stockham transform whose
// earlier stages (span Ns0 so far) are already done and held in `a`.
// Same conventions as fft_rows_lds.
template <typename T, int DIR, int L, int Ra, int Rb, int Rc>
__device__ cx<T>* fft_tail_lds(cx<T>* a, cx<T>* b, int ntr, int ld, const cx<T>* tw, int tid,
                               int nt, int Ns0) {
  cx<T>* src = a;
  cx<T>* dst = b;
  int Ns = Ns0;
  auto stage = [&](auto rc) {
    constexpr int R = decltype(rc)::value;
    if constexpr (R > 1) {
      const int tasks = L / R;
      for (int t = tid; t < ntr * tasks; t += nt) {
        const int row = t / tasks, j = t - row * tasks;
        stockham_stage<R, DIR, T>(src + row * ld, dst + row * ld, L, Ns, j, tw);
      }
      __syncthreads();
      cx<T>* x = src;
      src = dst;
      dst = x;
      Ns *= R;
    }
  };
  stage(std::integral_constant<int, Ra>{});
  stage(std::integral_constant<int, Rb>{});
  stage(std::integral_constant<int, Rc>{});
  return src;
}

// Copy a precomputed roots table (e^{+2 pi i t / L}, t < L) into LDS.
template <typename T>
__device__ __forceinline__ void load_roots(cx<T>* tw, const cx<T>* __restrict__ twg, int L,
                                           int tid, int nt) {
  for (int t = tid; t < L; t += nt) tw[t] = twg[t];
}

// Canonical frequency of FFT bin b (numpy fftfreq order).
__device__ __forceinline__ int nu_of_bin(int b, int n) { return b <= (n - 1) / 2 ? b : b - n; }

// Frequency index nu (any integer) -> FFT bin in [0, n).
__device__ __forceinline__ int bin_of(int nu, int n) {
  int b = nu % n;
  return b < 0 ? b + n : b;
}

}  // namespace lb
