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
