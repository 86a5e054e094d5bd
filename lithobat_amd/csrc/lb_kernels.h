// Host-side launchers of the device kernels (internal to the library).
#pragma once
#include "lb_common.h"

namespace lb {

// ---- roots-of-unity tables (lb_util.hip) -----------------------------------
// e^{+2 pi i t / L} for t < L on the current device.  ensure_roots builds a
// table once (it allocates: call it at plan creation, not while capturing);
// roots_table only looks one up (nullptr when absent).
template <typename T>
hipError_t ensure_roots(int L);
template <typename T>
const cx<T>* roots_table(int L);

// the table for L, built on first use
template <typename T>
inline hipError_t table_for(int L, const cx<T>** tw) {
  *tw = roots_table<T>(L);
  if (*tw) return hipSuccess;
  hipError_t e = ensure_roots<T>(L);
  if (e != hipSuccess) return e;
  *tw = roots_table<T>(L);
  return *tw ? hipSuccess : hipErrorInvalidValue;
}

// ---- band transforms (lb_band.hip) ---------------------------------------
// band_size_supported: the sides with radix plans; band_analysis / band_synthesis
// take the direct-sum path of lb_band_any.hip for any other side up to
// kBandAnyMax and for bands wider than the grid.
bool band_size_supported(int n);
constexpr int kBandAnyMax = 4096;
bool band_any_supported(int n);
template <typename T>
hipError_t band_analysis_any(int N, const void* in, int cplx, long long img_stride,
                             long long slice_stride, int nsum, int nimg, int nu0, int Wb, T s,
                             cx<T>* RT, cx<T>* X, hipStream_t st);
template <typename T>
hipError_t band_synthesis_any(int N, const cx<T>* Y, int nimg, int nu0, int Wb, int herm, T s,
                              int cout, cx<T>* TT, T* out, T* peak, hipStream_t st);

// Row stride of the band x N intermediates (RT / TT, [img][y][v]): the band
// width rounded up to 16 complex values, so rows start on 128-byte lines
// (FP32) and the row-major copies move whole lines.
__host__ __device__ inline int band_ld(int Wb) { return Wb < 16 ? Wb : (Wb + 15) / 16 * 16; }

// X[img][u][v] = s * fft2(f_img)[nu0+u][nu0+v]; f_img = sum over nsum slices
// (f real, or complex interleaved when cplx).  RT: scratch img x N x band_ld(Wb).
template <typename T>
hipError_t band_analysis(int N, const void* in, int cplx, long long img_stride,
                         long long slice_stride, int nsum, int nimg, int nu0, int Wb, T s,
                         cx<T>* RT, cx<T>* X, hipStream_t st);

// band_analysis of the ILT mask sigmoid(clip(beta theta, +-60)) of real theta
// [img][N][N] (contiguous images), the sigmoid applied as the row pass loads theta;
// hipErrorNotSupported where the side has no persistent row pass.
template <typename T>
hipError_t band_analysis_sigmoid(int N, const T* theta, int nimg, int nu0, int Wb, T s, T beta,
                                 cx<T>* RT, cx<T>* X, hipStream_t st);

// out[img] = s * Re sum X e^{+...} (cout = 0; X Hermitianised when herm) or the
// complex sum (cout = 1, out holds complex).  TT: scratch img x N x band_ld(Wb).
// peak (nullable, real output only): running max per image, must start at 0.
template <typename T>
hipError_t band_synthesis(int N, const cx<T>* Y, int nimg, int nu0, int Wb, int herm, T s,
                          int cout, cx<T>* TT, T* out, T* peak, hipStream_t st);

// ---- coherent-system kernels (lb_field.hip) -------------------------------
// A work item is a range of coherent systems that all image one mask into one
// image.  Item i covers image i / n_chunks, systems
// [plane * sys_per_plane + chunk * chunk_sz, ...) of plane (image % n_planes),
// mask image / n_planes.
struct ItemMap {
  int n_chunks;
  int n_planes;
  int sys_per_plane;
  int chunk_sz;
  int n_items;
  // spread: the n_img * sys_per_plane systems of the call, in image order,
  // split into `blocks` equal runs, one per resident block; item k * blocks
  // + b is the k-th image segment of block b's run (empty past its end)
  int spread;
  int blocks;
  long long total;
};

// Block whose spread run [b total / blocks, (b+1) total / blocks) holds system q.
__host__ __device__ inline int spread_run_of(long long q, long long total, int blocks) {
  return (int)(((q + 1) * blocks + total - 1) / total) - 1;
}

struct FieldArgs {
  const void* X;      // band spectra [n_masks][Wb][Wb] complex
  int Wb;
  int wrapN;          // > 0: band spans the whole grid, index mod wrapN
  const void* H;      // transfer tables [n_sys][W][W] complex
  int W;
  const int2* box;    // [n_sys] band index (row, col) of each box origin
  const void* wsys;   // [n_sys] real weights
  ItemMap map;
  int* queue;         // work-item counter, zero at launch (null: static item order)
};

struct FieldConfig {
  int M;       // coarse grid side
  int nt;      // threads per block
  int csplit;  // column splits; 0: staged (row tables through device memory)
  int grid;    // blocks (persistent)
};

#include <vector>
// every compiled coherent-system kernel shape, default-preference order
const std::vector<FieldConfig>& field_configs();
bool field_config_supported(int precision, const FieldConfig& c);

// One work item of a multi-plan (uneven clip) launch; *_off are byte offsets
// into the workspace (-1: unused), the rest are plan-owned device tables.
struct ItemDesc {
  long long X_off;
  int Wb, wrapN;
  const void* H;
  int W;
  const int2* box;
  const void* wsys;
  int s0, s1;
  long long part_off, gM_off, gX_off;
};
// items: device array; Wmax sizes LDS; forward writes partials, adjoint gX partials
hipError_t field_multi(int precision, const FieldConfig& c, const ItemDesc* items, int n_items,
                       int Wmax, char* ws, int* queue, bool adjoint, hipStream_t st);
size_t field_smem_bytes(int precision, const FieldConfig& c, int W, bool adjoint);

// partial[item][M][M] = sum over the item's systems of w |E|^2 on the coarse grid
hipError_t field_forward(int precision, const FieldConfig& c, const FieldArgs& a, void* partial,
                         hipStream_t st);
// The staged forward (FieldConfig csplit 0): images [img0, img0 + n_img) of the
// item map, their row tables (n_img * sys_per_plane of staged_table_bytes
// each) in G; writes the same partials as field_forward.
size_t staged_table_bytes(int precision, int M, int W);
hipError_t field_forward_staged(int precision, const FieldConfig& c, const FieldArgs& a,
                                void* partial, void* G, int img0, int n_img, int cus,
                                hipStream_t st);
// The staged adjoint: band-gradient partials gXp[item * nrc + chunk][Wb][Wb],
// nrc = staged_row_chunks(c, W) row chunks per item (the transposed row pass
// of each chunk accumulates into its own partial).
int staged_row_chunks(const FieldConfig& c, int W);
hipError_t field_adjoint_staged(int precision, const FieldConfig& c, const FieldArgs& a,
                                const void* gM, void* gXp, void* G, int img0, int n_img, int cus,
                                hipStream_t st);
// gX[item][Wb][Wb] = sum over the item's systems of 2 w conj(H) FFT(gM E) on the box
hipError_t field_adjoint(int precision, const FieldConfig& c, const FieldArgs& a,
                         const void* gM, void* gX, hipStream_t st);

// ---- small helpers (lb_util.hip) ------------------------------------------
template <typename T>
hipError_t fill(T* p, T v, long long n, hipStream_t st);
// out[g][e] = sum_{k < cnt} in[(g*cnt + k)][e], e < len (deterministic order).
// scratch (nullable) holds groups * kSegmentSplits * len values for the split path.
constexpr int kSegmentSplits = 16;
int segment_splits(int groups, int cnt, long long len);  // scratch slices the split path uses
// Sums the spread-layout partials (ItemMap::spread) of each group of `ipg`
// consecutive images: out[g] = sum over the segments of images g*ipg ..
// (g+1)*ipg - 1, in block order (deterministic).
template <typename T>
hipError_t spread_sum(const T* in, int groups, int ipg, int spp, long long total, int blocks,
                      int segs, long long len, T* out, T* peak, hipStream_t st);
template <typename T>
hipError_t segment_sum(const T* in, int groups, int cnt, long long len, T* out, T* peak,
                       T* scratch, hipStream_t st);
// The ILT parameter behind a mask gradient (opc/ilt.py:381-389 without the
// Gaussian): d/dtheta = d/dm (beta m (1 - m) g), m = sigmoid(clip(beta theta,
// +-60)), g read from device memory (null: 1).
template <typename T>
struct ThetaEpi {
  const T* theta;
  T beta;
  const T* g;
};
// band_synthesis (herm = 1, real output) of a mask gradient written as
// d/dtheta (ThetaEpi) in the same pass; hipErrorNotSupported, nothing
// launched, unless the band is symmetric about 0 inside |nu| < N / 16 and
// theta and grad_theta are 16-byte aligned.  TT as band_synthesis.
template <typename T>
hipError_t band_synthesis_theta(int N, const cx<T>* Y, int nimg, int nu0, int Wb, T s, cx<T>* TT,
                                ThetaEpi<T> te, T* grad_theta, hipStream_t st);
// Per-pixel relaxed-resist objective (opc/ilt.py:311-330): v = sign (scale I -
// level), z = sigmoid(a v), cost (z - t)^2, dI = dscale (z - t) z (1 - z) with
// dscale = sign 2 a scale / npix.
template <typename T>
struct ResistPix {
  T scale, level, a, sign, dscale;
};
// The deterministic reduction of resist_objective's block partials (part_cost,
// part_err [n_img][nblk]) into per-image costs (img_cost: n_img doubles of
// scratch) and the optional outputs.
template <typename T>
hipError_t resist_final(const double* part_cost, const int* part_err, int n_img, int nblk,
                        long long npix, double* img_cost, T* cost, T* total, int* mismatch,
                        hipStream_t st);
// Real-output band synthesis (band_synthesis with herm = 1, s = 1) fused with
// the relaxed-resist objective and the band analysis (nu0, Wb, s = 1) of the
// intensity derivative: GY[img][u][v] as band_analysis of dI would give it,
// per-image costs (cost), their sum (total) and mismatch counts; image and dI
// (nullable) receive the intensity and its derivative.  TT and RT are
// scratch of band_ld((Wb + 1) / 2) * N complex per image; part_cost holds
// synth_objective_part_count(N, nimg) doubles and part_err as many ints.
// hipErrorNotSupported, nothing launched, unless the band is symmetric about
// 0 inside |nu| < N / 16 and target, image and dI are 16-byte aligned.
int synth_objective_part_count(int N, int nimg);
bool synth_objective_supported(int N, int nu0, int Wb);  // the band conditions above
template <typename T>
hipError_t synth_objective(int N, const cx<T>* Y, int nimg, int nu0, int Wb, cx<T>* TT,
                           const T* target, ResistPix<T> rp, T* image, T* dI, cx<T>* RT,
                           cx<T>* GY, double* part_cost, int* part_err, T* cost, T* total,
                           int* mismatch, hipStream_t st);
// The coarse partials of each image (spread layout when `spread`, else `cnt`
// consecutive partials per image) summed and taken to their band spectrum on
// the M x M grid in one pass: Y[img][u][v] = s * sum_{y,x} I[y][x]
// e^{-2 pi i ((u - Q) y + (v - Q) x) / M}, u, v < 2Q + 1 (the result of
// spread_sum / segment_sum followed by band_analysis with nu0 = -Q).
// peak (nullable): set to 0 per image, ready for the synthesis's running maximum.
// hipErrorNotSupported, with nothing launched, when M, Q or the LDS rule it out.
template <typename T>
hipError_t coarse_band(int M, const T* parts, int spread, int cnt, int spp, long long total,
                       int blocks, int segs, int nimg, int Q, T s, cx<T>* Y, T* peak,
                       hipStream_t st);

template <typename T>
hipError_t scatter_band(const cx<T>* in, int K, int Wb, int nu0, int N, cx<T>* out,
                        hipStream_t st);
template <typename T>
hipError_t band_mul(const cx<T>* X, const cx<T>* Hk, int n_masks, int K, long long len, cx<T>* Y,
                    hipStream_t st);

// Per-system dense spectra on the whole grid (band order, nu0 = -N/2) for the
// dense comparison form, and the weighted |E|^2 accumulation over them.
template <typename T>
hipError_t dense_spectra(const cx<T>* X, int N, const cx<T>* H, int W, const int2* box, int j0,
                         int nsys, int nu0m, cx<T>* spec, hipStream_t st);
template <typename T>
hipError_t accum_norm2(const cx<T>* E, int nsys, long long len, const T* w, int j0, T* img,
                       hipStream_t st);
// E_k *= g (real) for a batch of fields; A += w conj(H) G over each box (atomic)
template <typename T>
hipError_t mul_real(cx<T>* E, int nsys, long long len, const T* g, hipStream_t st);
// Y_k = g E_k (copy); A (+)= sum_k w conj(H) X_k over each system's box on a band
// (deterministic order): the adjoint from given fields.
template <typename T>
hipError_t mul_real_into(const cx<T>* E, int nsys, long long len, const T* g, cx<T>* Y,
                         hipStream_t st);
template <typename T>
hipError_t fields_adj_band(const cx<T>* X, int Wb, int wrapN, const cx<T>* H, int W,
                           const int2* box, const T* w, int j0, int nsys, int accumulate,
                           cx<T>* A, hipStream_t st);
template <typename T>
hipError_t dense_adjoint_accum(const cx<T>* G, int N, const cx<T>* H, int W, const int2* box,
                               const T* w, int j0, int nsys, int nu0m, cx<T>* A, hipStream_t st);

// ---- Gaussian blur (lb_blur.hip) ---------------------------------------------
// scipy.ndimage.gaussian_filter(mode="reflect", truncate=4.0) of n_img n x n
// planes; tmp holds n_img n^2 values, w 2R+1 doubles (R = gauss_radius(sigma)).
int gauss_radius(double sigma);
template <typename T>
hipError_t gaussian_blur(const T* in, T* out, int n_img, int n, double sigma, T* tmp, double* w,
                         hipStream_t st);

// ---- ILT objective (lb_objective.hip) ------------------------------------------
// Relaxed-resist cost, its derivative, the signed field and the pattern
// mismatch of n_img images of npix pixels (d_f, signed_f, cost, total and
// mismatch may be null); part_cost / part_err hold n_img * resist_blocks(npix)
// partials.  Mask sigmoid m = sigmoid(clip(beta x, +-60)) and its adjoint.
int resist_blocks(long long npix);
template <typename T>
hipError_t resist_objective(const T* f, const T* target, int n_img, long long npix, T scale,
                            T level, T a, T sign, T* cost, T* total, T* d_f, T* signed_f,
                            int* mismatch, double* part_cost, int* part_err, hipStream_t st);
template <typename T>
hipError_t mask_sigmoid(const T* x, long long n, T beta, const double* pbeta, T* m,
                        hipStream_t st);
template <typename T>
hipError_t mask_sigmoid_grad(const T* gm, const T* m, long long n, T beta, const double* pbeta,
                             T* gx, hipStream_t st);
// gx = gscale[0] * gm * beta m (1 - m), m = sigmoid(clip(beta x, +-60)) from x (gscale null: 1)
template <typename T>
hipError_t mask_sigmoid_grad_theta(const T* gm, const T* x, long long n, T beta,
                                   const double* pbeta, const T* gscale, T* gx, hipStream_t st);
template <typename T>
hipError_t mask_penalties(const T* m, int n_img, int ny, int nx, T wtv, T wd, T* cost, T* total,
                          T* grad, double* part_cost, int* part_zero, hipStream_t st);
template <typename T>
hipError_t adam_step(T* theta, T* m1, T* m2, const T* g, long long n, T b1, T b2, T o1, T o2,
                     T eps, T step, T c1, T c2, const double* pc, hipStream_t st);
template <typename T>
hipError_t volume_lerp(const T* raw, int n_masks, int n_planes, long long npix, int nz,
                       const int* jz, const T* tz, const T* scale, T* vol, hipStream_t st);
// s = [beta, it, 1 - b1^it, 1 - b2^it, beta_next] advanced by one iteration
hipError_t ilt_schedule(double* s, double growth, double b1, double b2, hipStream_t st);

// ---- pupil tables (lb_pupil.hip) ------------------------------------------
struct PupilSys {
  double fsx, fsy;       // source-point frequency offset [1/m]
  double weight;         // multiplies H (1 for Abbe systems, sqrt(w_s) for the TCC rows)
  double jx, jy;         // input Jones vector (vector model)
  double defocus;        // [m]
  double n_image;
  int comp;              // -1 scalar, 0..2 = x, y, z component
  int box_y, box_x;      // frequency index (box coordinates) of the box origin
  int pad;
  double sin_x, sin_y;   // reflective mask: mask-side sines of the illumination
};

constexpr int kMaxZernike = 21;
struct PupilParams {
  int N;
  int W;
  double val;            // 1 / (N * pixel_size), the fftfreq step
  double k_max;          // NA / wavelength
  double wavelength;
  double NA;
  int exact;             // exact (non-paraxial) defocus
  int obliquity;
  int n_zern;
  int zn[kMaxZernike], zm[kMaxZernike];
  double znorm[kMaxZernike], zcoef[kMaxZernike], zrad[kMaxZernike][3];
  // reflective mask (n_refl > 0): blank reflection table on the device
  int n_refl;
  double refl_doff;
  const double* refl_theta;
  const double* refl_mag;
  const double* refl_phase;
};

// H[j][a][b] (complex T) and dc[j] = weight^2-free DC power of system j.
template <typename T>
hipError_t pupil_tables(const PupilSys* sys_dev, int n_sys, const PupilParams& pp, cx<T>* H,
                        double* dc, hipStream_t st);

// ---- SOCS decomposition (lb_socs.hip), FP64 ------------------------------
// G[s][t] = sum_f conj(A[s][f]) A[t][f]; work: gram_work_bytes(S, F) of scratch
size_t gram_work_bytes(int S, long long F);
hipError_t gram_f64(const cx<double>* A, int S, long long F, cx<double>* G, void* work,
                    hipStream_t st);
// Hermitian eigensolve (Householder tridiagonalisation; implicit QL, on two
// halves merged by divide and conquer from 32 rows on): G destroyed, evals (S,
// unordered), V = eigenvectors (columns); work: eigh_work_bytes(S) of scratch;
// counts (8 ints, 4 per QL part) = {QL sweeps, rotations, status (0 ok, 1 no
// convergence, 2 rotation log full), unused}
size_t eigh_work_bytes(int S);
hipError_t hermitian_eigh_f64(cx<double>* G, cx<double>* V, int S, void* work, double* evals,
                              int* counts, hipStream_t st);
// phi[k][f] = sum_s A[s][f] V[s][idx[k]] * inv_sqrt[k]
hipError_t socs_phi_f64(const cx<double>* A, const cx<double>* V, int S, long long F,
                        const int* idx, const double* inv_sqrt, int K, cx<double>* phi,
                        hipStream_t st);
hipError_t cast_c128_c64(const cx<double>* in, cx<float>* out, long long n, hipStream_t st);

}  // namespace lb
