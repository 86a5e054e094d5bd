/*
 * lithobat.h -- C ABI of the MI355X (gfx950) partially coherent imaging
 * library: the Abbe source-point sum, its sum-of-coherent-systems (SOCS /
 * Hopkins) form, and the adjoint of both for inverse lithography.
 *
 * Plain C types, device pointers and sizes only.  All device buffers are
 * owned by the caller except plans, which the library allocates and frees.
 * Every call returns an lb_status; on failure lb_last_error() (thread-local)
 * describes it.  Calls that take a hipStream_t (passed as void*) are
 * asynchronous on that stream.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * reference checkout, src/litho_sim/):
 *   lb_abbe_plan_create   expose/aerial_image.py:275-304 (source_points) and
 *                         :392-419 (per-call geometry set-up of _abbe_sum)
 *   lb_forward            expose/aerial_image.py:351-494 (_abbe_sum) behind
 *                         compute_aerial_image :502-548 and
 *                         compute_aerial_planes :551-587; for SOCS plans
 *                         expose/hopkins.py:205-218 (SOCSKernels.raw) and
 *                         opc/ilt.py:208-232 (socs_fields, intensity part)
 *   lb_backward           opc/ilt.py:235-266 (socs_gradient) -- the adjoint
 *                         of lb_forward, recomputing the fields
 *   lb_socs_plan_create   expose/hopkins.py:103-198 (SOCSKernels.build)
 *   lb_socs_kernels       the SOCSKernels.kernels attribute (hopkins.py:87)
 *   lb_fields             opc/ilt.py:208-232 (socs_fields, the fields part)
 *   lb_plan_clear         the `clear` scalar of _abbe_sum / SOCSKernels
 */
#ifndef LITHOBAT_H_
#define LITHOBAT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  LB_OK = 0,
  LB_ERR_ARG = 1,          /* bad argument, size or shape */
  LB_ERR_UNSUPPORTED = 2,  /* valid request this build does not implement */
  LB_ERR_HIP = 3,          /* HIP runtime error */
  LB_ERR_OOM = 4,          /* device allocation failed */
  LB_ERR_WORKSPACE = 5     /* workspace too small */
} lb_status;

typedef enum {
  LB_POL_UNPOLARISED = 0,
  LB_POL_X = 1,
  LB_POL_Y = 2,
  LB_POL_TE = 3,
  LB_POL_TM = 4
} lb_polarisation;

typedef struct lb_plan_s* lb_plan;
typedef struct lb_batch_s* lb_batch;

/* Optics, grid and source of one imaging set-up (OpticsConfig + GridConfig +
 * SourcePoints of the reference, core/config.py:383-406, 741-744;
 * expose/aerial_image.py:249-272). */
typedef struct {
  int32_t n;               /* grid side N (pixels) */
  double pixel_size;       /* [m] */
  double wavelength;       /* [m] */
  double na;
  int32_t vector;          /* 0 scalar, 1 vector (three field components) */
  int32_t polarisation;    /* lb_polarisation, vector model only */
  int32_t obliquity;       /* sqrt(cos theta) factor, vector model only */
  int32_t exact_defocus;   /* non-paraxial defocus OPD (forced on for vector) */
  int32_t n_planes;        /* image planes imaged together */
  const double* defocus;   /* [n_planes] defocus [m] */
  const double* n_image;   /* [n_planes] image-medium index */
  int32_t n_zernike;       /* Noll terms, j in 1..21 (others contribute 0) */
  const int32_t* zernike_j;
  const double* zernike_c; /* [n_zernike] coefficients in waves */
  int32_t n_source;        /* source points */
  const double* src_weight;
  const double* src_fx;    /* [1/m] */
  const double* src_fy;    /* [1/m] */
  const double* src_phi;   /* azimuth [rad] */
  int32_t precision;       /* 32: complex64 / FP32 accumulate; 64: complex128 */
  int32_t coarse_m;        /* 0 = automatic coarse-grid side, else forced */
  int32_t chunk;           /* 0 = automatic systems per work item */
  /* Reflective-mask model (MultilayerSpectra, expose/m3d/provider.py:190-293):
   * source point s images the spectrum M(f) r(|sin_s + refl_doff f|), r the
   * blank reflection interpolated linearly in angle (asin of the sine) from
   * the table, 0 beyond sine 1, and its clear field is scaled by |r|^2 at the
   * specular order.  n_refl = 0: thin mask (Abbe plans only). */
  int32_t n_refl;           /* angle-table points (0 = thin mask) */
  const double* refl_theta; /* [n_refl] angles [rad], increasing */
  const double* refl_mag;   /* [n_refl] |r| */
  const double* refl_phase; /* [n_refl] unwrapped arg r [rad] */
  double refl_doff;         /* wavelength / reduction: sine per unit frequency */
  const double* src_sin_x;  /* [n_source] mask-side direction sines of each */
  const double* src_sin_y;  /*   source point's illumination */
} lb_imaging_desc;

typedef struct {
  int32_t n;               /* N */
  int32_t m;               /* coarse grid side (== n when imaged directly) */
  int32_t box;             /* W: pupil box side in frequency bins */
  int32_t band;            /* side of the mask-spectrum band */
  int32_t n_planes;
  int32_t n_systems;       /* coherent systems per plane */
  int32_t precision;
  int32_t n_kernels;       /* SOCS plans: kernels kept (else 0) */
  int32_t block_threads;   /* chosen coherent-system kernel shape */
  int32_t column_splits;
  int32_t chunk;           /* systems per work item */
  int32_t grid_blocks;     /* persistent blocks of the coherent-system kernel */
  int32_t adj_block_threads; /* the adjoint kernel's shape */
  int32_t adj_column_splits;
  int32_t adj_chunk;       /* systems per adjoint work item for one mask */
  int32_t socs_sweeps;     /* SOCS plans: QL sweeps of the eigensolve (else 0) */
} lb_plan_info;

/* Library version (major * 10000 + minor * 100 + patch). */
int32_t lb_version(void);
/* Last error message of this thread ("" when none). */
const char* lb_last_error(void);

/* Abbe plan: builds the source table, pupil boxes and the per-system transfer
 * tables (FP64 on the device) for every plane, on the current device. */
lb_status lb_abbe_plan_create(const lb_imaging_desc* desc, lb_plan* out);

/* SOCS plan (scalar model, one plane): TCC eigen-decomposition in FP64 on the
 * device, kernels kept where lambda > tol * lambda_max and lambda > 0, capped at
 * max_kernels (< 0: no cap).  The kernels are used in desc->precision. */
lb_status lb_socs_plan_create(const lb_imaging_desc* desc, double tol, int32_t max_kernels,
                              lb_plan* out);

/* _build_fft_pupil (expose/aerial_image.py:169-209): the scalar pupil P(f - f_s) =
 * exp(i W) on the in-pupil bins of the whole N x N grid, zeros elsewhere, in FFT order
 * (DC at [0][0]), complex128, for source offset (fsx, fsy) [1/m] and the defocus / image
 * index of `plane`; W is the defocus OPD plus the Zernike terms, as the plans use.  Any
 * N >= 1.  Allocates and synchronises (a set-up call, not for graph capture). */
lb_status lb_pupil_dense(const lb_imaging_desc* desc, double fsx, double fsy, int32_t plane,
                         void* out_c128, void* stream);

lb_status lb_plan_destroy(lb_plan plan);
lb_status lb_plan_info_get(lb_plan plan, lb_plan_info* info);

/* Open-frame throughput per plane (host array of n_planes). */
lb_status lb_plan_clear(lb_plan plan, double* clear_per_plane);

/* SOCS plans: eigenvalues of the kept kernels (host, n_kernels), the
 * trace-normalised energy kept, and the band size n_band. */
lb_status lb_socs_weights(lb_plan plan, double* weights, double* energy_kept,
                          int32_t* n_band, int32_t* n_source);
/* SOCS plans: kernels on the full grid, FFT order, complex128, [K][N][N]
 * (device, caller allocated). */
lb_status lb_socs_kernels(lb_plan plan, void* kernels_c128, void* stream);

/* Workspace bytes lb_forward / lb_backward need for n_masks masks. */
size_t lb_workspace_bytes(lb_plan plan, int32_t n_masks, int32_t backward);

/* Raw (un-normalised, un-dosed) images of n_masks masks on every plane:
 * images[mask][plane][N][N] (real, precision of the plan) and the per-image
 * peak (may be NULL).  masks: [n_masks][N][N] real, or interleaved complex
 * when mask_complex != 0, in the plan's precision. */
lb_status lb_forward(lb_plan plan, const void* masks, int32_t mask_complex, int32_t n_masks,
                     void* images, void* peak, void* workspace, size_t workspace_bytes,
                     void* stream);

/* No counterpart in the reference: lb_forward for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The workspace is a fixed device array of the library,
 * LB_STATIC_SCRATCH_BYTES on each GPU it runs on (held until the process
 * ends), instead of a caller buffer; the call allocates nothing.  Same kernels
 * and outputs as lb_forward.  When lb_forward's workspace for the call would
 * not fit, the work items take more coherent systems each (fewer coarse
 * partial images; the images agree with lb_forward's to rounding), and staged
 * plans take smaller passes; a call that does not fit even then returns
 * LB_ERR_WORKSPACE.  Every call on one GPU shares that GPU's array: calls that
 * overlap in time on one GPU (other streams or threads) race. */
#define LB_STATIC_SCRATCH_BYTES ((size_t)64 << 20)
lb_status lb_forward_static(lb_plan plan, const void* masks, int32_t mask_complex,
                            int32_t n_masks, void* images, void* peak, void* stream);

/* d_masks[mask][N][N] = d/d mask of sum_{plane,x} d_images[mask][plane][x] * I(x)
 * for real masks (fields recomputed, not stored). */
lb_status lb_backward(lb_plan plan, const void* masks, int32_t n_masks, const void* d_images,
                      void* d_masks, void* workspace, size_t workspace_bytes, void* stream);

/* The same forward and adjoint from the masks' band spectra (the
 * ThinMaskSpectra of expose/aerial_image.py's provider, restricted to the
 * band every system's box lies in): lb_mask_spectrum computes them
 * (spectrum: [n_masks][Wb][Wb] complex, lb_mask_spectrum_bytes), and
 * lb_forward_spectrum / lb_backward_spectrum image from them and run the
 * adjoint from them, so an autograd step analyses each mask once (the
 * reference recomputes the fields too, opc/ilt.py:235-266).  Workspaces as
 * lb_forward / lb_backward (lb_mask_spectrum: the forward's). */
size_t lb_mask_spectrum_bytes(lb_plan plan, int32_t n_masks);
lb_status lb_mask_spectrum(lb_plan plan, const void* masks, int32_t mask_complex,
                           int32_t n_masks, void* spectrum, void* workspace,
                           size_t workspace_bytes, void* stream);
lb_status lb_forward_spectrum(lb_plan plan, const void* spectrum, int32_t n_masks, void* images,
                              void* peak, void* workspace, size_t workspace_bytes, void* stream);
lb_status lb_backward_spectrum(lb_plan plan, const void* spectrum, int32_t n_masks,
                               const void* d_images, void* d_masks, void* workspace,
                               size_t workspace_bytes, void* stream);

/* lb_mask_spectrum of the ILT mask m = sigmoid(clip(beta * theta, -60, 60))
 * (opc/ilt.py:368-378, _mask_of without its Gaussian) from the real parameter
 * theta [n_masks][N][N]: the sigmoid is applied as the band analysis loads
 * theta, so no mask array is written.  LB_ERR_UNSUPPORTED for a grid side
 * without the persistent analysis row pass (then lb_mask_sigmoid followed by
 * lb_mask_spectrum gives the same spectra).  Workspace as lb_mask_spectrum. */
lb_status lb_mask_spectrum_sigmoid(lb_plan plan, const void* theta, int32_t n_masks, double beta,
                                   void* spectrum, void* workspace, size_t workspace_bytes,
                                   void* stream);

/* lb_backward_spectrum for complex masks: d_masks[mask][N][N] (interleaved complex) =
 * dJ/dRe m + i dJ/dIm m of J = sum d_images * I, the gradient torch's autograd uses for a
 * complex input; its real part is the reference's projection (opc/ilt.py:235-266 takes
 * 2 Re of the adjoint). */
lb_status lb_backward_spectrum_complex(lb_plan plan, const void* spectrum, int32_t n_masks,
                                       const void* d_images, void* d_masks, void* workspace,
                                       size_t workspace_bytes, void* stream);

/* Coherent fields of each system on the full grid, complex, [n_masks][n_sys][N][N]
 * for plane 0 (SOCS: one per kernel, the `fields` of socs_fields). */
lb_status lb_fields(lb_plan plan, const void* masks, int32_t mask_complex, int32_t n_masks,
                    void* fields, void* workspace, size_t workspace_bytes, void* stream);
size_t lb_fields_workspace_bytes(lb_plan plan, int32_t n_masks);

/* socs_gradient(kernels, fields, sensitivity) of opc/ilt.py:235-266 on the fields it is
 * given (SOCS plans): d_masks[mask][N][N] = 2 Re sum_k lambda_k ifft2(conj(phi_k)
 * fft2(sensitivity[mask] * fields[mask][k])) for fields [n_masks][K][N][N] (complex) and
 * sensitivity [n_masks][N][N] (real), in the plan's precision; complex_out != 0 writes the
 * complex adjoint 2 sum_k ... before its real part (interleaved complex d_masks).  The
 * fields need not come from lb_fields.  Workspace: lb_fields_backward_workspace_bytes. */
size_t lb_fields_backward_workspace_bytes(lb_plan plan);
lb_status lb_fields_backward(lb_plan plan, const void* fields, int32_t n_masks,
                             const void* sensitivity, void* d_masks, int32_t complex_out,
                             void* workspace, size_t workspace_bytes, void* stream);

/* The reference's own dataflow on the device, for comparison: the raw image
 * of one real mask on one plane as one dense N x N inverse transform per
 * coherent system, `batch` systems per transform launch, |E|^2 accumulated
 * with the source weights (expose/aerial_image.py:421-488 with _Batch
 * :312-348).  Same result as lb_forward; no coarse grid. */
size_t lb_dense_workspace_bytes(lb_plan plan, int32_t batch);
lb_status lb_forward_dense(lb_plan plan, const void* mask, int32_t plane, int32_t batch,
                           void* image, void* workspace, size_t workspace_bytes, void* stream);
/* The matching dense adjoint (Abbe plans, real masks): d_mask = d/dmask of
 * sum(d_image * I) for one mask and plane, per coherent system as the
 * reference's adjoint runs it (opc/ilt.py:235-266: fft2 of the sensitivity
 * times the field, conj(pupil), one inverse transform at the end).  Same
 * result as lb_backward; same workspace size as lb_forward_dense. */
lb_status lb_backward_dense(lb_plan plan, const void* mask, int32_t plane, int32_t batch,
                            const void* d_image, void* d_mask, void* workspace,
                            size_t workspace_bytes, void* stream);

/* Gaussian blur of n_img n x n planes (FP32 or FP64, in and out may alias):
 * the PEB / ILT blur of opc/ilt.py:274-289 (scipy.ndimage.gaussian_filter with
 * mode "reflect", truncate 4.0), separable, weights computed on the device.
 * sigma < 1e-3 copies the input, as the reference returns it unblurred.  It is
 * its own adjoint (symmetric taps, symmetric extension), which is how the
 * reference's chain rule (ilt.py:336, 381-389) uses it. */
size_t lb_gaussian_blur_workspace_bytes(int32_t n_img, int32_t n, double sigma, int32_t precision);
lb_status lb_gaussian_blur(const void* in, void* out, int32_t n_img, int32_t n, double sigma,
                           int32_t precision, void* workspace, size_t workspace_bytes,
                           void* stream);

/* The relaxed-resist objective of opc/ilt.py:311-337 (_Objective.__call__,
 * without the TV / discreteness penalties) over n_img images of npix pixels
 * each, in one pass: f = scale * intensity (pass the PEB-blurred field with
 * scale 1 when sigma_px > 0), signed = sign * (f - level),
 * z = 1 / (1 + exp(-clip(a * signed, -60, 60))), a = steepness / max(|level|,
 * 1e-12); cost[b] = sum (z - target)^2 / npix and total = sum_b cost[b];
 * d_intensity = sign * (2 / npix) * (z - target) * a * z (1 - z) * scale (the
 * reference's d_intensity before its blur); mismatch[b] =
 * count_nonzero((signed > 0) != (target > 0.5)) (ilt.py:357).  intensity and
 * target are [n_img][npix] in the given precision; cost and total are
 * device scalars of that precision; every output may be NULL.  Costs are
 * reduced in a fixed order (deterministic). */
size_t lb_resist_objective_workspace_bytes(int32_t n_img, int64_t npix);
lb_status lb_resist_objective(const void* intensity, const void* target, int32_t n_img,
                              int64_t npix, double scale, double level, double steepness,
                              double sign, void* cost, void* total, void* d_intensity,
                              void* signed_field, int32_t* mismatch, int32_t precision,
                              void* workspace, size_t workspace_bytes, void* stream);

/* The mask penalties of opc/ilt.py:339-355 (the TV and discreteness terms of
 * _Objective.__call__) over n_img masks [ny][nx], periodic neighbours:
 *   cost_b = weight_tv sum (gx^2 + gy^2) / npix + weight_discrete sum m (1 - m) / npix,
 *   grad   = weight_tv (-2 / npix) lap(m) + weight_discrete (1 - 2 m) / npix,
 * one pass plus a deterministic two-level reduction.  cost [n_img], total
 * (0-d) and grad (like mask) are optional.  Workspace: the size
 * lb_resist_objective_workspace_bytes(n_img, ny * nx) returns. */
lb_status lb_mask_penalties(const void* mask, int32_t n_img, int32_t ny, int32_t nx,
                            double weight_tv, double weight_discrete, void* cost, void* total,
                            void* grad, int32_t precision, void* workspace, size_t ws_bytes,
                            void* stream);

/* One Adam step of opc/ilt.py:587-591 in place over n values:
 *   mom1 = b1 mom1 + (1 - b1) g,  mom2 = b2 mom2 + (1 - b2) g^2,
 *   theta -= step (mom1 / (1 - b1^it)) / (sqrt(mom2 / (1 - b2^it)) + eps),
 * iteration = it >= 1 (the bias corrections are formed in float64). */
lb_status lb_adam_step(void* theta, void* mom1, void* mom2, const void* grad, int64_t n,
                       int32_t iteration, double step, double beta1, double beta2, double eps,
                       int32_t precision, void* stream);

/* The voxel interpolation of develop/resist3d.py:172-185 (exposure_volume):
 * for n_masks stacks of n_planes raw images [npix] and nz voxels (bottom-up),
 * voxel z mixes planes plane_index[z] and plane_index[z] + 1 (each scaled by
 * its image normalisation plane_scale[mask][plane]) with the weight
 * weight[z] of the second: a + w (c - a).  plane_index (int32), weight and
 * plane_scale are device arrays; volume is [n_masks][nz][npix].  The caller
 * guarantees plane_index[z] + 1 < n_planes. */
lb_status lb_volume_interp(const void* raw, int32_t n_masks, int32_t n_planes, int64_t npix,
                           int32_t nz, const int32_t* plane_index, const void* weight,
                           const void* plane_scale, void* volume, int32_t precision,
                           void* stream);

/* Device-scalar forms for solver loops captured into a graph (no counterpart
 * in the reference, whose loop runs on the host: opc/ilt.py:583-592).
 * lb_ilt_schedule advances schedule = [beta, it, 1 - beta1^it, 1 - beta2^it,
 * beta_next] (float64, device) by one iteration: it += 1, beta = beta_next,
 * beta_next = beta * growth, and the two bias corrections.  The _dev entry
 * points read beta (lb_mask_sigmoid_dev / _grad_dev) or the two corrections
 * (lb_adam_step_dev: corrections[0..1]) from device memory at run time. */
lb_status lb_ilt_schedule(double* schedule, double growth, double beta1, double beta2,
                          void* stream);
lb_status lb_adam_step_dev(void* theta, void* mom1, void* mom2, const void* grad, int64_t n,
                           const double* corrections, double step, double beta1, double beta2,
                           double eps, int32_t precision, void* stream);
lb_status lb_mask_sigmoid_dev(const void* x, int64_t n, const double* beta, void* mask,
                              int32_t precision, void* stream);
lb_status lb_mask_sigmoid_grad_dev(const void* grad_mask, const void* mask, int64_t n,
                                   const double* beta, void* grad_x, int32_t precision,
                                   void* stream);

/* The ILT mask parameterisation (opc/ilt.py:368-389, _mask_of / _grad_theta
 * without their Gaussian, which lb_gaussian_blur applies): mask =
 * 1 / (1 + exp(-clip(beta * x, -60, 60))) and grad_x = grad_mask * beta *
 * mask * (1 - mask), over n values. */
lb_status lb_mask_sigmoid(const void* x, int64_t n, double beta, void* mask, int32_t precision,
                          void* stream);
lb_status lb_mask_sigmoid_grad(const void* grad_mask, const void* mask, int64_t n, double beta,
                               void* grad_x, int32_t precision, void* stream);
/* opc/ilt.py:381-389 (_grad_theta without its Gaussian) from theta rather than
 * the mask: grad_x = g * grad_mask * beta * m * (1 - m) with m recomputed as
 * lb_mask_sigmoid does, g a scalar read from device memory (the incoming
 * gradient of a scalar loss, of the same precision; null: 1), so neither the
 * mask nor a loss-scaled intensity derivative needs a pass of its own. */
lb_status lb_mask_sigmoid_grad_theta(const void* grad_mask, const void* theta, int64_t n,
                                     double beta, const void* g, void* grad_x,
                                     int32_t precision, void* stream);

/* The relaxed-resist objective of the images a spectrum gives (opc/ilt.py:
 * 311-330, _Objective's image term; lb_forward_spectrum then
 * lb_resist_objective), evaluated where the images are synthesised: per image
 * the cost, their total and the mismatch counts exactly as lb_resist_objective
 * returns them (each output nullable), and sens_band [n_masks * n_planes][Wb]
 * [Wb] complex (lb_sensitivity_band_bytes; Wb = 2 box - 1) = the band
 * spectrum of the intensity derivative d_intensity, from which
 * lb_backward_sensitivity_band runs the adjoint (opc/ilt.py:235-266) as
 * lb_backward_spectrum does from d_intensity.  target [n_img][N][N] real;
 * images and d_images (nullable) also receive the intensity and its
 * derivative.  Neither array needs a pass of its own.  LB_ERR_UNSUPPORTED for
 * a plan imaged directly on its N x N grid, an image band reaching |nu| >=
 * N / 16, or a pointer that is not 16-byte aligned (then lb_forward_spectrum,
 * lb_resist_objective and lb_backward_spectrum give the same results).
 * lb_forward_objective_workspace_bytes returns 0 when unsupported. */
size_t lb_forward_objective_workspace_bytes(lb_plan plan, int32_t n_masks);
size_t lb_sensitivity_band_bytes(lb_plan plan, int32_t n_masks);
lb_status lb_forward_objective(lb_plan plan, const void* spectrum, int32_t n_masks,
                               const void* target, double scale, double level, double steepness,
                               double sign, void* cost, void* total, int32_t* mismatch,
                               void* sens_band, void* images, void* d_images, void* workspace,
                               size_t workspace_bytes, void* stream);
lb_status lb_backward_sensitivity_band(lb_plan plan, const void* spectrum, int32_t n_masks,
                                       const void* sens_band, void* d_masks, void* workspace,
                                       size_t workspace_bytes, void* stream);
/* lb_backward_sensitivity_band followed by lb_mask_sigmoid_grad_theta in one
 * pass: the gradient of the ILT parameter theta [n_masks][N][N] (the mask is
 * sigmoid(clip(beta theta, +-60)), opc/ilt.py:368-389 without the Gaussian),
 * grad_theta = d/dm (beta m (1 - m) g), g a scalar in device memory (null: 1),
 * written by the mask-gradient synthesis itself.  LB_ERR_UNSUPPORTED for a
 * plan imaged directly, a mask band reaching |nu| >= N / 16, or theta /
 * grad_theta not 16-byte aligned (then the two calls give the same result).
 * Workspace as lb_backward_spectrum. */
lb_status lb_backward_objective_theta(lb_plan plan, const void* spectrum, int32_t n_masks,
                                      const void* sens_band, const void* theta, double beta,
                                      const void* g, void* grad_theta, void* workspace,
                                      size_t workspace_bytes, void* stream);

/* No counterpart in the reference: the coherent-system kernel shape of a
 * batch's forward or adjoint launch (single-launch shapes only; the work
 * items and workspace do not depend on it). */
lb_status lb_batch_set_shape(lb_batch batch, int32_t adjoint, int32_t block_threads,
                             int32_t column_splits);

/* lb_batch_backward after an lb_batch_forward of the same masks: the masks'
 * band spectra are taken from that call's workspace (forward_workspace,
 * untouched since) instead of recomputed, so the masks are not passed. */
lb_status lb_batch_backward_reuse(lb_batch batch, const void* const* d_images,
                                  void* const* d_masks, void* workspace,
                                  size_t workspace_bytes, const void* forward_workspace,
                                  void* stream);

/* Batches of clips (no direct counterpart in the reference, which images one
 * mask per call of compute_aerial_image / runs one run_ilt per clip): plans of
 * different grid sizes and optics -- one mask each -- imaged together.  The
 * band transforms run per clip; all clips' coherent systems run in one
 * persistent launch.  Plans must share precision and coarse grid side (create
 * them with the same coarse_m) and must outlive the batch.  Arrays are host
 * arrays of n_plans device pointers; peaks may be NULL. */
lb_status lb_batch_create(const lb_plan* plans, int32_t n_plans, lb_batch* out);
lb_status lb_batch_destroy(lb_batch batch);
size_t lb_batch_workspace_bytes(lb_batch batch, int32_t backward);
lb_status lb_batch_forward(lb_batch batch, const void* const* masks, void* const* images,
                           void* const* peaks, void* workspace, size_t workspace_bytes,
                           void* stream);
lb_status lb_batch_backward(lb_batch batch, const void* const* masks,
                            const void* const* d_images, void* const* d_masks, void* workspace,
                            size_t workspace_bytes, void* stream);

/* Kernel-shape control for the autotuner (no counterpart in the reference):
 * list the coherent-system kernel shapes (threads, column splits) that fit
 * this plan, and select one for the forward (adjoint = 0) or adjoint
 * (adjoint = 1) kernel; chunk >= 0 also sets the systems per work item
 * (0 = automatic), and chunk = LB_CHUNK_SPREAD splits the call's systems,
 * in image order, into one equal run per resident block of the kernel (at
 * most a few image segments per block; shapes with column splits only).
 * block_threads = 0 restores the plan's own default shape for that kernel
 * (and automatic work items unless chunk > 0 or LB_CHUNK_SPREAD). */
#define LB_CHUNK_SPREAD (-2)
lb_status lb_plan_shapes(lb_plan plan, int32_t adjoint, int32_t* shapes_pairs,
                         int32_t max_shapes, int32_t* n_shapes);
lb_status lb_plan_set_shape(lb_plan plan, int32_t adjoint, int32_t block_threads,
                            int32_t column_splits, int32_t chunk);
/* The work items a call of n_masks masks runs: systems per item and item count. */
lb_status lb_plan_work_items(lb_plan plan, int32_t n_masks, int32_t adjoint, int32_t* chunk,
                             int32_t* n_items);

/* Timing aid (no counterpart in the reference): when enabled, the forward and
 * adjoint entry points record a HIP event pair on their stream around each
 * launch (or staged pass) of the coherent-system kernel; lb_profile_read
 * (forward launches) and lb_profile_read_adjoint wait for them and return the
 * launch count and summed kernel time.  Enabling resets the record. */
lb_status lb_profile_enable(lb_plan plan, int32_t on);
lb_status lb_profile_read(lb_plan plan, int32_t* launches, double* total_ms);
lb_status lb_profile_read_adjoint(lb_plan plan, int32_t* launches, double* total_ms);

#ifdef __cplusplus
}
#endif

#endif /* LITHOBAT_H_ */
