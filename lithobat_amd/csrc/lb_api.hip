// C ABI of the imaging library (declared in include/lithobat.h).
//
// A plan is one optical set-up: its coherent systems (Abbe: source point x
// plane x field component; SOCS: eigen-kernel), their transfer tables on the
// device, the box / band geometry and the coherent-system kernel shape.
// lb_forward / lb_backward are pure launch sequences on the caller's stream
// over a caller-provided workspace: no allocation, no host synchronisation,
// so they can be captured into a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/lithobat.h"
#include "lb_kernels.h"

using namespace lb;

namespace {

thread_local std::string g_err;

lb_status fail(lb_status s, const std::string& msg) {
  g_err = msg;
  return s;
}

#define LB_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(e_ == hipErrorOutOfMemory ? LB_ERR_OOM : LB_ERR_HIP,                  \
                  std::string(#call) + ": " + hipGetErrorString(e_));                   \
  } while (0)

constexpr size_t kLdsLimit = 160 * 1024 - 64;  // dynamic LDS; the kernels' static slots take the rest
constexpr size_t kStagedPassBytes = 160u << 20;  // row tables per staged pass (last-level cache 256 MB)
const int kFieldSizes[] = {16, 32, 48, 64, 80, 96, 128, 256};
// Sizes the automatic choice picks from: the 80-point grid (radix 10 x 8)
// runs slower per image than the 96-point one on the same boxes, so it is
// used only when requested.
const int kAutoFieldSizes[] = {16, 32, 48, 64, 96, 128, 256};

// The fixed device workspace of lb_forward_static, for callers that cannot
// allocate.  Each GPU the library runs on holds its own copy for the life of
// the process.  Host code never passes the array's name to a kernel (on the
// host it is only a stand-in for the symbol): set_geometry_common, the plan
// set-up, looks up each device's address of it once with that device current,
// and the entry point passes that address.
static __device__ __align__(256) unsigned char scratch[LB_STATIC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// This device's address of `scratch` (looked up once per device).
lb_status static_scratch(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(LB_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(LB_ERR_UNSUPPORTED, "device index beyond the static scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(LB_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return LB_OK;
}

// Noll j -> (n, m), expose/pupil.py:56-78.
const int kNoll[22][2] = {{0, 0},  {0, 0},  {1, 1},  {1, -1}, {2, 0},  {2, -2}, {2, 2},  {3, -1},
                          {3, 1},  {3, -3}, {3, 3},  {4, 0},  {4, 2},  {4, -2}, {4, 4},  {4, -4},
                          {5, 1},  {5, -1}, {5, 3},  {5, -3}, {5, 5},  {5, -5}};

double factorial(int n) {
  double f = 1;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

struct Plan {
  int precision = 32;
  int N = 0;
  double px = 0, lam = 0, na = 0, k_max = 0, val = 0;
  int vector = 0, pol = 0, obliq = 1, exact = 0;
  std::vector<double> defocus, n_image;
  PupilParams pp{};
  int S = 0;
  // geometry
  int W = 0, M = 0, direct = 0;
  int QI = 0, WbI = 0;
  int nu0m = 0, Wbm = 0, wrapm = 0;
  int n_planes = 1, spp = 0;  // systems per plane
  int chunk_fwd = 0, chunk_adj = 0;  // systems per work item (0 = automatic)
  FieldConfig fwd{}, adj{};
  int cus = 256;
  // device tables
  void* H = nullptr;
  int2* box = nullptr;
  void* wsys = nullptr;
  std::vector<double> clear;
  // SOCS
  int socs = 0;
  int K = 0;
  std::vector<double> weights;
  double energy = 1.0;
  int n_band = 0;
  int eig_sweeps = 0;  // QL sweeps of the SOCS eigensolve
  void* phi128 = nullptr;  // [K][W][W] complex128
  double* refl = nullptr;  // reflective mask: [3][n_refl] angle, |r|, arg r
  // optional event timing of the coherent-system kernel (lb_profile_*)
  int profile = 0;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, int>> ev_pairs;  // (start, stop) indices per forward launch
  std::vector<std::pair<int, int>> ev_pairs_adj;  // the same for the adjoint launches

  ~Plan() {
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    if (H) (void)hipFree(H);
    if (box) (void)hipFree(box);
    if (wsys) (void)hipFree(wsys);
    if (phi128) (void)hipFree(phi128);
    if (refl) (void)hipFree(refl);
  }
  size_t csz() const { return precision == 64 ? 16 : 8; }
  size_t rsz() const { return precision == 64 ? 8 : 4; }
};

bool field_size(int m) {
  for (int s : kFieldSizes)
    if (s == m) return true;
  return false;
}

// Coherent-system kernel shape for (M, W): the first of field_configs() (the
// staged shapes on M = 256, else no column split, then the smallest split)
// whose LDS fits; the adjoint takes single-launch shapes only.
bool default_field_config(const Plan& p, bool adjoint, FieldConfig* out) {
  for (const FieldConfig& c : field_configs()) {
    if (c.M != p.M) continue;
    if (field_smem_bytes(p.precision, c, p.W, adjoint) <= kLdsLimit) {
      *out = c;
      return true;
    }
  }
  return false;
}

bool pick_field_configs(Plan& p) {
  return default_field_config(p, false, &p.fwd) && default_field_config(p, true, &p.adj);
}

lb_status set_geometry_common(Plan& p, const lb_imaging_desc* d) {
  p.precision = d->precision == 64 ? 64 : 32;
  p.N = d->n;
  p.px = d->pixel_size;
  p.lam = d->wavelength;
  p.na = d->na;
  p.k_max = d->na / d->wavelength;  // aerial_image.py:131
  p.val = 1.0 / ((double)d->n * d->pixel_size);  // numpy fftfreq step
  p.vector = d->vector ? 1 : 0;
  p.pol = d->polarisation;
  p.obliq = d->obliquity ? 1 : 0;
  p.exact = (d->exact_defocus || d->vector) ? 1 : 0;  // aerial_image.py:392
  p.chunk_fwd = p.chunk_adj = d->chunk > 0 ? d->chunk : 0;
  PupilParams& pp = p.pp;
  std::memset(&pp, 0, sizeof(pp));
  pp.N = p.N;
  pp.val = p.val;
  pp.k_max = p.k_max;
  pp.wavelength = p.lam;
  pp.NA = p.na;
  pp.exact = p.exact;
  pp.obliquity = p.obliq;
  int nz = 0;
  for (int t = 0; t < d->n_zernike; ++t) {
    const int j = d->zernike_j[t];
    if (j < 1 || j > 21) continue;  // pupil.py:128-130: unknown Noll index contributes 0
    if (nz >= kMaxZernike) return fail(LB_ERR_ARG, "too many Zernike terms");
    const int n = kNoll[j][0], m = kNoll[j][1], ma = std::abs(m);
    pp.zn[nz] = n;
    pp.zm[nz] = m;
    pp.znorm[nz] = m == 0 ? std::sqrt((double)(n + 1)) : std::sqrt(2.0 * (n + 1));
    pp.zcoef[nz] = d->zernike_c[t];
    for (int s = 0; s < 3; ++s) pp.zrad[nz][s] = 0.0;
    for (int s = 0; s <= (n - ma) / 2; ++s) {  // pupil.py:89-99
      pp.zrad[nz][s] = ((s % 2) ? -1.0 : 1.0) * factorial(n - s) /
                       (factorial(s) * factorial((n + ma) / 2 - s) * factorial((n - ma) / 2 - s));
    }
    ++nz;
  }
  pp.n_zern = nz;
  if (d->n_refl > 0) {  // reflective mask: the blank's reflection table on the device
    const int nr = d->n_refl;
    if (nr < 2 || !d->refl_theta || !d->refl_mag || !d->refl_phase || !d->src_sin_x ||
        !d->src_sin_y)
      return fail(LB_ERR_ARG, "reflective mask: need >= 2 table points and the source sines");
    for (int i = 1; i < nr; ++i)
      if (!(d->refl_theta[i] > d->refl_theta[i - 1]))
        return fail(LB_ERR_ARG, "reflective mask: table angles must increase");
    std::vector<double> tab(3 * (size_t)nr);
    std::copy(d->refl_theta, d->refl_theta + nr, tab.begin());
    std::copy(d->refl_mag, d->refl_mag + nr, tab.begin() + nr);
    std::copy(d->refl_phase, d->refl_phase + nr, tab.begin() + 2 * nr);
    LB_HIP(hipMalloc(&p.refl, sizeof(double) * tab.size()));
    LB_HIP(hipMemcpy(p.refl, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
    pp.n_refl = nr;
    pp.refl_doff = d->refl_doff;
    pp.refl_theta = p.refl;
    pp.refl_mag = p.refl + nr;
    pp.refl_phase = p.refl + 2 * nr;
  }
  int dev = 0;
  LB_HIP(hipGetDevice(&dev));
  LB_HIP(hipDeviceGetAttribute(&p.cus, hipDeviceAttributeMultiprocessorCount, dev));
  void* sp = nullptr;
  return static_scratch(&sp);  // this device's lb_forward_static workspace, once per device
}

lb_status check_desc(const lb_imaging_desc* d) {
  if (!d) return fail(LB_ERR_ARG, "null descriptor");
  if (d->n < 1) return fail(LB_ERR_ARG, "grid side must be >= 1");
  if (!band_any_supported(d->n))
    return fail(LB_ERR_UNSUPPORTED, "grid side " + std::to_string(d->n) + " exceeds " +
                                        std::to_string(kBandAnyMax));
  if (!(d->pixel_size > 0) || !(d->wavelength > 0) || !(d->na > 0))
    return fail(LB_ERR_ARG, "pixel_size, wavelength and NA must be > 0");
  if (d->n_planes < 1 || !d->defocus || !d->n_image) return fail(LB_ERR_ARG, "need >= 1 plane");
  if (d->n_source < 0) return fail(LB_ERR_ARG, "negative source count");
  if (d->n_source > 0 && (!d->src_weight || !d->src_fx || !d->src_fy || !d->src_phi))
    return fail(LB_ERR_ARG, "null source arrays");
  if (d->polarisation < 0 || d->polarisation > 4) return fail(LB_ERR_ARG, "bad polarisation");
  if (d->n_zernike > 0 && (!d->zernike_j || !d->zernike_c))
    return fail(LB_ERR_ARG, "null Zernike arrays");
  return LB_OK;
}

// Box side, coarse grid side, image band.
lb_status choose_grid(Plan& p, int Wbox, int forced_m) {
  p.W = std::min(Wbox, p.N);
  int M = 0;
  if (forced_m > 0) {
    if (forced_m != p.N && forced_m < 2 * p.W - 1)
      return fail(LB_ERR_ARG, "coarse grid side must be >= 2W-1 or equal N");
    M = forced_m;
  } else {
    for (int s : kAutoFieldSizes)
      if (s >= 2 * p.W - 1) {
        M = s;
        break;
      }
    // a coarse grid no smaller than the image grid images on the image grid itself where
    // that is a coherent-system size; otherwise it stays coarser than needed and the band
    // synthesis folds its frequencies onto the image grid
    if (M == 0 || (M >= p.N && field_size(p.N))) M = p.N;
  }
  if (!field_size(M))
    return fail(LB_ERR_UNSUPPORTED, "coarse grid side " + std::to_string(M) +
                                        " unsupported (pupil box " + std::to_string(p.W) +
                                        " bins on a " + std::to_string(p.N) + " grid)");
  p.M = M;
  p.direct = (M == p.N) ? 1 : 0;
  p.QI = p.W - 1;
  p.WbI = 2 * p.W - 1;
  if (!p.direct && !band_size_supported(M))
    return fail(LB_ERR_UNSUPPORTED, "coarse grid side not a band size");
  if (!pick_field_configs(p))
    return fail(LB_ERR_UNSUPPORTED, "no coherent-system kernel shape fits LDS for M=" +
                                        std::to_string(M) + " W=" + std::to_string(p.W));
  // roots tables the launches will use (allocated here, never while capturing)
  for (int L : {p.N, p.M}) {
    hipError_t e = p.precision == 64 ? ensure_roots<double>(L) : ensure_roots<float>(L);
    if (e != hipSuccess) return fail(LB_ERR_HIP, std::string("roots table: ") + hipGetErrorString(e));
  }
  return LB_OK;
}

template <typename T>
lb_status upload_tables(Plan& p, const std::vector<PupilSys>& sys, const std::vector<int2>& boxes,
                        const std::vector<double>& w, std::vector<double>* dc_host) {
  const int n = (int)sys.size();
  PupilSys* sys_dev = nullptr;
  double* dc = nullptr;
  LB_HIP(hipMalloc(&p.H, sizeof(cx<T>) * (size_t)std::max(n, 1) * p.W * p.W));
  LB_HIP(hipMalloc(&p.box, sizeof(int2) * std::max(n, 1)));
  LB_HIP(hipMalloc(&p.wsys, sizeof(T) * std::max(n, 1)));
  LB_HIP(hipMalloc(&sys_dev, sizeof(PupilSys) * std::max(n, 1)));
  LB_HIP(hipMalloc(&dc, sizeof(double) * std::max(n, 1)));
  std::vector<T> wt(w.begin(), w.end());
  lb_status rc = LB_OK;
  if (n > 0) {
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMemcpy(p.box, boxes.data(), sizeof(int2) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p.wsys, wt.data(), sizeof(T) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(sys_dev, sys.data(), sizeof(PupilSys) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(dc, 0, sizeof(double) * n);
    p.pp.W = p.W;
    if (e == hipSuccess) e = pupil_tables<T>(sys_dev, n, p.pp, (cx<T>*)p.H, dc, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess && dc_host) {
      dc_host->resize(n);
      e = hipMemcpy(dc_host->data(), dc, sizeof(double) * n, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) rc = fail(LB_ERR_HIP, std::string("pupil tables: ") + hipGetErrorString(e));
  }
  (void)hipFree(sys_dev);
  (void)hipFree(dc);
  return rc;
}

// Bounding box (a0, a1, b0, b1) of the non-zero transfer entries of each
// system, read back from the device tables; a1 < 0 for an all-zero table.
template <typename T>
lb_status table_support(const Plan& p, int n, std::vector<int4>* sup) {
  const size_t WW = (size_t)p.W * p.W;
  std::vector<cx<T>> h((size_t)n * WW);
  hipError_t e = hipMemcpy(h.data(), p.H, sizeof(cx<T>) * h.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(LB_ERR_HIP, std::string("table readback: ") + hipGetErrorString(e));
  sup->assign(n, make_int4(p.W, -1, p.W, -1));
  for (int j = 0; j < n; ++j) {
    int4& b = (*sup)[j];
    const cx<T>* t = h.data() + (size_t)j * WW;
    for (int a = 0; a < p.W; ++a)
      for (int c = 0; c < p.W; ++c)
        if (t[a * p.W + c].re != 0 || t[a * p.W + c].im != 0) {
          b.x = std::min(b.x, a);
          b.y = std::max(b.y, a);
          b.z = std::min(b.z, c);
          b.w = std::max(b.w, c);
        }
  }
  return LB_OK;
}

void free_tables(Plan& p) {
  if (p.H) (void)hipFree(p.H);
  if (p.box) (void)hipFree(p.box);
  if (p.wsys) (void)hipFree(p.wsys);
  p.H = nullptr;
  p.box = nullptr;
  p.wsys = nullptr;
}

// ---------------------------------------------------------------- workspace
struct Layout {
  size_t off = 0;
  size_t take(size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  }
};

struct Work {
  int n_masks, n_img, n_chunks, chunk, n_items;
  // spread layout (chunk LB_CHUNK_SPREAD, fused kernels): one run of systems
  // per resident block (runs of them), up to n_chunks image segments each
  int spread, runs;
  int pass_img;  // staged path: images whose row tables are held at once
  int nrc;       // staged adjoint: band-gradient partials per item (row chunks)
  size_t Xm, RTm, partial, IM, seg, RTc, Y, TT, gM, RTg, GY, TTc, gX, GX, TTm, queue, G;
  size_t total;
};

// Blocks of a fused coherent-system kernel resident on the whole GPU at once
// (LDS and threads per CU).
int resident_blocks(const Plan& p, const FieldConfig& fc, bool adjoint) {
  const size_t sm = field_smem_bytes(p.precision, fc, p.W, adjoint);
  int occ = (int)std::max<size_t>(1, kLdsLimit / std::max<size_t>(sm, 1));
  occ = std::min(occ, std::max(1, 2048 / fc.nt));
  return p.cus * occ;
}

// use: the coherent-system kernel shape the work is laid out for (default:
// the plan's own forward or adjoint shape).
Work layout(const Plan& p, int n_masks, bool backward, int chunk_forced = 0,
            const FieldConfig* use = nullptr, size_t staged_pass = kStagedPassBytes) {
  Work w{};
  w.n_masks = n_masks;
  w.n_img = n_masks * p.n_planes;
  int chunk = chunk_forced > 0 ? chunk_forced : (backward ? p.chunk_adj : p.chunk_fwd);
  const FieldConfig& fc = use ? *use : (backward ? p.adj : p.fwd);
  if (chunk == LB_CHUNK_SPREAD && fc.csplit > 0 && p.spp > 0) {
    w.spread = 1;
    w.runs = resident_blocks(p, fc, backward);
    const long long total = (long long)w.n_img * p.spp;
    const long long run = (total + w.runs - 1) / w.runs;
    w.n_chunks = (int)((run + p.spp - 1) / p.spp) + 1;  // image segments per run, at most
    w.chunk = LB_CHUNK_SPREAD;
    w.n_items = w.n_chunks * w.runs;
  } else if (chunk <= 0) {
    const long long total = (long long)w.n_img * p.spp;
    const long long target = 4LL * p.cus;
    int n_chunks = (int)std::max<long long>(1, std::min<long long>(p.spp, (target + w.n_img - 1) / std::max(w.n_img, 1)));
    (void)total;
    chunk = (p.spp + n_chunks - 1) / n_chunks;
  }
  if (!w.spread) {
    chunk = std::max(1, std::min(chunk, std::max(p.spp, 1)));
    w.chunk = chunk;
    w.n_chunks = std::max(1, (p.spp + chunk - 1) / chunk);
    w.n_items = w.n_img * w.n_chunks;
  }
  const size_t C = p.csz(), R = p.rsz();
  Layout L;
  w.queue = L.take(256);  // work-item counter of the coherent-system launch
  w.Xm = L.take(C * n_masks * p.Wbm * p.Wbm);
  w.RTm = L.take(C * (size_t)n_masks * band_ld(p.Wbm) * p.N);
  w.partial = L.take(R * (size_t)w.n_items * p.M * p.M);
  w.nrc = backward && fc.csplit == 0 ? staged_row_chunks(fc, p.W) : 1;
  if (fc.csplit == 0) {
    // row tables of the staged path, a pass of images at a time, sized to
    // stay in the last-level cache between its launches
    const size_t per_img = staged_table_bytes(p.precision, p.M, p.W) * (size_t)std::max(p.spp, 1);
    w.pass_img = (int)std::max<size_t>(1, std::min<size_t>((size_t)w.n_img,
                                                           staged_pass / per_img));
    w.G = L.take(per_img * (size_t)w.pass_img);
  }
  if (!w.spread) {
    // split-sum scratch: coarse partials (forward) or band gradients (backward)
    const long long mm = (long long)p.M * p.M, bb = 2LL * p.Wbm * p.Wbm;
    size_t seg = R * (size_t)w.n_img * segment_splits(w.n_img, w.n_chunks, mm) * mm;
    if (backward)
      seg = std::max<size_t>(seg, R * (size_t)n_masks *
                                      segment_splits(n_masks, p.n_planes * w.n_chunks * w.nrc,
                                                     bb) * bb);
    w.seg = L.take(seg);
  }
  if (!p.direct) w.IM = L.take(R * (size_t)w.n_img * p.M * p.M);
  if (!p.direct) {
    w.RTc = L.take(C * (size_t)w.n_img * band_ld(p.WbI) * p.M);
    w.Y = L.take(C * (size_t)w.n_img * p.WbI * p.WbI);
    w.TT = L.take(C * (size_t)w.n_img * band_ld(p.WbI) * p.N);
  }
  if (backward) {
    if (!p.direct) {
      w.gM = L.take(R * (size_t)w.n_img * p.M * p.M);
      w.RTg = L.take(C * (size_t)w.n_img * band_ld(p.WbI) * p.N);
      w.GY = L.take(C * (size_t)w.n_img * p.WbI * p.WbI);
      w.TTc = L.take(C * (size_t)w.n_img * band_ld(p.WbI) * p.M);
    }
    w.gX = L.take(C * (size_t)w.n_items * w.nrc * p.Wbm * p.Wbm);
    w.GX = L.take(C * (size_t)n_masks * p.Wbm * p.Wbm);
    w.TTm = L.take(C * (size_t)n_masks * band_ld(p.Wbm) * p.N);
  }
  w.total = L.off;
  return w;
}

FieldArgs field_args(const Plan& p, const void* Xm, const Work& w) {
  FieldArgs a{};
  a.X = Xm;
  a.Wb = p.Wbm;
  a.wrapN = p.wrapm;
  a.H = p.H;
  a.W = p.W;
  a.box = p.box;
  a.wsys = p.wsys;
  a.map = ItemMap{w.n_chunks, p.n_planes, p.spp, w.chunk, w.n_items, w.spread, w.runs,
                  (long long)w.n_img * p.spp};
  return a;
}

int grid_for(const Plan& p, const FieldConfig& fc, int n_items, bool adjoint) {
  return std::max(1, std::min(n_items, resident_blocks(p, fc, adjoint)));
}

hipError_t profile_events(Plan& p, int* e0, int* e1, bool adjoint = false) {
  for (int k = 0; k < 2; ++k) {
    hipEvent_t e;
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return r;
    p.ev_pool.push_back(e);
  }
  *e0 = (int)p.ev_pool.size() - 2;
  *e1 = (int)p.ev_pool.size() - 1;
  (adjoint ? p.ev_pairs_adj : p.ev_pairs).emplace_back(*e0, *e1);
  return hipSuccess;
}

// Forward, phase 1: band spectra of the masks.
template <typename T>
lb_status fwd_pre(Plan& p, const void* masks, int cplx, int n_masks, char* ws, const Work& w,
                  hipStream_t st) {
  const T invN2 = (T)(1.0 / ((double)p.N * (double)p.N));
  LB_HIP(band_analysis<T>(p.N, masks, cplx, (long long)p.N * p.N * (cplx ? 2 : 1), 0, 1, n_masks,
                          p.nu0m, p.Wbm, invN2, (cx<T>*)(ws + w.RTm), (cx<T>*)(ws + w.Xm), st));
  return LB_OK;
}

// The relaxed-resist objective evaluated where the images are synthesised
// (lb_forward_objective): targets [n_img][N][N], outputs as
// lb_resist_objective's, the band spectrum of the intensity derivative
// (sens_band) and, optionally, the derivative itself; rt / pc / pe are
// workspace offsets of the scratch that pass uses.
struct ObjArgs {
  const void* target;
  double scale, level, a, sign;
  void* cost;
  void* total;
  int32_t* mismatch;
  void* sens_band;
  void* d_images;
  size_t rt, pc, pe;
};

// Forward, phase 3: reduce the coarse partials, lift them to the N x N
// images (or sum them directly when the coarse grid is the image grid).
template <typename T>
lb_status fwd_post(Plan& p, void* images, void* peak, char* ws, const Work& w, hipStream_t st,
                   const ObjArgs* obj = nullptr) {
  T* partial = (T*)(ws + w.partial);
  // the per-image peak is a running maximum (atomics) in the final pass:
  // computed only when the caller asks for it
  T* pk = (T*)peak;
  const long long total = (long long)w.n_img * p.spp;
  if (p.direct) {
    if (pk) LB_HIP(fill<T>(pk, (T)0, w.n_img, st));
    if (w.spread)
      LB_HIP(spread_sum<T>(partial, w.n_img, 1, p.spp, total, w.runs, w.n_chunks,
                           (long long)p.N * p.N, (T*)images, pk, st));
    else
      LB_HIP(segment_sum<T>(partial, w.n_img, w.n_chunks, (long long)p.N * p.N, (T*)images, pk,
                            (T*)(ws + w.seg), st));
  } else {
    const T invM2 = (T)(1.0 / ((double)p.M * (double)p.M));
    cx<T>* Y = (cx<T>*)(ws + w.Y);
    T* IM = (T*)(ws + w.IM);
    // the partial sums and the coarse band spectrum in one pass (one block per
    // image) when the images are enough to occupy the device; fewer images
    // spread their partial sums over more blocks in the separate passes
    const hipError_t fused =
        2 * w.n_img >= p.cus
            ? coarse_band<T>(p.M, partial, w.spread, w.n_chunks, p.spp, total, w.runs,
                             w.n_chunks, w.n_img, p.QI, invM2, Y, pk, st)
            : hipErrorNotSupported;
    if (fused != hipErrorNotSupported) {
      LB_HIP(fused);  // the peaks are zeroed there
    } else {
      if (pk) LB_HIP(fill<T>(pk, (T)0, w.n_img, st));
      // one work item per image: its partial is the image's coarse intensity
      const T* im = w.spread || w.n_chunks > 1 ? IM : partial;
      if (w.spread)
        LB_HIP(spread_sum<T>(partial, w.n_img, 1, p.spp, total, w.runs, w.n_chunks,
                             (long long)p.M * p.M, IM, (T*)nullptr, st));
      else if (w.n_chunks > 1)
        LB_HIP(segment_sum<T>(partial, w.n_img, w.n_chunks, (long long)p.M * p.M, IM,
                              (T*)nullptr, (T*)(ws + w.seg), st));
      LB_HIP(band_analysis<T>(p.M, im, 0, (long long)p.M * p.M, 0, 1, w.n_img, -p.QI, p.WbI,
                              invM2, (cx<T>*)(ws + w.RTc), Y, st));
    }
    if (obj) {
      const double npix = (double)p.N * (double)p.N;
      const ResistPix<T> rp{(T)obj->scale, (T)obj->level, (T)obj->a, (T)obj->sign,
                            (T)(obj->sign * 2.0 / npix * obj->a * obj->scale)};
      LB_HIP(synth_objective<T>(p.N, Y, w.n_img, -p.QI, p.WbI, (cx<T>*)(ws + w.TT),
                                (const T*)obj->target, rp, (T*)images, (T*)obj->d_images,
                                (cx<T>*)(ws + obj->rt), (cx<T>*)obj->sens_band,
                                (double*)(ws + obj->pc), (int*)(ws + obj->pe), (T*)obj->cost,
                                (T*)obj->total, obj->mismatch, st));
    } else {
      LB_HIP(band_synthesis<T>(p.N, Y, w.n_img, -p.QI, p.WbI, 1, (T)1, 0, (cx<T>*)(ws + w.TT),
                               (T*)images, pk, st));
    }
  }
  return LB_OK;
}

// spectrum: the masks' band spectra from lb_mask_spectrum (then masks is
// unused), or null to compute them here.
template <typename T>
lb_status forward_t(Plan& p, const void* masks, int cplx, int n_masks, void* images, void* peak,
                    char* ws, const Work& w, hipStream_t st, const void* spectrum = nullptr,
                    bool may_profile = true, const ObjArgs* obj = nullptr) {
  if (!spectrum) {
    lb_status rc = fwd_pre<T>(p, masks, cplx, n_masks, ws, w, st);
    if (rc != LB_OK) return rc;
  }
  FieldConfig fc = p.fwd;
  fc.grid = w.spread ? w.runs : grid_for(p, fc, w.n_items, false);
  FieldArgs a = field_args(p, spectrum ? spectrum : ws + w.Xm, w);
  // spread runs are static (block b takes items b, b + G, ...): no queue
  a.queue = w.spread ? nullptr : (int*)(ws + w.queue);
  if (a.queue) LB_HIP(fill<float>((float*)a.queue, 0.0f, 1, st));
  int e0 = -1, e1 = -1;
  const bool prof = p.profile && may_profile;
  if (prof) {
    LB_HIP(profile_events(p, &e0, &e1));
    LB_HIP(hipEventRecord(p.ev_pool[e0], st));
  }
  if (fc.csplit == 0) {
    for (int i0 = 0; i0 < w.n_img; i0 += w.pass_img)
      LB_HIP(field_forward_staged(p.precision, fc, a, ws + w.partial, ws + w.G, i0,
                                  std::min(w.pass_img, w.n_img - i0), p.cus, st));
  } else {
    LB_HIP(field_forward(p.precision, fc, a, ws + w.partial, st));
  }
  if (prof) LB_HIP(hipEventRecord(p.ev_pool[e1], st));
  return fwd_post<T>(p, images, peak, ws, w, st, obj);
}

// Backward, phase 1: band spectra of the masks and the coarse sensitivity
// gM (the exact adjoint of the lift from M x M to N x N).
// gy: the band spectrum of dimg (lb_forward_objective's sens_band), given
// instead of dimg (coarse grids only).
template <typename T>
lb_status bwd_pre(Plan& p, const void* masks, int n_masks, const void* dimg, char* ws,
                  const Work& w, const T** gM, hipStream_t st, bool have_spectrum = false,
                  const void* gy = nullptr) {
  const T invN2 = (T)(1.0 / ((double)p.N * (double)p.N));
  if (!have_spectrum)
    LB_HIP(band_analysis<T>(p.N, masks, 0, (long long)p.N * p.N, 0, 1, n_masks, p.nu0m, p.Wbm,
                            invN2, (cx<T>*)(ws + w.RTm), (cx<T>*)(ws + w.Xm), st));
  *gM = (const T*)dimg;
  if (!p.direct) {
    const T invM2 = (T)(1.0 / ((double)p.M * (double)p.M));
    const cx<T>* GY = (const cx<T>*)gy;
    if (!GY) {
      LB_HIP(band_analysis<T>(p.N, dimg, 0, (long long)p.N * p.N, 0, 1, w.n_img, -p.QI, p.WbI,
                              (T)1, (cx<T>*)(ws + w.RTg), (cx<T>*)(ws + w.GY), st));
      GY = (const cx<T>*)(ws + w.GY);
    }
    LB_HIP(band_synthesis<T>(p.M, GY, w.n_img, -p.QI, p.WbI, 1, invM2, 0, (cx<T>*)(ws + w.TTc),
                             (T*)(ws + w.gM), (T*)nullptr, st));
    *gM = (const T*)(ws + w.gM);
  }
  return LB_OK;
}

// Backward, phase 3: sum the band-gradient partials per mask and synthesise
// the real mask gradient.
// cout: the complex gradient d/dRe + i d/dIm (complex masks) instead of its real part.
// te: write d/dtheta of the ILT parameter instead of the mask gradient
// (band_synthesis_theta; real masks over a symmetric band inside N/16).
template <typename T>
lb_status bwd_post(Plan& p, int n_masks, void* dmask, char* ws, const Work& w, hipStream_t st,
                   bool cout = false, const ThetaEpi<T>* te = nullptr) {
  const T invN2 = (T)(1.0 / ((double)p.N * (double)p.N));
  cx<T>* GX = (cx<T>*)(ws + w.GX);
  if (w.spread)
    LB_HIP(spread_sum<T>((const T*)(ws + w.gX), n_masks, p.n_planes, p.spp,
                         (long long)w.n_img * p.spp, w.runs, w.n_chunks, 2LL * p.Wbm * p.Wbm,
                         (T*)GX, (T*)nullptr, st));
  else
    LB_HIP(segment_sum<T>((const T*)(ws + w.gX), n_masks, p.n_planes * w.n_chunks * w.nrc,
                          2LL * p.Wbm * p.Wbm, (T*)GX, (T*)nullptr, (T*)(ws + w.seg), st));
  if (te)
    LB_HIP(band_synthesis_theta<T>(p.N, GX, n_masks, p.nu0m, p.Wbm, invN2, (cx<T>*)(ws + w.TTm),
                                   *te, (T*)dmask, st));
  else
    LB_HIP(band_synthesis<T>(p.N, GX, n_masks, p.nu0m, p.Wbm, cout ? 0 : 1, invN2, cout ? 1 : 0,
                             (cx<T>*)(ws + w.TTm), (T*)dmask, (T*)nullptr, st));
  return LB_OK;
}

template <typename T>
lb_status backward_t(Plan& p, const void* masks, int n_masks, const void* dimg, void* dmask,
                     char* ws, const Work& w, hipStream_t st, const void* spectrum = nullptr,
                     bool cout = false, const void* gy = nullptr,
                     const ThetaEpi<T>* te = nullptr) {
  const T* gM = nullptr;
  lb_status rc = bwd_pre<T>(p, masks, n_masks, dimg, ws, w, &gM, st, spectrum != nullptr, gy);
  if (rc != LB_OK) return rc;
  FieldConfig fc = p.adj;
  fc.grid = w.spread ? w.runs : grid_for(p, fc, w.n_items, true);
  FieldArgs a = field_args(p, spectrum ? spectrum : ws + w.Xm, w);
  a.queue = w.spread ? nullptr : (int*)(ws + w.queue);
  if (a.queue) LB_HIP(fill<float>((float*)a.queue, 0.0f, 1, st));
  int e0 = -1, e1 = -1;
  if (p.profile) {
    LB_HIP(profile_events(p, &e0, &e1, true));
    LB_HIP(hipEventRecord(p.ev_pool[e0], st));
  }
  if (fc.csplit == 0) {
    for (int i0 = 0; i0 < w.n_img; i0 += w.pass_img)
      LB_HIP(field_adjoint_staged(p.precision, fc, a, gM, ws + w.gX, ws + w.G, i0,
                                  std::min(w.pass_img, w.n_img - i0), p.cus, st));
  } else {
    LB_HIP(field_adjoint(p.precision, fc, a, gM, ws + w.gX, st));
  }
  if (p.profile) LB_HIP(hipEventRecord(p.ev_pool[e1], st));
  return bwd_post<T>(p, n_masks, dmask, ws, w, st, cout, te);
}

// mask band covering every box: symmetric [-Q, Q], or the whole grid
void set_mask_band(Plan& p, int Q) {
  if (2 * Q + 1 > p.N) {
    p.nu0m = -(p.N / 2);
    p.Wbm = p.N;
    p.wrapm = p.N;
  } else {
    p.nu0m = -Q;
    p.Wbm = 2 * Q + 1;
    p.wrapm = 0;
  }
}

}  // namespace

struct lb_plan_s {
  Plan p;
};

// Several plans (clips of different sizes / optics) imaged together: their
// band passes run per clip, their coherent systems in ONE persistent launch
// of the coherent-system kernel over a descriptor table of work items.
// Clips with the same band geometry (size, mask band, coarse grid and its
// band, planes, systems per plane) form a group: their band passes run as one
// multi-mask pass over a group workspace laid out like a plan's for that many
// masks.  A group's masks and images are used in place when the caller's
// buffers are consecutive in memory (clip after clip), else through staging
// copies in the group workspace.
struct BatchGroup {
  std::vector<int> members;  // clip indices, in order
  Work wf, wb;
  size_t base_f = 0, base_b = 0;
  size_t stage_f = 0, stage_b = 0;  // staging offsets (masks, then images / image gradients)
};

struct lb_batch_s {
  std::vector<lb_plan> plans;  // not owned
  int precision = 32, M = 0, Wmax = 0, cus = 256, chunk = 1;
  FieldConfig fwd{}, adj{};
  std::vector<BatchGroup> groups;
  size_t ws_f = 0, ws_b = 0;
  ItemDesc* items_f = nullptr;
  ItemDesc* items_b = nullptr;
  int n_items_f = 0, n_items_b = 0;
  // the groups' band passes run concurrently between the coherent-system
  // launches: one side stream each (default priority, non-blocking), forked
  // from and joined to the caller's stream with events (graph capture records
  // them as parallel branches)
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork_ev = nullptr;
  ~lb_batch_s() {
    if (items_f) (void)hipFree(items_f);
    if (items_b) (void)hipFree(items_b);
    for (hipStream_t s : side) (void)hipStreamDestroy(s);
    for (hipEvent_t e : joins) (void)hipEventDestroy(e);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
  }
};

constexpr size_t kBatchSideStreams = 4;  // groups beyond this share streams round-robin

// body(group, stream) for every group of the batch: on the caller's stream
// when there is one group (or no side streams), else each group on a side
// stream between a fork from and a join to the caller's stream.  The join
// also runs when a group fails, so no forked work is left outside it.
template <class F>
static lb_status for_groups(lb_batch_s& b, hipStream_t st, F&& body) {
  const size_t G = b.groups.size();
  if (G < 2 || b.side.empty()) {
    for (BatchGroup& g : b.groups) {
      lb_status rc = body(g, st);
      if (rc != LB_OK) return rc;
    }
    return LB_OK;
  }
  const size_t S = std::min(G, b.side.size());
  LB_HIP(hipEventRecord(b.fork_ev, st));
  for (size_t k = 0; k < S; ++k) LB_HIP(hipStreamWaitEvent(b.side[k], b.fork_ev, 0));
  lb_status rc = LB_OK;
  for (size_t g = 0; g < G && rc == LB_OK; ++g) rc = body(b.groups[g], b.side[g % S]);
  for (size_t k = 0; k < S; ++k) {
    LB_HIP(hipEventRecord(b.joins[k], b.side[k]));
    LB_HIP(hipStreamWaitEvent(st, b.joins[k], 0));
  }
  return rc;
}

extern "C" {

int32_t lb_version(void) { return 100; }

const char* lb_last_error(void) { return g_err.c_str(); }

lb_status lb_abbe_plan_create(const lb_imaging_desc* d, lb_plan* out) {
  g_err.clear();
  if (!out) return fail(LB_ERR_ARG, "null output plan");
  *out = nullptr;
  lb_status rc = check_desc(d);
  if (rc != LB_OK) return rc;
  lb_plan h = new lb_plan_s();
  Plan& p = h->p;
  rc = set_geometry_common(p, d);
  if (rc != LB_OK) {
    delete h;
    return rc;
  }
  p.S = d->n_source;
  p.n_planes = d->n_planes;
  p.defocus.assign(d->defocus, d->defocus + d->n_planes);
  p.n_image.assign(d->n_image, d->n_image + d->n_planes);
  const double r = p.k_max / p.val;  // pupil radius in bins
  const int Wbox = (int)std::floor(2.0 * r) + 3;
  rc = choose_grid(p, Wbox, 0);  // the requested coarse grid applies to the final boxes
  if (rc != LB_OK) {
    delete h;
    return rc;
  }
  // boxes per source point (box coordinates: origin frequency index)
  std::vector<int> oy(p.S), ox(p.S);
  int Q = 0;
  for (int s = 0; s < p.S; ++s) {
    oy[s] = (int)std::floor(d->src_fy[s] / p.val - r);
    ox[s] = (int)std::floor(d->src_fx[s] / p.val - r);
    Q = std::max({Q, std::abs(oy[s]), std::abs(oy[s] + p.W - 1), std::abs(ox[s]),
                  std::abs(ox[s] + p.W - 1)});
  }
  set_mask_band(p, Q);
  // Jones states, pupil.py:272-285
  struct St {
    double jx, jy, w;
  };
  auto states = [&](double phi_s) {
    std::vector<St> v;
    switch (p.pol) {
      case LB_POL_X: v.push_back({1.0, 0.0, 1.0}); break;
      case LB_POL_Y: v.push_back({0.0, 1.0, 1.0}); break;
      case LB_POL_TE: v.push_back({-std::sin(phi_s), std::cos(phi_s), 1.0}); break;
      case LB_POL_TM: v.push_back({std::cos(phi_s), std::sin(phi_s), 1.0}); break;
      default:
        v.push_back({1.0, 0.0, 0.5});
        v.push_back({0.0, 1.0, 0.5});
    }
    return v;
  };
  const int per_source = p.vector ? (p.pol == LB_POL_UNPOLARISED ? 6 : 3) : 1;
  p.spp = p.S * per_source;
  std::vector<PupilSys> sys;
  std::vector<int2> boxes;
  std::vector<double> wts;
  std::vector<int> plane_of;
  sys.reserve((size_t)p.spp * p.n_planes);
  for (int pl = 0; pl < p.n_planes; ++pl) {
    for (int s = 0; s < p.S; ++s) {
      PupilSys base{};
      base.fsx = d->src_fx[s];
      base.fsy = d->src_fy[s];
      base.weight = 1.0;
      base.defocus = p.defocus[pl];
      base.n_image = p.n_image[pl];
      base.box_y = oy[s];
      base.box_x = ox[s];
      if (d->n_refl > 0) {
        base.sin_x = d->src_sin_x[s];
        base.sin_y = d->src_sin_y[s];
      }
      const int2 bi = make_int2(oy[s] - p.nu0m, ox[s] - p.nu0m);
      if (!p.vector) {
        base.comp = -1;
        sys.push_back(base);
        boxes.push_back(bi);
        wts.push_back(d->src_weight[s]);
        plane_of.push_back(pl);
        continue;
      }
      for (const St& stt : states(d->src_phi[s])) {
        for (int c = 0; c < 3; ++c) {
          PupilSys v = base;
          v.jx = stt.jx;
          v.jy = stt.jy;
          v.comp = c;
          sys.push_back(v);
          boxes.push_back(bi);
          wts.push_back(d->src_weight[s] * stt.w);
          plane_of.push_back(pl);
        }
      }
    }
  }
  std::vector<double> dc;
  rc = p.precision == 64 ? upload_tables<double>(p, sys, boxes, wts, &dc)
                         : upload_tables<float>(p, sys, boxes, wts, &dc);
  if (rc != LB_OK) {
    delete h;
    return rc;
  }
  // Tighten every box to its table's non-zero support: a disc of radius r
  // covers at most floor(2r) + 1 bins per axis, two fewer than the box above,
  // and the coarse grid only has to hold 2 W - 1 points for the common span W.
  {
    std::vector<int4> sup;
    rc = p.precision == 64 ? table_support<double>(p, (int)sys.size(), &sup)
                           : table_support<float>(p, (int)sys.size(), &sup);
    if (rc != LB_OK) {
      delete h;
      return rc;
    }
    int Wt = 1;
    for (const int4& b : sup)
      if (b.y >= 0) Wt = std::max({Wt, b.y - b.x + 1, b.w - b.z + 1});
    if (Wt < p.W || d->coarse_m != 0) {
      free_tables(p);
      for (size_t j = 0; j < sys.size(); ++j)
        if (sup[j].y >= 0) {
          sys[j].box_y += sup[j].x;
          sys[j].box_x += sup[j].z;
        }
      rc = choose_grid(p, std::min(Wt, p.W), d->coarse_m);
      if (rc != LB_OK) {
        delete h;
        return rc;
      }
      int Qt = 0;
      for (const PupilSys& v : sys)
        Qt = std::max({Qt, std::abs(v.box_y), std::abs(v.box_y + p.W - 1), std::abs(v.box_x),
                       std::abs(v.box_x + p.W - 1)});
      set_mask_band(p, Qt);
      for (size_t j = 0; j < sys.size(); ++j)
        boxes[j] = make_int2(sys[j].box_y - p.nu0m, sys[j].box_x - p.nu0m);
      rc = p.precision == 64 ? upload_tables<double>(p, sys, boxes, wts, nullptr)
                             : upload_tables<float>(p, sys, boxes, wts, nullptr);
      if (rc != LB_OK) {
        delete h;
        return rc;
      }
    }
  }
  p.clear.assign(p.n_planes, 0.0);
  for (size_t j = 0; j < sys.size(); ++j) p.clear[plane_of[j]] += wts[j] * dc[j];
  *out = h;
  return LB_OK;
}

lb_status lb_socs_plan_create(const lb_imaging_desc* d, double tol, int32_t max_kernels,
                              lb_plan* out) {
  g_err.clear();
  if (!out) return fail(LB_ERR_ARG, "null output plan");
  *out = nullptr;
  lb_status rc = check_desc(d);
  if (rc != LB_OK) return rc;
  if (d->vector || d->n_refl > 0)
    return fail(LB_ERR_ARG, "SOCS kernels apply to the scalar thin-mask process only");
  if (d->n_planes != 1) return fail(LB_ERR_ARG, "SOCS plans image one plane");
  lb_plan h = new lb_plan_s();
  Plan& p = h->p;
  rc = set_geometry_common(p, d);
  if (rc != LB_OK) {
    delete h;
    return rc;
  }
  p.socs = 1;
  p.S = d->n_source;
  p.n_planes = 1;
  p.defocus.assign(1, d->defocus[0]);
  p.n_image.assign(1, d->n_image[0]);
  const double r = p.k_max / p.val;
  // the band |f| <= 2 NA / lambda (hopkins.py:144-147), as a square box
  int Qb = (int)std::floor(2.0 * r * (1.0 + 1e-9)) + 1;
  set_mask_band(p, Qb);
  const int Wt = p.Wbm;
  rc = choose_grid(p, Wt, d->coarse_m);
  if (rc != LB_OK) {
    delete h;
    return rc;
  }
  if (p.W != Wt) {
    delete h;
    return fail(LB_ERR_UNSUPPORTED, "SOCS band wider than the grid");
  }
  // n_band as the reference counts it
  {
    int cnt = 0;
    const double lim = 2.0 * p.k_max + 1e-9 * p.k_max;
    for (int i = 0; i < p.N; ++i)
      for (int j = 0; j < p.N; ++j) {
        const int nuy = i <= (p.N - 1) / 2 ? i : i - p.N;
        const int nux = j <= (p.N - 1) / 2 ? j : j - p.N;
        if (std::hypot(nux * p.val, nuy * p.val) <= lim) ++cnt;
      }
    p.n_band = cnt;
  }
  const int S = p.S;
  const long long F = (long long)Wt * Wt;
  // A[s][f] = sqrt(w_s) P_s(f) on the band box, FP64 (hopkins.py:155-167)
  std::vector<PupilSys> sys(S);
  for (int s = 0; s < S; ++s) {
    PupilSys v{};
    v.fsx = d->src_fx[s];
    v.fsy = d->src_fy[s];
    v.weight = std::sqrt(d->src_weight[s]);
    v.defocus = p.defocus[0];
    v.n_image = p.n_image[0];
    v.comp = -1;
    v.box_y = p.nu0m;
    v.box_x = p.nu0m;
    sys[s] = v;
  }
  cx<double>* A = nullptr;
  cx<double>* G = nullptr;
  cx<double>* V = nullptr;
  double* ev = nullptr;
  int* counts = nullptr;
  void* ework = nullptr;
  PupilSys* sys_dev = nullptr;
  double* dc = nullptr;
  int* idx_dev = nullptr;
  double* isq_dev = nullptr;
  auto cleanup = [&]() {
    (void)hipFree(A);
    (void)hipFree(G);
    (void)hipFree(V);
    (void)hipFree(ev);
    (void)hipFree(counts);
    (void)hipFree(ework);
    (void)hipFree(sys_dev);
    (void)hipFree(dc);
    (void)hipFree(idx_dev);
    (void)hipFree(isq_dev);
  };
#define LB_HIP_C(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      cleanup();                                                                        \
      delete h;                                                                         \
      return fail(e_ == hipErrorOutOfMemory ? LB_ERR_OOM : LB_ERR_HIP,                  \
                  std::string(#call) + ": " + hipGetErrorString(e_));                   \
    }                                                                                   \
  } while (0)
  const int Sa = std::max(S, 1);
  LB_HIP_C(hipMalloc(&A, sizeof(cx<double>) * Sa * F));
  LB_HIP_C(hipMalloc(&G, sizeof(cx<double>) * Sa * Sa));
  LB_HIP_C(hipMalloc(&V, sizeof(cx<double>) * Sa * Sa));
  LB_HIP_C(hipMalloc(&ev, sizeof(double) * Sa));
  LB_HIP_C(hipMalloc(&counts, 8 * sizeof(int)));
  LB_HIP_C(hipMalloc(&ework, std::max(eigh_work_bytes(Sa), gram_work_bytes(Sa, F))));
  LB_HIP_C(hipMalloc(&sys_dev, sizeof(PupilSys) * Sa));
  LB_HIP_C(hipMalloc(&dc, sizeof(double) * Sa));
  std::vector<double> evals(S);
  if (S > 0) {
    LB_HIP_C(hipMemcpy(sys_dev, sys.data(), sizeof(PupilSys) * S, hipMemcpyHostToDevice));
    p.pp.W = Wt;
    LB_HIP_C(pupil_tables<double>(sys_dev, S, p.pp, A, dc, 0));
    LB_HIP_C(gram_f64(A, S, F, G, ework, 0));
    LB_HIP_C(hermitian_eigh_f64(G, V, S, ework, ev, counts, 0));
    LB_HIP_C(hipDeviceSynchronize());
    LB_HIP_C(hipMemcpy(evals.data(), ev, sizeof(double) * S, hipMemcpyDeviceToHost));
    int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    LB_HIP_C(hipMemcpy(cnt, counts, 8 * sizeof(int), hipMemcpyDeviceToHost));
    const int bad = cnt[2] != 0 ? cnt[2] : cnt[6];
    if (bad != 0) {
      cleanup();
      delete h;
      return fail(LB_ERR_HIP, bad == 1 ? "SOCS eigensolve: the QL iteration did not converge"
                                       : "SOCS eigensolve: rotation log full");
    }
    p.eig_sweeps = cnt[0] + cnt[4];
  }
  // order, truncation: hopkins.py:172-182
  std::vector<int> order(S);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return evals[a] > evals[b]; });
  double trace = 0;
  for (double e : evals) trace += std::max(e, 0.0);
  const double lmax = S ? evals[order[0]] : 0.0;
  std::vector<int> keep;
  std::vector<double> isq;
  for (int i = 0; i < S; ++i) {
    const double e = evals[order[i]];
    if (!(e > tol * lmax) || !(e > 0.0)) continue;
    if (max_kernels >= 0 && (int)keep.size() >= max_kernels) break;
    keep.push_back(order[i]);
    p.weights.push_back(e);
    isq.push_back(1.0 / std::sqrt(e));
  }
  // (the reference applies the cap by position after the tolerance; kernels
  // failing the tolerance are all at the tail of the descending order)
  p.K = (int)keep.size();
  double kept = 0;
  for (double e : p.weights) kept += e;
  p.energy = trace > 0 ? kept / trace : 1.0;
  const int Ka = std::max(p.K, 1);
  LB_HIP_C(hipMalloc(&p.phi128, sizeof(cx<double>) * Ka * F));
  LB_HIP_C(hipMalloc(&idx_dev, sizeof(int) * Ka));
  LB_HIP_C(hipMalloc(&isq_dev, sizeof(double) * Ka));
  if (p.K > 0) {
    LB_HIP_C(hipMemcpy(idx_dev, keep.data(), sizeof(int) * p.K, hipMemcpyHostToDevice));
    LB_HIP_C(hipMemcpy(isq_dev, isq.data(), sizeof(double) * p.K, hipMemcpyHostToDevice));
    LB_HIP_C(socs_phi_f64(A, V, S, F, idx_dev, isq_dev, p.K, (cx<double>*)p.phi128, 0));
  }
  // systems: one per kernel, weight lambda_k, box = the whole band
  p.spp = p.K;
  const size_t csz = p.csz();
  LB_HIP_C(hipMalloc(&p.H, csz * Ka * F));
  LB_HIP_C(hipMalloc(&p.box, sizeof(int2) * Ka));
  LB_HIP_C(hipMalloc(&p.wsys, p.rsz() * Ka));
  if (p.K > 0) {
    if (p.precision == 64) {
      LB_HIP_C(hipMemcpy(p.H, p.phi128, csz * p.K * F, hipMemcpyDeviceToDevice));
      LB_HIP_C(hipMemcpy(p.wsys, p.weights.data(), sizeof(double) * p.K, hipMemcpyHostToDevice));
    } else {
      LB_HIP_C(cast_c128_c64((const cx<double>*)p.phi128, (cx<float>*)p.H, (long long)p.K * F, 0));
      std::vector<float> wf(p.weights.begin(), p.weights.end());
      LB_HIP_C(hipMemcpy(p.wsys, wf.data(), sizeof(float) * p.K, hipMemcpyHostToDevice));
    }
    std::vector<int2> boxes(p.K, make_int2(0, 0));
    LB_HIP_C(hipMemcpy(p.box, boxes.data(), sizeof(int2) * p.K, hipMemcpyHostToDevice));
    LB_HIP_C(hipDeviceSynchronize());
  }
  // clear = sum_k lambda_k |phi_k(0)|^2 (hopkins.py:190-191)
  double clear = 0;
  if (p.K > 0) {
    const long long dc_idx = (long long)(-p.nu0m) * Wt + (-p.nu0m);
    std::vector<cx<double>> at(p.K);
    for (int k = 0; k < p.K; ++k)
      LB_HIP_C(hipMemcpy(&at[k], (cx<double>*)p.phi128 + (long long)k * F + dc_idx,
                         sizeof(cx<double>), hipMemcpyDeviceToHost));
    for (int k = 0; k < p.K; ++k)
      clear += p.weights[k] * (at[k].re * at[k].re + at[k].im * at[k].im);
  }
  p.clear.assign(1, clear);
#undef LB_HIP_C
  cleanup();
  *out = h;
  return LB_OK;
}

lb_status lb_pupil_dense(const lb_imaging_desc* d, double fsx, double fsy, int32_t plane,
                         void* out, void* stream) {
  g_err.clear();
  if (!d || !out) return fail(LB_ERR_ARG, "null descriptor or output");
  if (d->n < 1) return fail(LB_ERR_ARG, "grid side must be >= 1");
  if (!(d->pixel_size > 0) || !(d->wavelength > 0) || !(d->na > 0))
    return fail(LB_ERR_ARG, "pixel_size, wavelength and NA must be > 0");
  if (d->n_planes < 1 || !d->defocus || !d->n_image || plane < 0 || plane >= d->n_planes)
    return fail(LB_ERR_ARG, "plane out of range");
  if (d->n_zernike > 0 && (!d->zernike_j || !d->zernike_c))
    return fail(LB_ERR_ARG, "null Zernike arrays");
  lb_imaging_desc thin = *d;
  thin.n_refl = 0;  // the pupil alone: no mask-side reflection
  Plan p;
  lb_status rc = set_geometry_common(p, &thin);
  if (rc != LB_OK) return rc;
  PupilSys s{};
  s.fsx = fsx;
  s.fsy = fsy;
  s.weight = 1.0;
  s.defocus = d->defocus[plane];
  s.n_image = d->n_image[plane];
  s.comp = -1;
  s.box_y = s.box_x = 0;  // the whole grid from bin 0: FFT order
  p.pp.W = p.N;
  PupilSys* sys_dev = nullptr;
  double* dc = nullptr;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipMalloc(&sys_dev, sizeof(PupilSys));
  if (e == hipSuccess) e = hipMalloc(&dc, sizeof(double));
  if (e == hipSuccess) e = hipMemcpyAsync(sys_dev, &s, sizeof(PupilSys), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = pupil_tables<double>(sys_dev, 1, p.pp, (cx<double>*)out, dc, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(sys_dev);
  (void)hipFree(dc);
  if (e != hipSuccess) return fail(LB_ERR_HIP, std::string("dense pupil: ") + hipGetErrorString(e));
  return LB_OK;
}

lb_status lb_plan_destroy(lb_plan plan) {
  delete plan;
  return LB_OK;
}

lb_status lb_plan_info_get(lb_plan plan, lb_plan_info* info) {
  if (!plan || !info) return fail(LB_ERR_ARG, "null plan or info");
  const Plan& p = plan->p;
  Work w = layout(p, 1, false);
  info->n = p.N;
  info->m = p.M;
  info->box = p.W;
  info->band = p.Wbm;
  info->n_planes = p.n_planes;
  info->n_systems = p.spp;
  info->precision = p.precision;
  info->n_kernels = p.socs ? p.K : 0;
  info->block_threads = p.fwd.nt;
  info->column_splits = p.fwd.csplit;
  info->chunk = w.chunk;
  info->grid_blocks = w.spread ? w.runs : grid_for(p, p.fwd, w.n_items, false);
  info->adj_block_threads = p.adj.nt;
  info->adj_column_splits = p.adj.csplit;
  info->adj_chunk = layout(p, 1, true).chunk;
  info->socs_sweeps = p.socs ? p.eig_sweeps : 0;
  return LB_OK;
}

lb_status lb_plan_clear(lb_plan plan, double* clear) {
  if (!plan || !clear) return fail(LB_ERR_ARG, "null plan or output");
  for (size_t i = 0; i < plan->p.clear.size(); ++i) clear[i] = plan->p.clear[i];
  return LB_OK;
}

lb_status lb_socs_weights(lb_plan plan, double* weights, double* energy, int32_t* n_band,
                          int32_t* n_source) {
  if (!plan) return fail(LB_ERR_ARG, "null plan");
  const Plan& p = plan->p;
  if (!p.socs) return fail(LB_ERR_ARG, "not a SOCS plan");
  if (weights)
    for (int k = 0; k < p.K; ++k) weights[k] = p.weights[k];
  if (energy) *energy = p.energy;
  if (n_band) *n_band = p.n_band;
  if (n_source) *n_source = p.S;
  return LB_OK;
}

lb_status lb_plan_set_shape(lb_plan plan, int32_t adjoint, int32_t block_threads,
                            int32_t column_splits, int32_t chunk) {
  if (!plan) return fail(LB_ERR_ARG, "null plan");
  Plan& p = plan->p;
  if (block_threads == 0) {  // back to the plan's own choice and automatic work items
    FieldConfig d{};
    if (!default_field_config(p, adjoint != 0, &d))
      return fail(LB_ERR_UNSUPPORTED, "no default coherent-system kernel shape");
    (adjoint ? p.adj : p.fwd) = d;
    (adjoint ? p.chunk_adj : p.chunk_fwd) = chunk > 0 || chunk == LB_CHUNK_SPREAD ? chunk : 0;
    return LB_OK;
  }
  FieldConfig fc{p.M, block_threads, column_splits, 0};
  if (!field_config_supported(p.precision, fc))
    return fail(LB_ERR_UNSUPPORTED, "no coherent-system kernel of shape M=" +
                                        std::to_string(p.M) + " threads=" +
                                        std::to_string(block_threads) + " splits=" +
                                        std::to_string(column_splits));
  if (field_smem_bytes(p.precision, fc, p.W, adjoint != 0) > kLdsLimit)
    return fail(LB_ERR_UNSUPPORTED, "shape does not fit in LDS");
  if (adjoint)
    p.adj = fc;
  else
    p.fwd = fc;
  if (chunk >= 0 || chunk == LB_CHUNK_SPREAD) (adjoint ? p.chunk_adj : p.chunk_fwd) = chunk;
  return LB_OK;
}

lb_status lb_plan_work_items(lb_plan plan, int32_t n_masks, int32_t adjoint, int32_t* chunk,
                             int32_t* n_items) {
  if (!plan || !chunk || !n_items || n_masks < 1) return fail(LB_ERR_ARG, "bad argument");
  const Work w = layout(plan->p, n_masks, adjoint != 0);
  *chunk = w.chunk;
  *n_items = w.n_items;
  return LB_OK;
}

lb_status lb_plan_shapes(lb_plan plan, int32_t adjoint, int32_t* shapes, int32_t max_shapes,
                         int32_t* n_shapes) {
  if (!plan || !n_shapes) return fail(LB_ERR_ARG, "null argument");
  Plan& p = plan->p;
  int n = 0;
  for (const FieldConfig& c : field_configs()) {
    if (c.M != p.M) continue;
    if (field_smem_bytes(p.precision, c, p.W, adjoint != 0) > kLdsLimit) continue;
    if (shapes && n < max_shapes) {
      shapes[2 * n] = c.nt;
      shapes[2 * n + 1] = c.csplit;
    }
    ++n;
  }
  *n_shapes = n;
  return LB_OK;
}

lb_status lb_profile_enable(lb_plan plan, int32_t on) {
  if (!plan) return fail(LB_ERR_ARG, "null plan");
  Plan& p = plan->p;
  for (hipEvent_t e : p.ev_pool) (void)hipEventDestroy(e);
  p.ev_pool.clear();
  p.ev_pairs.clear();
  p.ev_pairs_adj.clear();
  p.profile = on ? 1 : 0;
  return LB_OK;
}

}  // extern "C"

static lb_status profile_sum(Plan& p, const std::vector<std::pair<int, int>>& pairs,
                             int32_t* launches, double* total_ms) {
  double t = 0;
  for (auto& pr : pairs) {
    LB_HIP(hipEventSynchronize(p.ev_pool[pr.second]));
    float ms = 0;
    LB_HIP(hipEventElapsedTime(&ms, p.ev_pool[pr.first], p.ev_pool[pr.second]));
    t += ms;
  }
  *launches = (int32_t)pairs.size();
  *total_ms = t;
  return LB_OK;
}

extern "C" {

lb_status lb_profile_read(lb_plan plan, int32_t* launches, double* total_ms) {
  if (!plan || !launches || !total_ms) return fail(LB_ERR_ARG, "null argument");
  return profile_sum(plan->p, plan->p.ev_pairs, launches, total_ms);
}

lb_status lb_profile_read_adjoint(lb_plan plan, int32_t* launches, double* total_ms) {
  if (!plan || !launches || !total_ms) return fail(LB_ERR_ARG, "null argument");
  return profile_sum(plan->p, plan->p.ev_pairs_adj, launches, total_ms);
}

size_t lb_workspace_bytes(lb_plan plan, int32_t n_masks, int32_t backward) {
  if (!plan || n_masks < 1) return 0;
  return layout(plan->p, n_masks, backward != 0).total;
}

lb_status lb_forward(lb_plan plan, const void* masks, int32_t cplx, int32_t n_masks, void* images,
                     void* peak, void* workspace, size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!plan || !masks || !images) return fail(LB_ERR_ARG, "null plan, masks or images");
  if (n_masks < 1) return fail(LB_ERR_ARG, "n_masks must be >= 1");
  Plan& p = plan->p;
  if (p.spp < 1) return fail(LB_ERR_ARG, "plan has no coherent systems (empty source)");
  Work w = layout(p, n_masks, false);
  if (!workspace || ws_bytes < w.total)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(w.total) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  return p.precision == 64
             ? forward_t<double>(p, masks, cplx, n_masks, images, peak, (char*)workspace, w, st)
             : forward_t<float>(p, masks, cplx, n_masks, images, peak, (char*)workspace, w, st);
}

lb_status lb_forward_static(lb_plan plan, const void* masks, int32_t cplx, int32_t n_masks,
                            void* images, void* peak, void* stream) {
  g_err.clear();
  if (!plan || !masks || !images) return fail(LB_ERR_ARG, "null plan, masks or images");
  if (n_masks < 1) return fail(LB_ERR_ARG, "n_masks must be >= 1");
  Plan& p = plan->p;
  if (p.spp < 1) return fail(LB_ERR_ARG, "plan has no coherent systems (empty source)");
  void* ws = nullptr;
  lb_status rc = static_scratch(&ws);
  if (rc != LB_OK) return rc;
  const size_t cap = LB_STATIC_SCRATCH_BYTES;
  // lb_forward's own layout when it fits; else more systems per work item
  // (fewer coarse partials), then smaller staged passes
  Work w = layout(p, n_masks, false);
  for (int chunk = std::max(w.chunk, 1); w.total > cap && chunk < p.spp;) {
    chunk = std::min(p.spp, chunk * 2);
    w = layout(p, n_masks, false, chunk);
  }
  for (size_t pass = kStagedPassBytes / 2; w.total > cap && p.fwd.csplit == 0 && pass >= (1u << 20);
       pass /= 2)
    w = layout(p, n_masks, false, p.spp, nullptr, pass);
  if (w.total > cap)
    return fail(LB_ERR_WORKSPACE, "lb_forward_static: " + std::to_string(n_masks) +
                                      " masks need " + std::to_string(w.total) +
                                      " bytes of workspace, beyond the static scratch of " +
                                      std::to_string(cap) + " bytes; use lb_forward");
  hipStream_t st = (hipStream_t)stream;
  return p.precision == 64
             ? forward_t<double>(p, masks, cplx, n_masks, images, peak, (char*)ws, w, st, nullptr,
                                 false)
             : forward_t<float>(p, masks, cplx, n_masks, images, peak, (char*)ws, w, st, nullptr,
                                false);
}

lb_status lb_backward(lb_plan plan, const void* masks, int32_t n_masks, const void* dimg,
                      void* dmask, void* workspace, size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!plan || !masks || !dimg || !dmask) return fail(LB_ERR_ARG, "null argument");
  if (n_masks < 1) return fail(LB_ERR_ARG, "n_masks must be >= 1");
  Plan& p = plan->p;
  if (p.spp < 1) return fail(LB_ERR_ARG, "plan has no coherent systems (empty source)");
  Work w = layout(p, n_masks, true);
  if (!workspace || ws_bytes < w.total)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(w.total) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  return p.precision == 64
             ? backward_t<double>(p, masks, n_masks, dimg, dmask, (char*)workspace, w, st)
             : backward_t<float>(p, masks, n_masks, dimg, dmask, (char*)workspace, w, st);
}

size_t lb_mask_spectrum_bytes(lb_plan plan, int32_t n_masks) {
  if (!plan || n_masks < 1) return 0;
  const Plan& p = plan->p;
  return p.csz() * (size_t)n_masks * p.Wbm * p.Wbm;
}

lb_status lb_mask_spectrum(lb_plan plan, const void* masks, int32_t cplx, int32_t n_masks,
                           void* spectrum, void* workspace, size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!plan || !masks || !spectrum) return fail(LB_ERR_ARG, "null plan, masks or spectrum");
  if (n_masks < 1) return fail(LB_ERR_ARG, "n_masks must be >= 1");
  Plan& p = plan->p;
  Work w = layout(p, n_masks, false);
  if (!workspace || ws_bytes < w.total)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(w.total) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)workspace;
  if (p.precision == 64) {
    const double invN2 = 1.0 / ((double)p.N * (double)p.N);
    LB_HIP(band_analysis<double>(p.N, masks, cplx, (long long)p.N * p.N * (cplx ? 2 : 1), 0, 1,
                                 n_masks, p.nu0m, p.Wbm, invN2, (cx<double>*)(ws + w.RTm),
                                 (cx<double>*)spectrum, st));
  } else {
    const float invN2 = (float)(1.0 / ((double)p.N * (double)p.N));
    LB_HIP(band_analysis<float>(p.N, masks, cplx, (long long)p.N * p.N * (cplx ? 2 : 1), 0, 1,
                                n_masks, p.nu0m, p.Wbm, invN2, (cx<float>*)(ws + w.RTm),
                                (cx<float>*)spectrum, st));
  }
  return LB_OK;
}

lb_status lb_mask_spectrum_sigmoid(lb_plan plan, const void* theta, int32_t n_masks, double beta,
                                   void* spectrum, void* workspace, size_t ws_bytes,
                                   void* stream) {
  g_err.clear();
  if (!plan || !theta || !spectrum) return fail(LB_ERR_ARG, "null plan, theta or spectrum");
  if (n_masks < 1) return fail(LB_ERR_ARG, "n_masks must be >= 1");
  Plan& p = plan->p;
  Work w = layout(p, n_masks, false);
  if (!workspace || ws_bytes < w.total)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(w.total) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)workspace;
  const double invN2 = 1.0 / ((double)p.N * (double)p.N);
  const hipError_t e =
      p.precision == 64
          ? band_analysis_sigmoid<double>(p.N, (const double*)theta, n_masks, p.nu0m, p.Wbm,
                                          invN2, beta, (cx<double>*)(ws + w.RTm),
                                          (cx<double>*)spectrum, st)
          : band_analysis_sigmoid<float>(p.N, (const float*)theta, n_masks, p.nu0m, p.Wbm,
                                         (float)invN2, (float)beta, (cx<float>*)(ws + w.RTm),
                                         (cx<float>*)spectrum, st);
  if (e == hipErrorNotSupported)
    return fail(LB_ERR_UNSUPPORTED, "no fused sigmoid analysis for a " + std::to_string(p.N) +
                                        "-point grid (apply lb_mask_sigmoid, then "
                                        "lb_mask_spectrum)");
  LB_HIP(e);
  return LB_OK;
}

lb_status lb_forward_spectrum(lb_plan plan, const void* spectrum, int32_t n_masks, void* images,
                              void* peak, void* workspace, size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!plan || !spectrum || !images) return fail(LB_ERR_ARG, "null plan, spectrum or images");
  if (n_masks < 1) return fail(LB_ERR_ARG, "n_masks must be >= 1");
  Plan& p = plan->p;
  if (p.spp < 1) return fail(LB_ERR_ARG, "plan has no coherent systems (empty source)");
  Work w = layout(p, n_masks, false);
  if (!workspace || ws_bytes < w.total)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(w.total) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  return p.precision == 64
             ? forward_t<double>(p, nullptr, 0, n_masks, images, peak, (char*)workspace, w, st,
                                 spectrum)
             : forward_t<float>(p, nullptr, 0, n_masks, images, peak, (char*)workspace, w, st,
                                spectrum);
}

lb_status lb_backward_spectrum(lb_plan plan, const void* spectrum, int32_t n_masks,
                               const void* dimg, void* dmask, void* workspace, size_t ws_bytes,
                               void* stream) {
  g_err.clear();
  if (!plan || !spectrum || !dimg || !dmask) return fail(LB_ERR_ARG, "null argument");
  if (n_masks < 1) return fail(LB_ERR_ARG, "n_masks must be >= 1");
  Plan& p = plan->p;
  if (p.spp < 1) return fail(LB_ERR_ARG, "plan has no coherent systems (empty source)");
  Work w = layout(p, n_masks, true);
  if (!workspace || ws_bytes < w.total)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(w.total) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  return p.precision == 64
             ? backward_t<double>(p, nullptr, n_masks, dimg, dmask, (char*)workspace, w, st,
                                  spectrum)
             : backward_t<float>(p, nullptr, n_masks, dimg, dmask, (char*)workspace, w, st,
                                 spectrum);
}

namespace {

// Workspace of lb_forward_objective: the forward's, then the derivative's
// row spectra and the cost partials.
struct ObjWork {
  Work w;
  size_t rt, pc, pe, total;
};

ObjWork objective_layout(const Plan& p, int n_masks) {
  ObjWork o;
  o.w = layout(p, n_masks, false);
  Layout L;
  L.off = o.w.total;
  const int parts = synth_objective_part_count(p.N, o.w.n_img);
  o.rt = L.take(p.csz() * (size_t)o.w.n_img * band_ld((p.WbI + 1) / 2) * p.N);
  o.pc = L.take(sizeof(double) * (size_t)parts);
  o.pe = L.take(sizeof(int) * (size_t)parts);
  o.total = L.off;
  return o;
}

bool objective_supported(const Plan& p) {
  return !p.direct && p.spp > 0 && synth_objective_supported(p.N, -p.QI, p.WbI);
}

}  // namespace

size_t lb_forward_objective_workspace_bytes(lb_plan plan, int32_t n_masks) {
  if (!plan || n_masks < 1 || !objective_supported(plan->p)) return 0;
  return objective_layout(plan->p, n_masks).total;
}

size_t lb_sensitivity_band_bytes(lb_plan plan, int32_t n_masks) {
  if (!plan || n_masks < 1 || plan->p.direct) return 0;
  const Plan& p = plan->p;
  return p.csz() * (size_t)n_masks * p.n_planes * p.WbI * p.WbI;
}

lb_status lb_forward_objective(lb_plan plan, const void* spectrum, int32_t n_masks,
                               const void* target, double scale, double level, double steepness,
                               double sign, void* cost, void* total, int32_t* mismatch,
                               void* sens_band, void* images, void* d_images, void* workspace,
                               size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!plan || !spectrum || !target || !sens_band)
    return fail(LB_ERR_ARG, "null plan, spectrum, target or sens_band");
  if (n_masks < 1) return fail(LB_ERR_ARG, "n_masks must be >= 1");
  Plan& p = plan->p;
  if (p.spp < 1) return fail(LB_ERR_ARG, "plan has no coherent systems (empty source)");
  auto al = [](const void* q) { return q == nullptr || ((uintptr_t)q & 15) == 0; };
  if (!objective_supported(p))
    return fail(LB_ERR_UNSUPPORTED,
                "the fused objective needs a coarse grid and an image band inside |nu| < N/16");
  if (!al(target) || !al(images) || !al(d_images))
    return fail(LB_ERR_UNSUPPORTED, "target, images and d_images must be 16-byte aligned");
  const ObjWork o = objective_layout(p, n_masks);
  if (!workspace || ws_bytes < o.total)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(o.total) + " bytes");
  const double a = steepness / std::max(std::fabs(level), 1e-12);  // as lb_resist_objective
  const ObjArgs obj{target, scale, level, a, sign, cost, total, mismatch, sens_band, d_images,
                    o.rt, o.pc, o.pe};
  hipStream_t st = (hipStream_t)stream;
  return p.precision == 64
             ? forward_t<double>(p, nullptr, 0, n_masks, images, nullptr, (char*)workspace, o.w,
                                 st, spectrum, true, &obj)
             : forward_t<float>(p, nullptr, 0, n_masks, images, nullptr, (char*)workspace, o.w,
                                st, spectrum, true, &obj);
}

lb_status lb_backward_sensitivity_band(lb_plan plan, const void* spectrum, int32_t n_masks,
                                       const void* sens_band, void* dmask, void* workspace,
                                       size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!plan || !spectrum || !sens_band || !dmask) return fail(LB_ERR_ARG, "null argument");
  if (n_masks < 1) return fail(LB_ERR_ARG, "n_masks must be >= 1");
  Plan& p = plan->p;
  if (p.spp < 1) return fail(LB_ERR_ARG, "plan has no coherent systems (empty source)");
  if (p.direct)
    return fail(LB_ERR_UNSUPPORTED, "the sensitivity band form needs a coarse grid");
  Work w = layout(p, n_masks, true);
  if (!workspace || ws_bytes < w.total)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(w.total) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  return p.precision == 64
             ? backward_t<double>(p, nullptr, n_masks, nullptr, dmask, (char*)workspace, w, st,
                                  spectrum, false, sens_band)
             : backward_t<float>(p, nullptr, n_masks, nullptr, dmask, (char*)workspace, w, st,
                                 spectrum, false, sens_band);
}

lb_status lb_backward_objective_theta(lb_plan plan, const void* spectrum, int32_t n_masks,
                                      const void* sens_band, const void* theta, double beta,
                                      const void* g, void* grad_theta, void* workspace,
                                      size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!plan || !spectrum || !sens_band || !theta || !grad_theta)
    return fail(LB_ERR_ARG, "null argument");
  if (n_masks < 1) return fail(LB_ERR_ARG, "n_masks must be >= 1");
  Plan& p = plan->p;
  if (p.spp < 1) return fail(LB_ERR_ARG, "plan has no coherent systems (empty source)");
  if (p.direct || !synth_objective_supported(p.N, p.nu0m, p.Wbm))
    return fail(LB_ERR_UNSUPPORTED,
                "the theta gradient pass needs a coarse grid and a mask band inside |nu| < N/16");
  if (((uintptr_t)theta & 15) || ((uintptr_t)grad_theta & 15))
    return fail(LB_ERR_UNSUPPORTED, "theta and grad_theta must be 16-byte aligned");
  Work w = layout(p, n_masks, true);
  if (!workspace || ws_bytes < w.total)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(w.total) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  if (p.precision == 64) {
    const ThetaEpi<double> te{(const double*)theta, beta, (const double*)g};
    return backward_t<double>(p, nullptr, n_masks, nullptr, grad_theta, (char*)workspace, w, st,
                              spectrum, false, sens_band, &te);
  }
  const ThetaEpi<float> te{(const float*)theta, (float)beta, (const float*)g};
  return backward_t<float>(p, nullptr, n_masks, nullptr, grad_theta, (char*)workspace, w, st,
                           spectrum, false, sens_band, &te);
}

lb_status lb_backward_spectrum_complex(lb_plan plan, const void* spectrum, int32_t n_masks,
                                       const void* dimg, void* dmask, void* workspace,
                                       size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!plan || !spectrum || !dimg || !dmask) return fail(LB_ERR_ARG, "null argument");
  if (n_masks < 1) return fail(LB_ERR_ARG, "n_masks must be >= 1");
  Plan& p = plan->p;
  if (p.spp < 1) return fail(LB_ERR_ARG, "plan has no coherent systems (empty source)");
  Work w = layout(p, n_masks, true);
  if (!workspace || ws_bytes < w.total)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(w.total) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  return p.precision == 64
             ? backward_t<double>(p, nullptr, n_masks, dimg, dmask, (char*)workspace, w, st,
                                  spectrum, true)
             : backward_t<float>(p, nullptr, n_masks, dimg, dmask, (char*)workspace, w, st,
                                 spectrum, true);
}

lb_status lb_socs_kernels(lb_plan plan, void* kernels, void* stream) {
  g_err.clear();
  if (!plan || !kernels) return fail(LB_ERR_ARG, "null argument");
  Plan& p = plan->p;
  if (!p.socs) return fail(LB_ERR_ARG, "not a SOCS plan");
  if (p.K == 0) return LB_OK;
  hipStream_t st = (hipStream_t)stream;
  LB_HIP(fill<double>((double*)kernels, 0.0, 2LL * p.K * p.N * p.N, st));
  LB_HIP(scatter_band<double>((const cx<double>*)p.phi128, p.K, p.Wbm, p.nu0m, p.N,
                              (cx<double>*)kernels, st));
  return LB_OK;
}

lb_status lb_batch_create(const lb_plan* plans, int32_t n_plans, lb_batch* out) {
  g_err.clear();
  if (!plans || n_plans < 1 || !out) return fail(LB_ERR_ARG, "need >= 1 plan and an output");
  *out = nullptr;
  for (int32_t i = 0; i < n_plans; ++i)
    if (!plans[i]) return fail(LB_ERR_ARG, "null plan in batch");
  lb_batch b = new lb_batch_s();
  b->plans.assign(plans, plans + n_plans);
  const Plan& p0 = plans[0]->p;
  b->precision = p0.precision;
  b->M = p0.M;
  b->cus = p0.cus;
  long long total_sys = 0;
  for (lb_plan h : b->plans) {
    const Plan& p = h->p;
    if (p.precision != b->precision || p.M != b->M) {
      delete b;
      return fail(LB_ERR_ARG, "batched plans must share precision and coarse grid side "
                              "(create them with the same coarse_m)");
    }
    if (p.direct) {
      delete b;
      return fail(LB_ERR_UNSUPPORTED, "batched plans need a coarse grid smaller than the clip");
    }
    if (p.spp < 1) {
      delete b;
      return fail(LB_ERR_ARG, "plan has no coherent systems");
    }
    b->Wmax = std::max(b->Wmax, p.W);
    total_sys += (long long)p.spp * p.n_planes;
  }
  bool got_f = false, got_a = false;
  for (const FieldConfig& c : field_configs()) {
    if (c.M != b->M || c.csplit == 0) continue;  // batches take the single-launch kernels
    if (!got_f && field_smem_bytes(b->precision, c, b->Wmax, false) <= kLdsLimit) {
      b->fwd = c;
      got_f = true;
    }
    if (!got_a && field_smem_bytes(b->precision, c, b->Wmax, true) <= kLdsLimit) {
      b->adj = c;
      got_a = true;
    }
  }
  if (!got_f || !got_a) {
    delete b;
    return fail(LB_ERR_UNSUPPORTED, "no coherent-system kernel shape fits the widest clip");
  }
  b->chunk = (int)std::max<long long>(1, (total_sys + 4LL * b->cus - 1) / (4LL * b->cus));
  // group the clips by band geometry
  auto same = [](const Plan& a, const Plan& c) {
    return a.N == c.N && a.nu0m == c.nu0m && a.Wbm == c.Wbm && a.wrapm == c.wrapm && a.M == c.M &&
           a.QI == c.QI && a.WbI == c.WbI && a.n_planes == c.n_planes && a.spp == c.spp;
  };
  for (int c = 0; c < n_plans; ++c) {
    BatchGroup* g = nullptr;
    for (BatchGroup& h : b->groups)
      if (same(b->plans[h.members[0]]->p, b->plans[c]->p)) g = &h;
    if (!g) {
      b->groups.emplace_back();
      g = &b->groups.back();
    }
    g->members.push_back(c);
  }
  std::vector<ItemDesc> itf, itb;
  size_t off_f = 0, off_b = 0;
  for (BatchGroup& g : b->groups) {
    const Plan& p = b->plans[g.members[0]]->p;
    const int cnt = (int)g.members.size();
    g.wf = layout(p, cnt, false, b->chunk, &b->fwd);
    g.wb = layout(p, cnt, true, b->chunk, &b->adj);
    const Work& wf = g.wf;
    const Work& wb = g.wb;
    g.base_f = off_f;
    g.base_b = off_b;
    const size_t R = p.rsz(), C = p.csz();
    const size_t nn = (size_t)p.N * p.N;
    g.stage_f = off_f + (wf.total + 255) / 256 * 256;
    g.stage_b = off_b + (wb.total + 255) / 256 * 256;
    for (int k = 0; k < cnt; ++k) {
      const Plan& q = b->plans[g.members[k]]->p;
      for (int pl = 0; pl < q.n_planes; ++pl)
        for (int ch = 0; ch < wf.n_chunks; ++ch) {
          const long long gi = (long long)k * q.n_planes + pl;  // image in the group
          const long long item = gi * wf.n_chunks + ch;
          ItemDesc d{};
          d.Wb = q.Wbm;
          d.wrapN = q.wrapm;
          d.H = q.H;
          d.W = q.W;
          d.box = q.box;
          d.wsys = q.wsys;
          d.s0 = pl * q.spp + ch * wf.chunk;
          d.s1 = std::min(d.s0 + wf.chunk, (pl + 1) * q.spp);
          d.X_off = (long long)(off_f + wf.Xm + (size_t)k * q.Wbm * q.Wbm * C);
          d.part_off = (long long)(off_f + wf.partial + item * q.M * q.M * R);
          d.gM_off = -1;
          d.gX_off = -1;
          itf.push_back(d);
          d.X_off = (long long)(off_b + wb.Xm + (size_t)k * q.Wbm * q.Wbm * C);
          d.part_off = -1;
          d.gM_off = (long long)(off_b + wb.gM + (size_t)gi * q.M * q.M * R);
          d.gX_off = (long long)(off_b + wb.gX + item * q.Wbm * q.Wbm * C);
          itb.push_back(d);
        }
    }
    // staging: masks, then images (forward) or image gradients and mask gradients (backward)
    const size_t masks_b = (R * cnt * nn + 255) / 256 * 256;
    const size_t imgs_b = (R * cnt * p.n_planes * nn + 255) / 256 * 256;
    const size_t peak_b = (R * cnt * p.n_planes + 255) / 256 * 256;
    off_f = g.stage_f + masks_b + imgs_b + peak_b;
    off_b = g.stage_b + masks_b + imgs_b + masks_b;
  }
  b->ws_f = off_f;
  b->ws_b = off_b;
  b->n_items_f = (int)itf.size();
  b->n_items_b = (int)itb.size();
  hipError_t e = hipMalloc(&b->items_f, sizeof(ItemDesc) * itf.size());
  if (e == hipSuccess) e = hipMalloc(&b->items_b, sizeof(ItemDesc) * itb.size());
  if (e == hipSuccess)
    e = hipMemcpy(b->items_f, itf.data(), sizeof(ItemDesc) * itf.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->items_b, itb.data(), sizeof(ItemDesc) * itb.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete b;
    return fail(e == hipErrorOutOfMemory ? LB_ERR_OOM : LB_ERR_HIP,
                std::string("batch tables: ") + hipGetErrorString(e));
  }
  if (b->groups.size() > 1) {
    const size_t S = std::min(b->groups.size(), kBatchSideStreams);
    e = hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming);
    for (size_t k = 0; k < S && e == hipSuccess; ++k) {
      hipStream_t sk = nullptr;
      hipEvent_t ek = nullptr;
      e = hipStreamCreateWithFlags(&sk, hipStreamNonBlocking);
      if (e == hipSuccess) b->side.push_back(sk);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&ek, hipEventDisableTiming);
      if (e == hipSuccess) b->joins.push_back(ek);
    }
    if (e != hipSuccess) {
      delete b;
      return fail(LB_ERR_HIP, std::string("batch streams: ") + hipGetErrorString(e));
    }
  }
  *out = b;
  return LB_OK;
}

lb_status lb_batch_set_shape(lb_batch batch, int32_t adjoint, int32_t block_threads,
                             int32_t column_splits) {
  g_err.clear();
  if (!batch) return fail(LB_ERR_ARG, "null batch");
  FieldConfig fc{batch->M, block_threads, column_splits, 0};
  if (column_splits < 1 || !field_config_supported(batch->precision, fc))
    return fail(LB_ERR_UNSUPPORTED, "no single-launch coherent-system kernel of shape M=" +
                                        std::to_string(batch->M) + " threads=" +
                                        std::to_string(block_threads) + " splits=" +
                                        std::to_string(column_splits));
  if (field_smem_bytes(batch->precision, fc, batch->Wmax, adjoint != 0) > kLdsLimit)
    return fail(LB_ERR_UNSUPPORTED, "shape does not fit in LDS for the widest clip");
  (adjoint ? batch->adj : batch->fwd) = fc;
  return LB_OK;
}

lb_status lb_batch_destroy(lb_batch batch) {
  delete batch;
  return LB_OK;
}

size_t lb_batch_workspace_bytes(lb_batch batch, int32_t backward) {
  if (!batch) return 0;
  return backward ? batch->ws_b : batch->ws_f;
}

static int batch_grid(const lb_batch_s& b, const FieldConfig& fc, int n_items, bool adjoint) {
  const size_t sm = field_smem_bytes(b.precision, fc, b.Wmax, adjoint);
  int occ = (int)std::max<size_t>(1, kLdsLimit / std::max<size_t>(sm, 1));
  occ = std::min(occ, std::max(1, 2048 / fc.nt));
  return std::max(1, std::min(n_items, b.cus * occ));
}

}  // extern "C"

// Whether the clips of a group sit consecutively in the caller's memory at
// `bytes` per clip (then the group runs on them in place).
static bool consecutive(const void* const* ptr, const std::vector<int>& members, size_t bytes) {
  for (size_t k = 1; k < members.size(); ++k)
    if ((const char*)ptr[members[k]] != (const char*)ptr[members[0]] + k * bytes) return false;
  return true;
}

static hipError_t gather(char* dst, const void* const* src, const std::vector<int>& members,
                         size_t bytes, hipStream_t st) {
  for (size_t k = 0; k < members.size(); ++k) {
    hipError_t e = hipMemcpyAsync(dst + k * bytes, src[members[k]], bytes,
                                  hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

static hipError_t scatter(void* const* dst, const char* src, const std::vector<int>& members,
                          size_t bytes, hipStream_t st) {
  for (size_t k = 0; k < members.size(); ++k) {
    hipError_t e = hipMemcpyAsync(dst[members[k]], src + k * bytes, bytes,
                                  hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename T>
static lb_status batch_forward_t(lb_batch_s& b, const void* const* masks, void* const* images,
                                 void* const* peaks, char* ws, hipStream_t st0) {
  lb_status rc0 = for_groups(b, st0, [&](BatchGroup& g, hipStream_t st) -> lb_status {
    const Plan& p = b.plans[g.members[0]]->p;
    const size_t mb = sizeof(T) * (size_t)p.N * p.N;
    const void* m0 = masks[g.members[0]];
    if (!consecutive(masks, g.members, mb)) {
      LB_HIP(gather(ws + g.stage_f, masks, g.members, mb, st));
      m0 = ws + g.stage_f;
    }
    return fwd_pre<T>(b.plans[g.members[0]]->p, m0, 0, (int)g.members.size(), ws + g.base_f,
                      g.wf, st);
  });
  if (rc0 != LB_OK) return rc0;
  hipStream_t st = st0;
  FieldConfig fc = b.fwd;
  fc.grid = batch_grid(b, fc, b.n_items_f, false);
  int* q = (int*)(ws + b.groups[0].base_f + b.groups[0].wf.queue);
  LB_HIP(fill<float>((float*)q, 0.0f, 1, st));
  LB_HIP(field_multi(b.precision, fc, b.items_f, b.n_items_f, b.Wmax, ws, q, false, st));
  return for_groups(b, st0, [&](BatchGroup& g, hipStream_t sg) -> lb_status {
    Plan& p = b.plans[g.members[0]]->p;
    const int cnt = (int)g.members.size();
    const size_t nn = (size_t)p.N * p.N;
    const size_t ib = sizeof(T) * p.n_planes * nn, pb = sizeof(T) * p.n_planes;
    char* stage_img = ws + g.stage_f + (sizeof(T) * cnt * nn + 255) / 256 * 256;
    char* stage_pk = stage_img + (sizeof(T) * cnt * p.n_planes * nn + 255) / 256 * 256;
    const bool img_in_place = consecutive(images, g.members, ib);
    const bool pk_in_place = !peaks || consecutive(peaks, g.members, pb);
    void* out = img_in_place ? images[g.members[0]] : (void*)stage_img;
    void* pk = peaks ? (pk_in_place ? peaks[g.members[0]] : (void*)stage_pk) : nullptr;
    lb_status rc = fwd_post<T>(p, out, pk, ws + g.base_f, g.wf, sg);
    if (rc != LB_OK) return rc;
    if (!img_in_place) LB_HIP(scatter(images, stage_img, g.members, ib, sg));
    if (peaks && !pk_in_place) LB_HIP(scatter(peaks, stage_pk, g.members, pb, sg));
    return LB_OK;
  });
}

// fws (nullable): the workspace of the lb_batch_forward call that imaged these
// masks; their band spectra are copied from it instead of recomputed.
template <typename T>
static lb_status batch_backward_t(lb_batch_s& b, const void* const* masks,
                                  const void* const* dimg, void* const* dmask, char* ws,
                                  hipStream_t st0, const char* fws = nullptr) {
  lb_status rc0 = for_groups(b, st0, [&](BatchGroup& g, hipStream_t st) -> lb_status {
    const Plan& p = b.plans[g.members[0]]->p;
    const int cnt = (int)g.members.size();
    const size_t nn = (size_t)p.N * p.N;
    const size_t mb = sizeof(T) * nn, ib = sizeof(T) * p.n_planes * nn;
    char* stage_m = ws + g.stage_b;
    char* stage_g = stage_m + (sizeof(T) * cnt * nn + 255) / 256 * 256;
    const void* m0 = masks[g.members[0]];
    if (fws) {
      LB_HIP(hipMemcpyAsync(ws + g.base_b + g.wb.Xm, fws + g.base_f + g.wf.Xm,
                            p.csz() * (size_t)cnt * p.Wbm * p.Wbm, hipMemcpyDeviceToDevice, st));
    } else if (!consecutive(masks, g.members, mb)) {
      LB_HIP(gather(stage_m, masks, g.members, mb, st));
      m0 = stage_m;
    }
    const void* g0 = dimg[g.members[0]];
    if (!consecutive(dimg, g.members, ib)) {
      LB_HIP(gather(stage_g, dimg, g.members, ib, st));
      g0 = stage_g;
    }
    const T* gM = nullptr;
    return bwd_pre<T>(b.plans[g.members[0]]->p, m0, cnt, g0, ws + g.base_b, g.wb, &gM, st,
                      fws != nullptr);
  });
  if (rc0 != LB_OK) return rc0;
  hipStream_t st = st0;
  FieldConfig fc = b.adj;
  fc.grid = batch_grid(b, fc, b.n_items_b, true);
  int* q = (int*)(ws + b.groups[0].base_b + b.groups[0].wb.queue);
  LB_HIP(fill<float>((float*)q, 0.0f, 1, st));
  LB_HIP(field_multi(b.precision, fc, b.items_b, b.n_items_b, b.Wmax, ws, q, true, st));
  return for_groups(b, st0, [&](BatchGroup& g, hipStream_t sg) -> lb_status {
    Plan& p = b.plans[g.members[0]]->p;
    const int cnt = (int)g.members.size();
    const size_t nn = (size_t)p.N * p.N;
    const size_t mb = sizeof(T) * nn;
    char* stage_d = ws + g.stage_b + (sizeof(T) * cnt * nn + 255) / 256 * 256 +
                    (sizeof(T) * cnt * p.n_planes * nn + 255) / 256 * 256;
    const bool in_place = consecutive(dmask, g.members, mb);
    void* out = in_place ? dmask[g.members[0]] : (void*)stage_d;
    lb_status rc = bwd_post<T>(p, cnt, out, ws + g.base_b, g.wb, sg);
    if (rc != LB_OK) return rc;
    if (!in_place) LB_HIP(scatter(dmask, stage_d, g.members, mb, sg));
    return LB_OK;
  });
}

extern "C" {

lb_status lb_batch_forward(lb_batch batch, const void* const* masks, void* const* images,
                           void* const* peaks, void* workspace, size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!batch || !masks || !images) return fail(LB_ERR_ARG, "null argument");
  for (size_t c = 0; c < batch->plans.size(); ++c)
    if (!masks[c] || !images[c]) return fail(LB_ERR_ARG, "null mask or image pointer");
  if (!workspace || ws_bytes < batch->ws_f)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(batch->ws_f) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  return batch->precision == 64
             ? batch_forward_t<double>(*batch, masks, images, peaks, (char*)workspace, st)
             : batch_forward_t<float>(*batch, masks, images, peaks, (char*)workspace, st);
}

lb_status lb_batch_backward_reuse(lb_batch batch, const void* const* d_images,
                                  void* const* d_masks, void* workspace, size_t ws_bytes,
                                  const void* forward_workspace, void* stream) {
  g_err.clear();
  if (!batch || !d_images || !d_masks || !forward_workspace) return fail(LB_ERR_ARG, "null argument");
  for (size_t c = 0; c < batch->plans.size(); ++c)
    if (!d_images[c] || !d_masks[c]) return fail(LB_ERR_ARG, "null pointer");
  if (!workspace || ws_bytes < batch->ws_b)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(batch->ws_b) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  std::vector<const void*> none(batch->plans.size(), forward_workspace);  // masks unused
  return batch->precision == 64
             ? batch_backward_t<double>(*batch, none.data(), d_images, d_masks, (char*)workspace,
                                        st, (const char*)forward_workspace)
             : batch_backward_t<float>(*batch, none.data(), d_images, d_masks, (char*)workspace,
                                       st, (const char*)forward_workspace);
}

lb_status lb_batch_backward(lb_batch batch, const void* const* masks,
                            const void* const* d_images, void* const* d_masks, void* workspace,
                            size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!batch || !masks || !d_images || !d_masks) return fail(LB_ERR_ARG, "null argument");
  for (size_t c = 0; c < batch->plans.size(); ++c)
    if (!masks[c] || !d_images[c] || !d_masks[c]) return fail(LB_ERR_ARG, "null pointer");
  if (!workspace || ws_bytes < batch->ws_b)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(batch->ws_b) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  return batch->precision == 64
             ? batch_backward_t<double>(*batch, masks, d_images, d_masks, (char*)workspace, st)
             : batch_backward_t<float>(*batch, masks, d_images, d_masks, (char*)workspace, st);
}

size_t lb_dense_workspace_bytes(lb_plan plan, int32_t batch) {
  if (!plan || batch < 1) return 0;
  const Plan& p = plan->p;
  Layout L;
  L.take(p.csz() * (size_t)p.N * p.N);      // whole-grid mask spectrum
  L.take(p.csz() * (size_t)p.N * p.N);      // analysis scratch
  L.take(p.csz() * (size_t)batch * p.N * p.N);  // dense spectra
  L.take(p.csz() * (size_t)batch * p.N * p.N);  // synthesis scratch
  L.take(p.csz() * (size_t)batch * p.N * p.N);  // fields
  L.take(p.csz() * (size_t)p.N * p.N);      // adjoint: spectrum-gradient accumulator
  return L.off;
}

}  // extern "C"

template <typename T>
static lb_status dense_t(Plan& p, const void* mask, int plane, int batch, void* image, char* ws,
                         hipStream_t st) {
  const int N = p.N;
  Layout L;
  cx<T>* Xf = (cx<T>*)(ws + L.take(p.csz() * (size_t)N * N));
  cx<T>* RT = (cx<T>*)(ws + L.take(p.csz() * (size_t)N * N));
  cx<T>* spec = (cx<T>*)(ws + L.take(p.csz() * (size_t)batch * N * N));
  cx<T>* TT = (cx<T>*)(ws + L.take(p.csz() * (size_t)batch * N * N));
  cx<T>* E = (cx<T>*)(ws + L.take(p.csz() * (size_t)batch * N * N));
  const T invN2 = (T)(1.0 / ((double)N * (double)N));
  // the mask spectrum on the whole grid (provider.py:177-178), scaled by 1/N^2
  LB_HIP(band_analysis<T>(N, mask, 0, (long long)N * N, 0, 1, 1, -(N / 2), N, invN2, RT, Xf, st));
  LB_HIP(fill<T>((T*)image, (T)0, (long long)N * N, st));
  const int s_begin = plane * p.spp, s_end = (plane + 1) * p.spp;
  for (int j0 = s_begin; j0 < s_end; j0 += batch) {
    const int ns = std::min(batch, s_end - j0);
    LB_HIP(dense_spectra<T>(Xf, N, (const cx<T>*)p.H, p.W, p.box, j0, ns, p.nu0m, spec, st));
    LB_HIP(band_synthesis<T>(N, spec, ns, -(N / 2), N, 0, (T)1, 1, TT, (T*)E, (T*)nullptr, st));
    LB_HIP(accum_norm2<T>(E, ns, (long long)N * N, (const T*)p.wsys, j0, (T*)image, st));
  }
  return LB_OK;
}

// The dense adjoint, per source point as the reference's adjoint computes it
// (opc/ilt.py:235-266 restated for Abbe systems): E_s on the whole grid,
// G_s = fft2(g E_s), A += w_s conj(H_s) G_s, d mask = 2 Re ifft2(A).
template <typename T>
static lb_status dense_bwd_t(Plan& p, const void* mask, int plane, int batch, const void* dimg,
                             void* dmask, char* ws, hipStream_t st) {
  const int N = p.N;
  const long long NN = (long long)N * N;
  Layout L;
  cx<T>* Xf = (cx<T>*)(ws + L.take(p.csz() * (size_t)NN));
  cx<T>* RT = (cx<T>*)(ws + L.take(p.csz() * (size_t)NN));
  cx<T>* spec = (cx<T>*)(ws + L.take(p.csz() * (size_t)batch * NN));
  cx<T>* TT = (cx<T>*)(ws + L.take(p.csz() * (size_t)batch * NN));
  cx<T>* E = (cx<T>*)(ws + L.take(p.csz() * (size_t)batch * NN));
  cx<T>* A = (cx<T>*)(ws + L.take(p.csz() * (size_t)NN));
  const T invN2 = (T)(1.0 / ((double)N * (double)N));
  LB_HIP(band_analysis<T>(N, mask, 0, NN, 0, 1, 1, -(N / 2), N, invN2, RT, Xf, st));
  LB_HIP(fill<T>((T*)A, (T)0, 2 * NN, st));
  const int s_begin = plane * p.spp, s_end = (plane + 1) * p.spp;
  for (int j0 = s_begin; j0 < s_end; j0 += batch) {
    const int ns = std::min(batch, s_end - j0);
    LB_HIP(dense_spectra<T>(Xf, N, (const cx<T>*)p.H, p.W, p.box, j0, ns, p.nu0m, spec, st));
    LB_HIP(band_synthesis<T>(N, spec, ns, -(N / 2), N, 0, (T)1, 1, TT, (T*)E, (T*)nullptr, st));
    LB_HIP(mul_real<T>(E, ns, NN, (const T*)dimg, st));
    LB_HIP(band_analysis<T>(N, E, 1, 2 * NN, 0, 1, ns, -(N / 2), N, (T)1, TT, spec, st));
    LB_HIP(dense_adjoint_accum<T>(spec, N, (const cx<T>*)p.H, p.W, p.box, (const T*)p.wsys, j0,
                                  ns, p.nu0m, A, st));
  }
  LB_HIP(band_synthesis<T>(N, A, 1, -(N / 2), N, 1, (T)2 * invN2, 0, TT, (T*)dmask, (T*)nullptr,
                           st));
  return LB_OK;
}

extern "C" {

lb_status lb_backward_dense(lb_plan plan, const void* mask, int32_t plane, int32_t batch,
                            const void* d_image, void* d_mask, void* workspace, size_t ws_bytes,
                            void* stream) {
  g_err.clear();
  if (!plan || !mask || !d_image || !d_mask) return fail(LB_ERR_ARG, "null argument");
  Plan& p = plan->p;
  if (p.socs) return fail(LB_ERR_UNSUPPORTED, "the dense adjoint is for Abbe plans");
  if (plane < 0 || plane >= p.n_planes) return fail(LB_ERR_ARG, "plane out of range");
  if (batch < 1) return fail(LB_ERR_ARG, "batch must be >= 1");
  const size_t need = lb_dense_workspace_bytes(plan, batch);
  if (!workspace || ws_bytes < need)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(need) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  return p.precision == 64 ? dense_bwd_t<double>(p, mask, plane, batch, d_image, d_mask,
                                                 (char*)workspace, st)
                           : dense_bwd_t<float>(p, mask, plane, batch, d_image, d_mask,
                                                (char*)workspace, st);
}

size_t lb_gaussian_blur_workspace_bytes(int32_t n_img, int32_t n, double sigma,
                                        int32_t precision) {
  if (n_img < 1 || n < 1 || !(sigma >= 0)) return 0;
  const size_t R = (size_t)gauss_radius(sigma);
  const size_t t = (precision == 64 ? 8 : 4) * (size_t)n_img * n * n;
  return ((t + 255) / 256) * 256 + sizeof(double) * (2 * R + 1);
}

lb_status lb_gaussian_blur(const void* in, void* out, int32_t n_img, int32_t n, double sigma,
                           int32_t precision, void* workspace, size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!in || !out) return fail(LB_ERR_ARG, "null argument");
  if (n_img < 1 || n < 1) return fail(LB_ERR_ARG, "n_img and n must be >= 1");
  if (precision != 32 && precision != 64) return fail(LB_ERR_ARG, "precision must be 32 or 64");
  if (!(sigma >= 0) || sigma > 1e4) return fail(LB_ERR_ARG, "sigma must be in [0, 1e4]");
  hipStream_t st = (hipStream_t)stream;
  const size_t bytes = (precision == 64 ? 8 : 4) * (size_t)n_img * n * n;
  if (sigma < 1e-3) {  // the reference returns its input unblurred below 1e-3 px
    if (in != out) LB_HIP(hipMemcpyAsync(out, in, bytes, hipMemcpyDeviceToDevice, st));
    return LB_OK;
  }
  const size_t need = lb_gaussian_blur_workspace_bytes(n_img, n, sigma, precision);
  if (!workspace || ws_bytes < need)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(need) + " bytes");
  char* ws = (char*)workspace;
  double* w = (double*)(ws + ((bytes + 255) / 256) * 256);
  if (precision == 64)
    LB_HIP(gaussian_blur<double>((const double*)in, (double*)out, n_img, n, sigma, (double*)ws, w,
                                 st));
  else
    LB_HIP(gaussian_blur<float>((const float*)in, (float*)out, n_img, n, sigma, (float*)ws, w,
                                st));
  return LB_OK;
}

size_t lb_resist_objective_workspace_bytes(int32_t n_img, int64_t npix) {
  if (n_img < 1 || npix < 1) return 0;
  const size_t k = (size_t)n_img * resist_blocks(npix);
  return ((k * sizeof(double) + n_img * sizeof(double) + 255) / 256) * 256 + k * sizeof(int);
}

lb_status lb_resist_objective(const void* intensity, const void* target, int32_t n_img,
                              int64_t npix, double scale, double level, double steepness,
                              double sign, void* cost, void* total, void* d_intensity,
                              void* signed_field, int32_t* mismatch, int32_t precision,
                              void* workspace, size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!intensity || !target) return fail(LB_ERR_ARG, "null argument");
  if (n_img < 1 || npix < 1) return fail(LB_ERR_ARG, "n_img and npix must be >= 1");
  if (precision != 32 && precision != 64) return fail(LB_ERR_ARG, "precision must be 32 or 64");
  const size_t need = lb_resist_objective_workspace_bytes(n_img, npix);
  if (!workspace || ws_bytes < need)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(need) + " bytes");
  const double a = steepness / std::max(std::fabs(level), 1e-12);
  const size_t k = (size_t)n_img * resist_blocks(npix);
  double* pc = (double*)workspace;  // k block partials, then n_img image costs
  int* pe = (int*)((char*)workspace + ((k * sizeof(double) + n_img * sizeof(double) + 255) / 256) *
                                          256);
  hipStream_t st = (hipStream_t)stream;
  if (precision == 64)
    LB_HIP(resist_objective<double>((const double*)intensity, (const double*)target, n_img, npix,
                                    scale, level, a, sign, (double*)cost, (double*)total,
                                    (double*)d_intensity, (double*)signed_field, mismatch, pc, pe,
                                    st));
  else
    LB_HIP(resist_objective<float>((const float*)intensity, (const float*)target, n_img, npix,
                                   (float)scale, (float)level, (float)a, (float)sign,
                                   (float*)cost, (float*)total, (float*)d_intensity,
                                   (float*)signed_field, mismatch, pc, pe, st));
  return LB_OK;
}

lb_status lb_mask_penalties(const void* mask, int32_t n_img, int32_t ny, int32_t nx,
                            double weight_tv, double weight_discrete, void* cost, void* total,
                            void* grad, int32_t precision, void* workspace, size_t ws_bytes,
                            void* stream) {
  g_err.clear();
  if (!mask) return fail(LB_ERR_ARG, "null argument");
  if (n_img < 1 || ny < 1 || nx < 1) return fail(LB_ERR_ARG, "n_img, ny and nx must be >= 1");
  if (precision != 32 && precision != 64) return fail(LB_ERR_ARG, "precision must be 32 or 64");
  const long long npix = (long long)ny * nx;
  const size_t need = lb_resist_objective_workspace_bytes(n_img, npix);
  if (!workspace || ws_bytes < need)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(need) + " bytes");
  const size_t k = (size_t)n_img * resist_blocks(npix);
  double* pc = (double*)workspace;
  int* pz = (int*)((char*)workspace + ((k * sizeof(double) + n_img * sizeof(double) + 255) / 256) *
                                          256);
  hipStream_t st = (hipStream_t)stream;
  if (precision == 64)
    LB_HIP(mask_penalties<double>((const double*)mask, n_img, ny, nx, weight_tv, weight_discrete,
                                  (double*)cost, (double*)total, (double*)grad, pc, pz, st));
  else
    LB_HIP(mask_penalties<float>((const float*)mask, n_img, ny, nx, (float)weight_tv,
                                 (float)weight_discrete, (float*)cost, (float*)total, (float*)grad,
                                 pc, pz, st));
  return LB_OK;
}

static lb_status adam_any(void* theta, void* mom1, void* mom2, const void* grad, int64_t n,
                          double c1, double c2, const double* pc, double step, double beta1,
                          double beta2, double eps, int32_t precision, void* stream) {
  g_err.clear();
  if (!theta || !mom1 || !mom2 || !grad) return fail(LB_ERR_ARG, "null argument");
  if (n < 0) return fail(LB_ERR_ARG, "n must be >= 0");
  if (precision != 32 && precision != 64) return fail(LB_ERR_ARG, "precision must be 32 or 64");
  hipStream_t st = (hipStream_t)stream;
  if (precision == 64)
    LB_HIP(adam_step<double>((double*)theta, (double*)mom1, (double*)mom2, (const double*)grad, n,
                             beta1, beta2, 1.0 - beta1, 1.0 - beta2, eps, step, c1, c2, pc, st));
  else
    LB_HIP(adam_step<float>((float*)theta, (float*)mom1, (float*)mom2, (const float*)grad, n,
                            (float)beta1, (float)beta2, (float)(1.0 - beta1), (float)(1.0 - beta2),
                            (float)eps, (float)step, (float)c1, (float)c2, pc, st));
  return LB_OK;
}

lb_status lb_adam_step(void* theta, void* mom1, void* mom2, const void* grad, int64_t n,
                       int32_t iteration, double step, double beta1, double beta2, double eps,
                       int32_t precision, void* stream) {
  if (iteration < 1) return fail(LB_ERR_ARG, "iteration counts from 1");
  const double c1 = 1.0 - std::pow(beta1, iteration), c2 = 1.0 - std::pow(beta2, iteration);
  return adam_any(theta, mom1, mom2, grad, n, c1, c2, nullptr, step, beta1, beta2, eps, precision,
                  stream);
}

lb_status lb_adam_step_dev(void* theta, void* mom1, void* mom2, const void* grad, int64_t n,
                           const double* corrections, double step, double beta1, double beta2,
                           double eps, int32_t precision, void* stream) {
  if (!corrections) return fail(LB_ERR_ARG, "null corrections");
  return adam_any(theta, mom1, mom2, grad, n, 0.0, 0.0, corrections, step, beta1, beta2, eps,
                  precision, stream);
}

lb_status lb_volume_interp(const void* raw, int32_t n_masks, int32_t n_planes, int64_t npix,
                           int32_t nz, const int32_t* plane_index, const void* weight,
                           const void* plane_scale, void* volume, int32_t precision,
                           void* stream) {
  g_err.clear();
  if (!raw || !plane_index || !weight || !plane_scale || !volume)
    return fail(LB_ERR_ARG, "null argument");
  if (n_masks < 1 || n_planes < 2 || npix < 1 || nz < 1)
    return fail(LB_ERR_ARG, "need n_masks >= 1, n_planes >= 2, npix >= 1, nz >= 1");
  if (precision != 32 && precision != 64) return fail(LB_ERR_ARG, "precision must be 32 or 64");
  hipStream_t st = (hipStream_t)stream;
  if (precision == 64)
    LB_HIP(volume_lerp<double>((const double*)raw, n_masks, n_planes, npix, nz, plane_index,
                               (const double*)weight, (const double*)plane_scale, (double*)volume,
                               st));
  else
    LB_HIP(volume_lerp<float>((const float*)raw, n_masks, n_planes, npix, nz, plane_index,
                              (const float*)weight, (const float*)plane_scale, (float*)volume,
                              st));
  return LB_OK;
}

lb_status lb_ilt_schedule(double* schedule, double growth, double beta1, double beta2,
                          void* stream) {
  g_err.clear();
  if (!schedule) return fail(LB_ERR_ARG, "null schedule");
  LB_HIP(ilt_schedule(schedule, growth, beta1, beta2, (hipStream_t)stream));
  return LB_OK;
}

static lb_status mask_sigmoid_any(const void* x, int64_t n, double beta, const double* pbeta,
                                  void* mask, int32_t precision, void* stream) {
  g_err.clear();
  if (!x || !mask) return fail(LB_ERR_ARG, "null argument");
  if (n < 0) return fail(LB_ERR_ARG, "n must be >= 0");
  if (precision != 32 && precision != 64) return fail(LB_ERR_ARG, "precision must be 32 or 64");
  hipStream_t st = (hipStream_t)stream;
  if (precision == 64)
    LB_HIP(mask_sigmoid<double>((const double*)x, n, beta, pbeta, (double*)mask, st));
  else
    LB_HIP(mask_sigmoid<float>((const float*)x, n, (float)beta, pbeta, (float*)mask, st));
  return LB_OK;
}

lb_status lb_mask_sigmoid(const void* x, int64_t n, double beta, void* mask, int32_t precision,
                          void* stream) {
  return mask_sigmoid_any(x, n, beta, nullptr, mask, precision, stream);
}

lb_status lb_mask_sigmoid_dev(const void* x, int64_t n, const double* beta, void* mask,
                              int32_t precision, void* stream) {
  if (!beta) return fail(LB_ERR_ARG, "null beta");
  return mask_sigmoid_any(x, n, 0.0, beta, mask, precision, stream);
}

static lb_status mask_sigmoid_grad_any(const void* grad_mask, const void* mask, int64_t n,
                                       double beta, const double* pbeta, void* grad_x,
                                       int32_t precision, void* stream) {
  g_err.clear();
  if (!grad_mask || !mask || !grad_x) return fail(LB_ERR_ARG, "null argument");
  if (n < 0) return fail(LB_ERR_ARG, "n must be >= 0");
  if (precision != 32 && precision != 64) return fail(LB_ERR_ARG, "precision must be 32 or 64");
  hipStream_t st = (hipStream_t)stream;
  if (precision == 64)
    LB_HIP(mask_sigmoid_grad<double>((const double*)grad_mask, (const double*)mask, n, beta, pbeta,
                                     (double*)grad_x, st));
  else
    LB_HIP(mask_sigmoid_grad<float>((const float*)grad_mask, (const float*)mask, n, (float)beta,
                                    pbeta, (float*)grad_x, st));
  return LB_OK;
}

lb_status lb_mask_sigmoid_grad(const void* grad_mask, const void* mask, int64_t n, double beta,
                               void* grad_x, int32_t precision, void* stream) {
  return mask_sigmoid_grad_any(grad_mask, mask, n, beta, nullptr, grad_x, precision, stream);
}

lb_status lb_mask_sigmoid_grad_dev(const void* grad_mask, const void* mask, int64_t n,
                                   const double* beta, void* grad_x, int32_t precision,
                                   void* stream) {
  if (!beta) return fail(LB_ERR_ARG, "null beta");
  return mask_sigmoid_grad_any(grad_mask, mask, n, 0.0, beta, grad_x, precision, stream);
}

lb_status lb_mask_sigmoid_grad_theta(const void* grad_mask, const void* theta, int64_t n,
                                     double beta, const void* gscale, void* grad_x,
                                     int32_t precision, void* stream) {
  g_err.clear();
  if (!grad_mask || !theta || !grad_x) return fail(LB_ERR_ARG, "null argument");
  if (n < 0) return fail(LB_ERR_ARG, "n must be >= 0");
  if (precision != 32 && precision != 64) return fail(LB_ERR_ARG, "precision must be 32 or 64");
  hipStream_t st = (hipStream_t)stream;
  if (precision == 64)
    LB_HIP(mask_sigmoid_grad_theta<double>((const double*)grad_mask, (const double*)theta, n, beta,
                                           nullptr, (const double*)gscale, (double*)grad_x, st));
  else
    LB_HIP(mask_sigmoid_grad_theta<float>((const float*)grad_mask, (const float*)theta, n,
                                          (float)beta, nullptr, (const float*)gscale,
                                          (float*)grad_x, st));
  return LB_OK;
}

lb_status lb_forward_dense(lb_plan plan, const void* mask, int32_t plane, int32_t batch,
                           void* image, void* workspace, size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!plan || !mask || !image) return fail(LB_ERR_ARG, "null argument");
  Plan& p = plan->p;
  if (plane < 0 || plane >= p.n_planes) return fail(LB_ERR_ARG, "plane out of range");
  if (batch < 1) return fail(LB_ERR_ARG, "batch must be >= 1");
  const size_t need = lb_dense_workspace_bytes(plan, batch);
  if (!workspace || ws_bytes < need)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(need) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  return p.precision == 64
             ? dense_t<double>(p, mask, plane, batch, image, (char*)workspace, st)
             : dense_t<float>(p, mask, plane, batch, image, (char*)workspace, st);
}

size_t lb_fields_workspace_bytes(lb_plan plan, int32_t n_masks) {
  if (!plan || n_masks < 1) return 0;
  const Plan& p = plan->p;
  Layout L;
  L.take(p.csz() * (size_t)n_masks * p.Wbm * p.Wbm);
  L.take(p.csz() * (size_t)n_masks * band_ld(p.Wbm) * p.N);
  L.take(p.csz() * (size_t)n_masks * p.spp * p.Wbm * p.Wbm);
  L.take(p.csz() * (size_t)n_masks * p.spp * band_ld(p.Wbm) * p.N);
  return L.off;
}

}  // extern "C"

template <typename T>
static lb_status fields_t(Plan& p, const void* masks, int cplx, int n_masks, void* fields,
                          char* ws, hipStream_t st) {
  Layout L;
  const size_t C = p.csz();
  cx<T>* Xm = (cx<T>*)(ws + L.take(C * (size_t)n_masks * p.Wbm * p.Wbm));
  cx<T>* RT = (cx<T>*)(ws + L.take(C * (size_t)n_masks * band_ld(p.Wbm) * p.N));
  cx<T>* Y = (cx<T>*)(ws + L.take(C * (size_t)n_masks * p.spp * p.Wbm * p.Wbm));
  cx<T>* TT = (cx<T>*)(ws + L.take(C * (size_t)n_masks * p.spp * band_ld(p.Wbm) * p.N));
  const T invN2 = (T)(1.0 / ((double)p.N * (double)p.N));
  LB_HIP(band_analysis<T>(p.N, masks, cplx, (long long)p.N * p.N * (cplx ? 2 : 1), 0, 1, n_masks,
                          p.nu0m, p.Wbm, invN2, RT, Xm, st));
  LB_HIP(band_mul<T>(Xm, (const cx<T>*)p.H, n_masks, p.spp, (long long)p.Wbm * p.Wbm, Y, st));
  LB_HIP(band_synthesis<T>(p.N, Y, n_masks * p.spp, p.nu0m, p.Wbm, 0, (T)1, 1, TT, (T*)fields,
                           (T*)nullptr, st));
  return LB_OK;
}

extern "C" {

lb_status lb_fields(lb_plan plan, const void* masks, int32_t cplx, int32_t n_masks, void* fields,
                    void* workspace, size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!plan || !masks || !fields) return fail(LB_ERR_ARG, "null argument");
  Plan& p = plan->p;
  if (!p.socs) return fail(LB_ERR_UNSUPPORTED, "fields are exported for SOCS plans only");
  if (n_masks < 1) return fail(LB_ERR_ARG, "n_masks must be >= 1");
  if (p.spp < 1) return LB_OK;
  const size_t need = lb_fields_workspace_bytes(plan, n_masks);
  if (!workspace || ws_bytes < need)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(need) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  return p.precision == 64 ? fields_t<double>(p, masks, cplx, n_masks, fields, (char*)workspace, st)
                           : fields_t<float>(p, masks, cplx, n_masks, fields, (char*)workspace, st);
}

}  // extern "C"

// systems whose products g E_k are transformed together by lb_fields_backward
constexpr int kFieldsAdjChunk = 16;

namespace {
struct FieldsAdjWork {
  size_t Y, RT, X, A, TT, total;
};
FieldsAdjWork fields_adj_layout(const Plan& p) {
  FieldsAdjWork f{};
  const size_t C = p.csz(), nn = (size_t)p.N * p.N;
  const int ch = std::max(1, std::min(kFieldsAdjChunk, p.spp));
  Layout L;
  f.Y = L.take(C * (size_t)ch * nn);
  f.RT = L.take(C * (size_t)ch * band_ld(p.Wbm) * p.N);
  f.X = L.take(C * (size_t)ch * p.Wbm * p.Wbm);
  f.A = L.take(C * (size_t)p.Wbm * p.Wbm);
  f.TT = L.take(C * (size_t)band_ld(p.Wbm) * p.N);
  f.total = L.off;
  return f;
}
}  // namespace

// socs_gradient (opc/ilt.py:235-266) on the fields given: per mask, for chunks of
// systems, Y_k = g E_k, X_k = fft2(Y_k) on the mask band, A += w_k conj(H_k) X_k on each
// box; then d mask = (2 / N^2) Re sum A e^{+...} (or the complex sum when cout).
template <typename T>
static lb_status fields_bwd_t(Plan& p, const void* fields, int n_masks, const void* sens,
                              void* dmask, bool cout, char* ws, hipStream_t st) {
  const FieldsAdjWork f = fields_adj_layout(p);
  const long long nn = (long long)p.N * p.N;
  cx<T>* Y = (cx<T>*)(ws + f.Y);
  cx<T>* RT = (cx<T>*)(ws + f.RT);
  cx<T>* X = (cx<T>*)(ws + f.X);
  cx<T>* A = (cx<T>*)(ws + f.A);
  cx<T>* TT = (cx<T>*)(ws + f.TT);
  const T s2 = (T)(2.0 / ((double)p.N * (double)p.N));
  for (int m = 0; m < n_masks; ++m) {
    const cx<T>* E = (const cx<T>*)fields + (long long)m * p.spp * nn;
    const T* g = (const T*)sens + (long long)m * nn;
    for (int j0 = 0; j0 < p.spp; j0 += kFieldsAdjChunk) {
      const int nk = std::min(kFieldsAdjChunk, p.spp - j0);
      LB_HIP(mul_real_into<T>(E + (long long)j0 * nn, nk, nn, g, Y, st));
      LB_HIP(band_analysis<T>(p.N, Y, 1, 2 * nn, 0, 1, nk, p.nu0m, p.Wbm, (T)1, RT, X, st));
      LB_HIP(fields_adj_band<T>(X, p.Wbm, p.wrapm, (const cx<T>*)p.H, p.W, p.box,
                                (const T*)p.wsys, j0, nk, j0 > 0 ? 1 : 0, A, st));
    }
    T* out = (T*)dmask + (long long)m * nn * (cout ? 2 : 1);
    LB_HIP(band_synthesis<T>(p.N, A, 1, p.nu0m, p.Wbm, cout ? 0 : 1, s2, cout ? 1 : 0, TT, out,
                             (T*)nullptr, st));
  }
  return LB_OK;
}

extern "C" {

size_t lb_fields_backward_workspace_bytes(lb_plan plan) {
  if (!plan) return 0;
  return fields_adj_layout(plan->p).total;
}

lb_status lb_fields_backward(lb_plan plan, const void* fields, int32_t n_masks,
                             const void* sensitivity, void* d_masks, int32_t complex_out,
                             void* workspace, size_t ws_bytes, void* stream) {
  g_err.clear();
  if (!plan || !fields || !sensitivity || !d_masks) return fail(LB_ERR_ARG, "null argument");
  Plan& p = plan->p;
  if (!p.socs) return fail(LB_ERR_UNSUPPORTED, "fields are exported for SOCS plans only");
  if (n_masks < 1) return fail(LB_ERR_ARG, "n_masks must be >= 1");
  if (p.spp < 1) return fail(LB_ERR_ARG, "plan has no kernels");
  const size_t need = lb_fields_backward_workspace_bytes(plan);
  if (!workspace || ws_bytes < need)
    return fail(LB_ERR_WORKSPACE, "workspace needs " + std::to_string(need) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  return p.precision == 64
             ? fields_bwd_t<double>(p, fields, n_masks, sensitivity, d_masks, complex_out != 0,
                                    (char*)workspace, st)
             : fields_bwd_t<float>(p, fields, n_masks, sensitivity, d_masks, complex_out != 0,
                                   (char*)workspace, st);
}

}  // extern "C"
