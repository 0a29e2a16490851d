// C ABI entry points (include/fastergs.h): argument checks, workspace
// carving and the launch sequence of each schedule.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fastergs.h"
#include "fgs_kernels.h"
#include "fgs_sched.h"
#include "fgs_layout.h"
#include "fgs_prof.h"

using namespace fgs;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_check(hipError_t e, const char* where) {
  if (e == hipSuccess) return FGS_OK;
  return fail(FGS_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define FGS_TRY(expr, where)                    \
  do {                                          \
    int rc__ = hip_check((expr), (where));      \
    if (rc__ != FGS_OK) return rc__;            \
  } while (0)

// Side streams for stages that do not depend on each other. Per host thread
// and device: two non-blocking streams of default priority and a few events,
// created on first use outside stream capture (a call made during capture
// before that runs every stage on the caller's stream). fork(i, e) makes
// side stream i wait for the work queued so far on the main stream; join(i,
// e) makes the main stream wait for side stream i. Under capture the side
// streams join the same graph through these events; every fork is joined
// before the entry point returns.
constexpr int SIDE_STREAMS = 2, SIDE_EVENTS = 10, SIDE_MAX_DEVICES = 64;
struct SideCtx {
  hipStream_t st[SIDE_STREAMS] = {};
  hipEvent_t ev[SIDE_EVENTS] = {};
  bool ready = false, failed = false;
};
thread_local SideCtx t_side[SIDE_MAX_DEVICES];

SideCtx* side_ctx(hipStream_t s) {
  int dev = -1;
  if (hipStreamGetDevice(s, &dev) != hipSuccess || dev < 0 || dev >= SIDE_MAX_DEVICES)
    return nullptr;
  SideCtx& c = t_side[dev];
  if (c.ready) return &c;
  if (c.failed) return nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return nullptr;
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess) return nullptr;
  bool ok = cur == dev || hipSetDevice(dev) == hipSuccess;
  for (int i = 0; ok && i < SIDE_STREAMS; ++i)
    ok = hipStreamCreateWithFlags(&c.st[i], hipStreamNonBlocking) == hipSuccess;
  for (int i = 0; ok && i < SIDE_EVENTS; ++i)
    ok = hipEventCreateWithFlags(&c.ev[i], hipEventDisableTiming) == hipSuccess;
  if (cur != dev) (void)hipSetDevice(cur);
  (void)hipGetLastError();
  if (!ok) {
    c.failed = true;
    return nullptr;
  }
  c.ready = true;
  return &c;
}

struct Fork {
  SideCtx* c;
  hipStream_t main;
  explicit Fork(hipStream_t s) : c(side_ctx(s)), main(s) {}
  hipStream_t side(int i) const { return c ? c->st[i] : main; }
  hipError_t fork(int i, int e) const {
    if (!c) return hipSuccess;
    hipError_t r = hipEventRecord(c->ev[e], main);
    return r != hipSuccess ? r : hipStreamWaitEvent(c->st[i], c->ev[e], 0);
  }
  // mark(i, e): event e at the current end of side stream i; wait(e): the
  // main stream waits for it; join = mark + wait.
  hipError_t mark(int i, int e) const { return c ? hipEventRecord(c->ev[e], c->st[i]) : hipSuccess; }
  hipError_t wait(int e) const { return c ? hipStreamWaitEvent(main, c->ev[e], 0) : hipSuccess; }
  hipError_t join(int i, int e) const {
    hipError_t r = mark(i, e);
    return r != hipSuccess ? r : wait(e);
  }
};

// Tile width of the options (0: the reference's 16; tiles hold 256 pixels).
int tile_w_of(const fgs_options* o) { return o && o->tile_w ? o->tile_w : TILE_W_REF; }
int grid_x_of(int W, int tw) { return grid_dim(W, tw); }
int grid_y_of(int H, int tw) { return grid_dim(H, tile_h_of(tw)); }
int tiles_of(int W, int H, int tw) { return grid_x_of(W, tw) * grid_y_of(H, tw); }

// Number of bits needed for tile ids (sort.cu:9-24 computes the same).
int tile_bits(int T) {
  int b = 0;
  while (b < 31 && (T >> b) != 0) ++b;
  return b;
}

// The plain schedule sorts 64-bit keys on [0, 32 + tile_bits); an odd
// number of 8-bit passes leaves the result in the alternate buffers.
bool plain_sorted_in_alt(int T) {
  const int passes = (32 + tile_bits(T) + 7) / 8;
  return (passes & 1) != 0;
}

// Whether a frame keeps what the backward reads (bucket checkpoints and
// per-pixel colours): not when rendered forward-only.
bool keeps_ckpt(const fgs_options* o) { return !(o && o->forward_only); }

int check_opts(const fgs_options* o) {
  if (!o) return fail(FGS_ERR_ARG, "options is NULL");
  if (o->schedule != FGS_SCHEDULE_PLAIN && o->schedule != FGS_SCHEDULE_FAST)
    return fail(FGS_ERR_ARG, "unknown schedule");
  const int b = o->bucket_size;
  if (!(b == 16 || b == 32 || b == 64 || b == 128 || b == 256))
    return fail(FGS_ERR_ARG, "bucket_size must be 16, 32, 64, 128 or 256");
  if (o->split_tiles < 1 || o->split_tiles > (1 << 24))
    return fail(FGS_ERR_ARG, "split_tiles must be in [1, 2^24]");
  if (o->tile_w != 0 && !tile_w_ok(o->tile_w))
    return fail(FGS_ERR_ARG, "tile_w must be 8, 16 or 32 (0: 16)");
  if (o->list_prefix < 0) return fail(FGS_ERR_ARG, "list_prefix must be >= 0");
  if (const fgs_tile_sched* ts = o->tile_sched) {
    if (ts->mode < FGS_TILES_RASTER || ts->mode > FGS_TILES_LEARNED)
      return fail(FGS_ERR_ARG, "unknown tile schedule mode");
    if (ts->mode == FGS_TILES_LEARNED && !ts->mlp_params)
      return fail(FGS_ERR_ARG, "learned tile schedule needs mlp_params");
    if (!ts->items != !ts->n_items)
      return fail(FGS_ERR_ARG, "a work list needs both items and n_items");
    if (ts->prev_features && (ts->mode != FGS_TILES_LEARNED || !ts->items))
      return fail(FGS_ERR_ARG, "prev_features needs the learned tile schedule and items");
  }
  return FGS_OK;
}

// The list passes read whole lists: a frame rendered with prefix lists is
// rendered again with list_prefix 0 first (the caller passes the frame's
// options to every call on it).
int check_whole_lists(const fgs_options* o) {
  if (o->list_prefix > 0)
    return fail(FGS_ERR_ARG,
                "the frame's lists may be prefix lists: render it again with list_prefix 0 first");
  return FGS_OK;
}

int check_cam(const fgs_camera* c, int tw) {
  if (!c) return fail(FGS_ERR_ARG, "camera is NULL");
  if (c->width <= 0 || c->height <= 0) return fail(FGS_ERR_ARG, "image size must be positive");
  if (grid_x_of(c->width, tw) > 65535 || grid_y_of(c->height, tw) > 65535)
    return fail(FGS_ERR_ARG, "image too large");
  if (!c->viewmatrix || !c->projmatrix || !c->campos || !c->bg)
    return fail(FGS_ERR_ARG, "camera matrices must be device pointers");
  if (c->sh_degree < 0 || c->sh_degree > 3) return fail(FGS_ERR_ARG, "sh_degree must be 0..3");
  return FGS_OK;
}

int check_gauss(const fgs_gaussians* g, const fgs_camera* c) {
  if (!g) return fail(FGS_ERR_ARG, "gaussians is NULL");
  if (g->P < 0) return fail(FGS_ERR_ARG, "P must be >= 0");
  if (g->P == 0) return FGS_OK;
  if (!g->means3D || !g->opacities || !g->scales || !g->rotations)
    return fail(FGS_ERR_ARG, "means3D, opacities, scales and rotations are required");
  if (!g->colors_precomp) {
    if (!g->dc) return fail(FGS_ERR_ARG, "dc is required when colours come from SH");
    const int need = sh_count(c->sh_degree) - 1;
    if (need > 0 && (!g->shs || g->max_coeffs < need))
      return fail(FGS_ERR_ARG, "shs has fewer coefficients than sh_degree needs");
  }
  return FGS_OK;
}

CamRef<float> make_cam(const fgs_camera* c, const fgs_gaussians* g, int tw) {
  CamRef<float> r;
  r.view = c->viewmatrix;
  r.proj = c->projmatrix;
  r.campos = c->campos;
  r.tanfovx = c->tanfovx;
  r.tanfovy = c->tanfovy;
  r.focal_x = (float)c->width / (2.0f * c->tanfovx);   // preprocess.cu:941-942
  r.focal_y = (float)c->height / (2.0f * c->tanfovy);
  r.scale_mod = c->scale_modifier;
  r.mult = c->mult;
  r.W = c->width;
  r.H = c->height;
  r.grid_x = grid_x_of(c->width, tw);
  r.grid_y = grid_y_of(c->height, tw);
  r.tile_w = tw;
  r.tile_h = tile_h_of(tw);
  r.deg = c->sh_degree;
  r.max_coeffs = (g->colors_precomp || !g->shs) ? 0 : g->max_coeffs;
  return r;
}

// A batch of views (FAST); a single view is a batch of one.
struct Batch {
  int V = 0;
  const fgs_camera* cams = nullptr;
  int T = 0;        // tiles of all views
  int tw = TILE_W_REF;  // tile width (one shape for every view)
  int64_t HW = 0;   // pixels of all views
  ViewSetup setup;  // kernel-argument description of the view table
};

// With check false the batch is only sized (workspace queries, views); the
// options give the tile shape either way (NULL: 16 x 16).
int make_batch(const fgs_batch* batch, Batch& b, const fgs_options* opt, bool check) {
  static_assert(sizeof(ViewRec) <= VIEW_REC_BYTES, "view record too large");
  static_assert(MAX_VIEWS <= VIEW_TABLE_VIEWS, "view table too small");
  if (!batch || !batch->cameras) return fail(FGS_ERR_ARG, "batch is NULL");
  if (batch->n_views < 1 || batch->n_views > MAX_VIEWS)
    return fail(FGS_ERR_ARG, "n_views must be in [1, 32]");
  if (check && opt->schedule == FGS_SCHEDULE_PLAIN && batch->n_views != 1)
    return fail(FGS_ERR_ARG, "the PLAIN schedule renders one view per call");
  b.V = batch->n_views;
  b.tw = tile_w_of(opt);
  if (!tile_w_ok(b.tw)) return fail(FGS_ERR_ARG, "tile_w must be 8, 16 or 32 (0: 16)");
  b.cams = batch->cameras;
  int64_t T = 0, HW = 0;
  b.setup.n = b.V;
  for (int v = 0; v < b.V; ++v) {
    const fgs_camera* c = &b.cams[v];
    if (check) {
      int rc = check_cam(c, b.tw);
      if (rc) return rc;
      if (v > 0 && (c->sh_degree != b.cams[0].sh_degree ||
                    c->scale_modifier != b.cams[0].scale_modifier || c->mult != b.cams[0].mult))
        return fail(FGS_ERR_ARG, "views of a batch must share sh_degree, scale_modifier and mult");
    } else if (c->width <= 0 || c->height <= 0) {
      return fail(FGS_ERR_ARG, "image size must be positive");
    }
    ViewSrc& src = b.setup.v[v];
    src.view = c->viewmatrix;
    src.proj = c->projmatrix;
    src.campos = c->campos;
    src.bg = c->bg;
    src.tanfovx = c->tanfovx;
    src.tanfovy = c->tanfovy;
    src.focal_x = (float)c->width / (2.0f * c->tanfovx);  // preprocess.cu:941-942
    src.focal_y = (float)c->height / (2.0f * c->tanfovy);
    src.W = c->width;
    src.H = c->height;
    src.tile_off = (uint32_t)T;
    src.pix_off = (uint32_t)HW;
    src.tile_w = b.tw;
    src.tile_h = tile_h_of(b.tw);
    T += tiles_of(c->width, c->height, b.tw);
    HW += (int64_t)c->width * c->height;
  }
  if (T >= (1 << 24) || HW >= (int64_t)1 << 32) return fail(FGS_ERR_ARG, "batch too large");
  b.T = (int)T;
  b.HW = HW;
  return FGS_OK;
}

GaussRef<float> make_gauss(const fgs_gaussians* g) {
  GaussRef<float> r;
  r.means = g->means3D;
  r.opac = g->opacities;
  r.scales = g->scales;
  r.rots = g->rotations;
  r.colors = g->colors_precomp;
  r.dc = g->colors_precomp ? nullptr : g->dc;
  r.shs = g->colors_precomp ? nullptr : g->shs;
  r.culling = g->culling;
  return r;
}

struct ProfRecord {
  const char* name;
  hipEvent_t start, stop;
};
struct Profiler {
  bool on = false;
  std::vector<ProfRecord> recs;
  std::vector<hipEvent_t> pool;
  hipEvent_t take() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
};
Profiler g_prof;

}  // namespace

namespace fgs {
bool prof_on() { return g_prof.on; }
hipEvent_t prof_begin(hipStream_t s) {
  hipEvent_t e = g_prof.take();
  if (e) hipEventRecord(e, s);
  return e;
}
void prof_end(const char* name, hipEvent_t start, hipStream_t s) {
  hipEvent_t e = g_prof.take();
  if (!e) return;
  hipEventRecord(e, s);
  g_prof.recs.push_back({name, start, e});
}
}  // namespace fgs

extern "C" {

int fgs_profile_enable(int on) {
  for (auto& r : g_prof.recs) {
    g_prof.pool.push_back(r.start);
    g_prof.pool.push_back(r.stop);
  }
  g_prof.recs.clear();
  g_prof.on = on != 0;
  return FGS_OK;
}

int fgs_profile_read(char* buf, size_t cap) {
  std::map<std::string, std::pair<double, int>> agg;
  for (auto& r : g_prof.recs) {
    if (hipEventSynchronize(r.stop) != hipSuccess) return fail(FGS_ERR_HIP, "profile sync");
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, r.start, r.stop) != hipSuccess) return fail(FGS_ERR_HIP, "profile time");
    auto& a = agg[r.name];
    a.first += ms;
    a.second += 1;
  }
  std::string out;
  char line[256];
  for (auto& kv : agg) {
    snprintf(line, sizeof(line), "%s %.6f %d\n", kv.first.c_str(), kv.second.first, kv.second.second);
    out += line;
  }
  if (buf && cap) {
    const size_t n = out.size() < cap - 1 ? out.size() : cap - 1;
    memcpy(buf, out.data(), n);
    buf[n] = 0;
  }
  return (int)out.size();
}

const char* fgs_last_error(void) { return g_err.c_str(); }

const char* fgs_build_info(void) {
  return "fastergs gfx950: wave64, -ffp-contract=off, own exp2/log2, IEEE div/sqrt";
}

size_t fgs_gauss_ws_size(int P, int width, int height, const fgs_options* opt) {
  return carve_gauss(nullptr, P, tiles_of(width, height, tile_w_of(opt)), true).total;
}

size_t fgs_pair_ws_size(int P, int width, int height, int64_t R, int64_t NB,
                        const fgs_options* opt) {
  (void)P;
  const bool fast = opt && opt->schedule == FGS_SCHEDULE_FAST;
  return carve_pair(nullptr, R, NB, (int64_t)width * height, tiles_of(width, height, tile_w_of(opt)), fast,
                    keeps_ckpt(opt)).total;
}

// raw sums (P x 4 doubles + P x 8 floats) and the split SH pass's mean term
// (P x 4 floats)
size_t fgs_grad_ws_size(int P) {
  return align_up((size_t)(P > 0 ? P : 1) *
                  (RAW_CON_STRIDE * sizeof(double) + (RAW_LIN_STRIDE + 4) * sizeof(float)));
}

size_t fgs_batch_gauss_ws_size(int P, const fgs_batch* batch, const fgs_options* opt) {
  Batch b;
  if (make_batch(batch, b, opt, false) != FGS_OK) return 0;
  return carve_gauss(nullptr, (int64_t)P * b.V, b.T, b.V == 1).total;
}

size_t fgs_batch_pair_ws_size(int P, const fgs_batch* batch, int64_t R, int64_t NB,
                              const fgs_options* opt) {
  (void)P;
  Batch b;
  if (make_batch(batch, b, opt, false) != FGS_OK) return 0;
  return carve_pair(nullptr, R, NB, b.HW, b.T, true, keeps_ckpt(opt)).total;
}

size_t fgs_batch_grad_ws_size(int P, int n_views) {
  const size_t p = (size_t)(P > 0 ? P : 1), v = (size_t)(n_views > 0 ? n_views : 1);
  return align_up(p * v * (RAW_CON_STRIDE * sizeof(double) + RAW_LIN_STRIDE * sizeof(float)) +
                  (v == 1 ? p * 4 * sizeof(float) : 0));
}

size_t fgs_sort_ws_size(int64_t n) { return sort_ws_bytes(n); }
size_t fgs_scan_ws_size(int64_t n) { return scan_ws_bytes(n); }

int fgs_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev, void* ws,
                 size_t ws_bytes, void* stream) {
  if (n < 0 || (n > 0 && (!in || !out))) return fail(FGS_ERR_ARG, "bad scan arguments");
  if (ws_bytes < scan_ws_bytes(n)) return fail(FGS_ERR_CAPACITY, "scan workspace too small");
  FGS_TRY(scan_u32(in, out, n, total_dev, ws, (hipStream_t)stream), "scan");
  return FGS_OK;
}

// ------------------------------------------------ no-allocation sort
// fgs_sort_pairs_noalloc keeps its workspace in a file-scope device array
// (scratch) instead of a caller buffer, for callers that cannot allocate.
// Each device holds its own copy (loaded with the library, kept until the
// process ends); host code uses the address hipGetSymbolAddress returns for
// the device, never the array's name (a host-side stand-in).
constexpr int64_t NOALLOC_SORT_MAX_N = int64_t(1) << 24;
constexpr size_t NOALLOC_SORT_BYTES = size_t(8) << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SORT_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The scratch array's address on the stream's device, looked up once per
// device with that device current.
int scratch_on(hipStream_t s, void** out) {
  int dev = -1, cur = -1;
  FGS_TRY(hipStreamGetDevice(s, &dev), "hipStreamGetDevice");
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(FGS_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    FGS_TRY(hipGetDevice(&cur), "hipGetDevice");
    if (cur != dev) FGS_TRY(hipSetDevice(dev), "hipSetDevice");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    FGS_TRY(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(FGS_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FGS_OK;
}
}  // namespace

int fgs_sort_pairs_noalloc(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                           int64_t n, int begin_bit, int end_bit, void* stream) {
  if (n < 0 || begin_bit < 0 || end_bit > 64 || (n > 0 && (!keys || !keys_alt)))
    return fail(FGS_ERR_ARG, "bad sort arguments");
  if ((vals == nullptr) != (vals_alt == nullptr)) return fail(FGS_ERR_ARG, "vals/vals_alt mismatch");
  if (n > NOALLOC_SORT_MAX_N)
    return fail(FGS_ERR_CAPACITY, "fgs_sort_pairs_noalloc sorts at most 2^24 pairs");
  if (sort_classic_ws_bytes(n) > NOALLOC_SORT_BYTES)
    return fail(FGS_ERR_CAPACITY, "sort scratch too small");
  hipStream_t s = (hipStream_t)stream;
  void* ws = nullptr;
  if (int rc = scratch_on(s, &ws)) return rc;
  bool alt = false;
  FGS_TRY(sort_kv_passes<uint64_t>(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, ws, s,
                                   alt),
          "sort");
  if (alt) {
    FGS_TRY(hipMemcpyAsync(keys, keys_alt, n * 8, hipMemcpyDeviceToDevice, s), "sort copy");
    if (vals) FGS_TRY(hipMemcpyAsync(vals, vals_alt, n * 4, hipMemcpyDeviceToDevice, s), "sort copy");
  }
  return FGS_OK;
}

int fgs_sort_check(const void* ws, int64_t n, void* stream) {
  if (!ws || n < 0) return fail(FGS_ERR_ARG, "bad sort check arguments");
  if (n <= 1 || n > ONESWEEP_MAX_KEYS) return FGS_OK;  // three-launch passes never wait
  // the fault word sits after the histograms, before the per-tile rows
  const SortWs w = carve_sort(const_cast<void*>(ws), 1);
  uint32_t fault = 0;
  hipStream_t s = (hipStream_t)stream;
  FGS_TRY(hipMemcpyAsync(&fault, w.counters + SORT_MAX_PASSES, sizeof(fault),
                         hipMemcpyDeviceToHost, s),
          "read sort fault");
  FGS_TRY(hipStreamSynchronize(s), "sort check sync");
  if (fault) return fail(FGS_ERR_HIP, "a look-back wait of the sort exceeded its bound");
  return FGS_OK;
}

int fgs_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                   int64_t n, int begin_bit, int end_bit, void* ws, size_t ws_bytes,
                   void* stream) {
  if (n < 0 || begin_bit < 0 || end_bit > 64 || (n > 0 && (!keys || !keys_alt)))
    return fail(FGS_ERR_ARG, "bad sort arguments");
  if ((vals == nullptr) != (vals_alt == nullptr)) return fail(FGS_ERR_ARG, "vals/vals_alt mismatch");
  if (n >= (int64_t)0xffffffffll) return fail(FGS_ERR_ARG, "sort supports < 2^32 elements");
  if (ws_bytes < sort_ws_bytes(n)) return fail(FGS_ERR_CAPACITY, "sort workspace too small");
  hipStream_t s = (hipStream_t)stream;
  bool alt = false;
  FGS_TRY(sort_kv<uint64_t>(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, ws, s, alt,
                            nullptr, nullptr, SORT_PAIRS),
          "sort");
  if (alt) {
    FGS_TRY(hipMemcpyAsync(keys, keys_alt, n * 8, hipMemcpyDeviceToDevice, s), "sort copy");
    if (vals) FGS_TRY(hipMemcpyAsync(vals, vals_alt, n * 4, hipMemcpyDeviceToDevice, s), "sort copy");
  }
  return FGS_OK;
}

// Arguments of the per-Gaussian passes over a carved gauss workspace.
PreArgs pre_args(const Batch& b, const fgs_gaussians* g, const fgs_options* opt, const GaussWs& w,
                 int32_t* radii) {
  PreArgs a;
  a.G = make_gauss(g);
  a.cam = make_cam(&b.cams[0], g, b.tw);
  a.geom = w.geom;
  a.splat = w.splat;
  a.dkey = w.dkey;
  a.pcount = w.pcount;
  a.radii = radii;
  a.P = g->P;
  a.split_tiles = opt->split_tiles;
  a.lflag = w.lflag;
  a.dsort_k = w.dsort_k0;
  a.dsort_v = w.dsort_v0;
  a.counts = w.counts;
  a.views = (const ViewRec*)w.views;
  a.V = b.V;
  // single FAST views that will be differentiated keep the SH Jacobians
  a.shjac = b.V == 1 && opt->schedule == FGS_SCHEDULE_FAST && !opt->forward_only ? w.shjac
                                                                                  : nullptr;
  return a;
}

int fgs_batch_preprocess(const fgs_batch* batch, const fgs_gaussians* g, const fgs_options* opt,
                         void* gauss_ws, size_t gauss_ws_bytes, int32_t* radii, void* stream) {
  Batch b;
  int rc;
  if ((rc = check_opts(opt)) || (rc = make_batch(batch, b, opt, true)) ||
      (rc = check_gauss(g, &b.cams[0])))
    return rc;
  const int64_t PV = (int64_t)g->P * b.V;
  // band list entries are record * LARGE_BANDS + band in 32 bits
  if (PV >= ((int64_t)1 << 29)) return fail(FGS_ERR_ARG, "P x n_views must be < 2^29");
  const GaussWs w = carve_gauss(gauss_ws, PV, b.T, b.V == 1);
  if (!gauss_ws || gauss_ws_bytes < w.total) return fail(FGS_ERR_CAPACITY, "gauss workspace too small");
  if (g->P > 0 && !radii) return fail(FGS_ERR_ARG, "radii is NULL");
  hipStream_t s = (hipStream_t)stream;
  const PreArgs a = pre_args(b, g, opt, w, radii);
  if (opt->schedule == FGS_SCHEDULE_PLAIN) {
    FGS_TRY(fill_u32(w.counts, 0u, GAUSS_SYNC_WORDS, s), "fill counts");
    if (g->P == 0) return FGS_OK;
    FGS_TRY(launch_preprocess_plain(a, s), "preprocess");
    FGS_TRY(scan_u32(w.pcount, w.poff, g->P, w.counts + 0, w.scratch, s), "pair scan");
    return FGS_OK;
  }
  // FAST: the view table, then per-Gaussian math + lane-packed exact counts
  // of small rects for every view
  FGS_TRY(launch_preprocess_fast(a, b.setup, (ViewRec*)w.views, w.counts, GAUSS_SYNC_WORDS, s),
          "preprocess");
  if (g->P == 0) return FGS_OK;
  // Independent chains after the geometry pass: on a side stream the
  // large-rect compaction and tile counts, then the SH colours (read only by
  // compositing), which fill the GPU that the depth sort's last,
  // latency-bound passes leave idle; on the main stream the depth sort of
  // the records, the longer chain (a cross-stream wait costs the waiting
  // stream 10-15 us, so the longer chain stays on the main stream and the
  // side stream has finished when it is joined). The pair offsets wait for
  // the tile counts only; the colours are joined at the end.
  const Fork f(s);
  // the learned schedule's work list for this frame from the previous
  // frame's tile features (cost model, longest predicted first), on side
  // stream 0 beside the whole preprocess
  const fgs_tile_sched* ts = opt->tile_sched;
  const bool list_side = ts && ts->prev_features;
  if (list_side) {
    FGS_TRY(f.fork(0, 8), "fork");
    FGS_TRY(launch_mlp_infer(ts->mlp_params, ts->prev_features, b.T, w.pred, true, f.side(0)),
            "cost model");
    FGS_TRY(launch_sched_items(b.T, nullptr, opt->bucket_size, w.pred, ts->items, ts->n_items,
                               f.side(0)),
            "work items");
  }
  FGS_TRY(f.fork(1, 1), "fork");
  // large rects, one list entry per row band (count in counts[3]), each band
  // walked by a whole wave
  FGS_TRY(expand_flags(w.lflag, PV, w.large_ids, w.counts + 3, w.cscratch, f.side(1)),
          "large bands");
  FGS_TRY(launch_large_tiles(false, g->P, b.V, (const ViewRec*)w.views, w.geom, w.large_ids,
                             w.counts, w.pcount, nullptr, nullptr, w.lband, 0u, nullptr, nullptr,
                             false, f.side(1)),
          "large tiles");
  FGS_TRY(f.mark(1, 2), "mark");
  FGS_TRY(launch_sh_color(a, f.side(1)), "colour");
  // records in (depth, index) order; pair offsets assigned in that order so
  // that a stable sort by (global) tile alone yields the canonical per-tile
  // order of every view
  bool alt = false;
  FGS_TRY(sort_kv<uint64_t>(w.dsort_k0, w.dsort_v0, w.dsort_k1, w.dsort_v1, PV, 32, 64,
                            w.scratch, s, alt, nullptr, nullptr, SORT_DEPTH, w.dsort_sync),
          "depth sort");
  if (alt) return fail(FGS_ERR_HIP, "depth sort parity");
  FGS_TRY(f.wait(2), "join");
  FGS_TRY(scan_pair_offsets(w.dsort_k0, w.dsort_v0, w.pcount, w.lrank, PV, w.poff, w.counts,
                            opt->bucket_size, b.T, w.scratch, s),
          "pair offsets");
  FGS_TRY(f.join(1, 3), "join");  // the colours
  if (list_side) FGS_TRY(f.join(0, 9), "join");
  return FGS_OK;
}

int fgs_batch_read_counts(const void* gauss_ws, int P, const fgs_batch* batch,
                          const fgs_options* opt, int64_t counts_host[2], void* stream) {
  Batch b;
  int rc;
  if ((rc = make_batch(batch, b, opt, false))) return rc;
  if (!gauss_ws || !counts_host) return fail(FGS_ERR_ARG, "NULL argument");
  const GaussWs w = carve_gauss(const_cast<void*>(gauss_ws), (int64_t)P * b.V, b.T, b.V == 1);
  // [0] pairs, [1] bucket-table capacity, up to the depth sort's fault flag
  uint32_t c[DSORT_FAULT + 1];
  hipStream_t s = (hipStream_t)stream;
  FGS_TRY(hipMemcpyAsync(c, w.counts, sizeof(c), hipMemcpyDeviceToHost, s), "read counts");
  FGS_TRY(hipStreamSynchronize(s), "read counts sync");
  if (opt && opt->schedule == FGS_SCHEDULE_FAST && c[DSORT_FAULT])
    return fail(FGS_ERR_HIP, "a look-back wait of the preprocess exceeded its bound");
  if (c[COUNT_LIST_FAULT])
    return fail(FGS_ERR_HIP, "a look-back wait of the list sort exceeded its bound");
  counts_host[0] = c[0];
  counts_host[1] = c[1];
  return FGS_OK;
}

int fgs_batch_render(const fgs_batch* batch, const fgs_gaussians* g, const fgs_options* opt,
                     void* gauss_ws, size_t gauss_ws_bytes, void* pair_ws, size_t pair_ws_bytes,
                     int64_t R, int64_t NB, float* out_color, float* T_final,
                     int32_t* last_contributor, void* stream) {
  Batch b;
  int rc;
  if ((rc = check_opts(opt)) || (rc = make_batch(batch, b, opt, true)) ||
      (rc = check_gauss(g, &b.cams[0])))
    return rc;
  if (!out_color || !T_final || !last_contributor) return fail(FGS_ERR_ARG, "NULL output");
  if (R < 0 || NB < 0 || R >= (int64_t)0xffffffffll) return fail(FGS_ERR_ARG, "bad R / NB");
  const int T = b.T;
  const GaussWs gw = carve_gauss(gauss_ws, (int64_t)g->P * b.V, T, b.V == 1);
  if (!gauss_ws || gauss_ws_bytes < gw.total) return fail(FGS_ERR_CAPACITY, "gauss workspace too small");
  const bool fast = opt->schedule == FGS_SCHEDULE_FAST;
  const PairWs pw = carve_pair(pair_ws, R, NB, b.HW, T, fast, keeps_ckpt(opt));
  if (!pair_ws || pair_ws_bytes < pw.total) return fail(FGS_ERR_CAPACITY, "pair workspace too small");
  hipStream_t s = (hipStream_t)stream;
  const fgs_camera* cam = &b.cams[0];
  const int W = cam->width, H = cam->height;
  const int gx = grid_x_of(W, b.tw);
  CompArgs c;
  c.counts = gw.counts;
  c.W = W;
  c.H = H;
  c.grid_x = gx;
  c.tile_w = b.tw;
  c.T = T;
  c.bucket = opt->bucket_size;
  c.ranges = gw.ranges;
  c.splat = gw.splat;
  c.bg = cam->bg;
  c.boff = gw.boff;
  c.out = out_color;
  c.tfinal = T_final;
  c.contrib = last_contributor;
  c.pixc = pw.pixc;
  c.maxc = gw.maxc;
  c.acc = pw.acc;
  c.items = nullptr;
  c.n_items = nullptr;
  c.slots = 0;
  c.split_len = 0;
  c.split_pred = nullptr;
  c.cycles = nullptr;
  c.queue = fast && opt->composite_queue ? gw.counts + COMPOSITE_QUEUE_OFF : nullptr;
  c.views = (const ViewRec*)gw.views;
  c.V = b.V;
  c.redo = nullptr;
  c.status = nullptr;
  c.gate = nullptr;
  c.zero = nullptr;
  c.zero_n = 0;
  if (fast) {
    const ViewRec* views = (const ViewRec*)gw.views;
    const int64_t NR = (int64_t)g->P * b.V;
    // prefix lists: the pairs of the list_prefix nearest records first. The
    // depth order interleaves the views' records; restricted to one view it
    // is that view's depth order, so a prefix of it is a prefix of every
    // view's order.
    const bool prefix = !opt->tile_sched && opt->list_prefix > 0 &&
                        (int64_t)opt->list_prefix < NR;
    const Fork f(s);
    // the backward's raw sums (all-zero bits) when the caller hands its
    // gradient workspace over now: stored by compositing, whose time is VALU
    // issue and which leaves HBM bandwidth free (a fill kernel beside the
    // list build, joined back, cost the step as much as one in the backward)
    c.zero = (float*)opt->zero_grad_ws;
    c.zero_n = opt->zero_grad_ws ? NR * (2 * RAW_CON_STRIDE + RAW_LIN_STRIDE) : 0;
    // emission in depth order, tile-only stable sort, lists, ranges, buckets.
    // R is the pair capacity; the pair count itself (counts[0], or the
    // prefix's counts[COUNT_PREFIX_N]) stays on the device, so nothing here
    // waits for the host. redo: the second, whole-list build of a prefix
    // frame, whose launches return at once unless compositing asked for it.
    const int tb = tile_bits(T - 1) > 0 ? tile_bits(T - 1) : 1;
    auto build_lists = [&](int mode) -> int {
      const bool redo = mode == EMIT_REDO;
      const uint32_t* gate = redo ? gw.counts + COUNT_REDO : nullptr;
      const uint32_t* n_sort = mode == EMIT_PREFIX ? gw.counts + COUNT_PREFIX_N
                               : redo              ? gw.counts + COUNT_REDO_N
                                                   : gw.counts;
      const int64_t limit = mode == EMIT_PREFIX ? (int64_t)opt->list_prefix : NR;
      if (mode == EMIT_WHOLE) {
        // small and large rects write disjoint pair blocks: side by side
        FGS_TRY(f.fork(0, 4), "fork");
        FGS_TRY(launch_emit_small(g->P, b.V, views, gw.geom, opt->split_tiles, gw.dsort_k0,
                                  gw.poff, (uint32_t)R, pw.t0, pw.r0, pw.tile16, f.side(0), NR,
                                  gw.counts, EMIT_WHOLE),
                "emit");
        FGS_TRY(launch_large_tiles(true, g->P, b.V, views, gw.geom, gw.large_ids, gw.counts,
                                   nullptr, gw.poff, gw.lrank, gw.lband, (uint32_t)R, pw.t0,
                                   pw.r0, pw.tile16, s),
                "emit large");
        FGS_TRY(f.join(0, 5), "join");
      } else if (mode == EMIT_PREFIX) {
        // the prefix's small and large rects side by side too
        FGS_TRY(f.fork(0, 4), "fork");
        FGS_TRY(launch_emit_small(g->P, b.V, views, gw.geom, opt->split_tiles, gw.dsort_k0,
                                  gw.poff, (uint32_t)R, pw.t0, pw.r0, pw.tile16, f.side(0), limit,
                                  gw.counts, EMIT_PREFIX),
                "emit");
        FGS_TRY(launch_large_tiles(true, g->P, b.V, views, gw.geom, gw.large_ids, gw.counts,
                                   nullptr, gw.poff, gw.lrank, gw.lband, (uint32_t)R, pw.t0,
                                   pw.r0, pw.tile16, s, (uint32_t)limit, nullptr),
                "emit large");
        FGS_TRY(f.join(0, 5), "join");
      } else {
        // the redo rarely runs: one stream (a cross-stream wait would cost
        // every frame)
        FGS_TRY(launch_emit_small(g->P, b.V, views, gw.geom, opt->split_tiles, gw.dsort_k0,
                                  gw.poff, (uint32_t)R, pw.t0, pw.r0, pw.tile16, s, limit,
                                  gw.counts, EMIT_REDO),
                "emit");
        FGS_TRY(launch_large_tiles(true, g->P, b.V, views, gw.geom, gw.large_ids, gw.counts,
                                   nullptr, gw.poff, gw.lrank, gw.lband, (uint32_t)R, pw.t0,
                                   pw.r0, pw.tile16, s, 0xffffffffu, gate),
                "emit large");
      }
      // stable sort of (tile id, record) by tile id: the last pass writes the
      // per-tile lists; ranges by binary search over the sorted tile ids
      bool alt = false;
      const SortKind kind = redo ? SORT_TILES_REDO : prefix ? SORT_TILES_PREFIX : SORT_TILES;
      uint32_t* sync = redo ? gw.counts + REDO_SYNC_OFF : nullptr;
      if (pw.tile16)
        FGS_TRY(sort_kv<uint16_t>((uint16_t*)pw.t0, pw.r0, (uint16_t*)pw.t1, pw.r1, R, 0, tb,
                                  pw.scratch, s, alt, n_sort, pw.v0, kind, sync,
                                  gw.counts + COUNT_LIST_FAULT),
                "tile sort");
      else
        FGS_TRY(sort_kv<uint32_t>((uint32_t*)pw.t0, pw.r0, (uint32_t*)pw.t1, pw.r1, R, 0, tb,
                                  pw.scratch, s, alt, n_sort, pw.v0, kind, sync,
                                  gw.counts + COUNT_LIST_FAULT),
                "tile sort");
      FGS_TRY(launch_tile_ranges(R, n_sort, alt ? pw.t1 : pw.t0, pw.tile16, gw.ranges, T,
                                 opt->bucket_size, gw.bcount, gw.maxc, gw.boff,
                                 gw.counts + COUNT_ARRIVE_RANGES, s, gate),
              "ranges and bucket offsets");
      return FGS_OK;
    };
    if ((rc = build_lists(prefix ? EMIT_PREFIX : EMIT_WHOLE))) return rc;
    // the bucket -> (tile, list range) tables are not built here: every pass
    // indexes buckets through boff; fgs_batch_bucket_tables builds them on
    // demand
    c.vals = pw.v0;
    const fgs_tile_sched* ts = opt->tile_sched;
    // this frame's tile features for the next frame's schedule: a small
    // kernel on a side stream that fits beside compositing
    const bool feat_side = ts && ts->features_out;
    if (feat_side) {
      FGS_TRY(f.fork(0, 8), "fork");
      FGS_TRY(launch_tile_features(T, gw.ranges, pw.v0, gw.geom, gw.splat, ts->features_out,
                                   f.side(0)),
              "tile features");
    }
    if (ts) {
      const bool learned = ts->mode == FGS_TILES_LEARNED;
      if (ts->items) {  // a given work list (built by fgs_preprocess or the caller)
        c.items = ts->items;
        c.n_items = ts->n_items;
      } else if (!feat_side) {  // with features_out and no list yet: tile order
        if (learned)
          FGS_TRY(launch_tile_predict(T, gw.ranges, pw.v0, gw.geom, gw.splat, ts->mlp_params,
                                      gw.feat, gw.pred, s),
                  "cost model");
        else if (ts->features)
          FGS_TRY(launch_tile_features(T, gw.ranges, pw.v0, gw.geom, gw.splat, gw.feat, s),
                  "tile features");
        if (learned && ts->order) {
          FGS_TRY(launch_sched_items(T, gw.ranges, opt->bucket_size, gw.pred, gw.items,
                                     gw.counts + 4, s),
                  "work items");
          c.items = gw.items;
          c.n_items = gw.counts + 4;
        } else if (ts->mode != FGS_TILES_RASTER) {
          c.slots = 1;
          c.split_len = ts->split_len;
          c.split_pred = learned ? gw.pred : nullptr;
        }
      }
      if (ts->log_cycles) c.cycles = gw.cycles;
    }
    if (prefix) {
      c.redo = gw.counts + COUNT_REDO;
      c.status = opt->list_status;
    }
    FGS_TRY(launch_composite_fast(c, s), "composite");
    if (feat_side) FGS_TRY(f.join(0, 9), "join");
    if (prefix) {
      // the whole lists and a second compositing of every tile, when some
      // pixel needed more than its prefix list (gated on the device)
      if ((rc = build_lists(EMIT_REDO))) return rc;
      c.redo = nullptr;
      c.status = nullptr;
      c.zero = nullptr;  // done by the first compositing
      c.zero_n = 0;
      c.gate = gw.counts + COUNT_REDO;
      FGS_TRY(launch_composite_fast(c, s), "composite again");
    }
    return FGS_OK;
  }

  // PLAIN: duplicate with keys, 64-bit radix sort, ranges, composite.
  FGS_TRY(launch_duplicate_plain(g->P, gw.geom, gw.dkey, gw.pcount, gw.poff, gx, b.tw, (uint32_t)R,
                                 pw.k0, pw.v0, s),
          "duplicate");
  bool alt = false;
  FGS_TRY(sort_kv<uint64_t>(pw.k0, pw.v0, pw.k1, pw.v1, R, 0, 32 + tile_bits(T), pw.scratch, s,
                            alt, gw.counts, nullptr, SORT_PAIRS, nullptr,
                            gw.counts + COUNT_LIST_FAULT),
          "sort");
  if (R > 0 && alt != plain_sorted_in_alt(T)) return fail(FGS_ERR_HIP, "sort parity");
  const uint64_t* keys = alt ? pw.k1 : pw.k0;
  c.vals = alt ? pw.v1 : pw.v0;
  FGS_TRY(launch_identify_ranges(R, gw.counts, keys, gw.ranges, T, s), "ranges");
  FGS_TRY(launch_composite_plain(c, s), "composite");
  return FGS_OK;
}

int fgs_batch_backward(const fgs_batch* batch, const fgs_gaussians* g, const fgs_options* opt,
                       void* gauss_ws, void* pair_ws, int64_t R, int64_t NB, const int32_t* radii,
                       const float* T_final, const int32_t* last_contributor,
                       const float* dL_dimage, void* grad_ws, const fgs_grads* out,
                       void* stream) {
  Batch b;
  int rc;
  if ((rc = check_opts(opt)) || (rc = make_batch(batch, b, opt, true)) ||
      (rc = check_gauss(g, &b.cams[0])))
    return rc;
  if (!out || !grad_ws || !dL_dimage || !T_final || !last_contributor)
    return fail(FGS_ERR_ARG, "NULL argument");
  if (opt->forward_only)
    return fail(FGS_ERR_ARG, "the frame was rendered forward-only (no checkpoints to differentiate)");
  if (g->P == 0) return FGS_OK;
  const fgs_camera* cam = &b.cams[0];
  const int T = b.T;
  const int64_t PV = (int64_t)g->P * b.V;
  const GaussWs gw = carve_gauss(gauss_ws, PV, T, b.V == 1);
  const bool fast = opt->schedule == FGS_SCHEDULE_FAST;
  const PairWs pw = carve_pair(pair_ws, R, NB, b.HW, T, fast, keeps_ckpt(opt));
  hipStream_t s = (hipStream_t)stream;
  // raw sums: PV conic rows (doubles), then PV rows of the other fields
  // (floats), then the split SH pass's P x 4 floats; cleared beside the
  // active-bucket scan (0.0 is all-zero bits)
  double* raw_con = (double*)grad_ws;
  float* raw_lin = (float*)(raw_con + (size_t)PV * RAW_CON_STRIDE);
  const Fork f(s);
  const bool zero = !opt->grad_ws_zeroed;  // else zeroed by the frame's preprocess
  if (zero) {
    FGS_TRY(f.fork(0, 6), "fork");
    FGS_TRY(fill_f32((float*)raw_con, 0.0f, PV * (2 * RAW_CON_STRIDE + RAW_LIN_STRIDE),
                     f.side(0)),
            "zero raw");
  }
  BwdArgs bw;
  bw.W = cam->width;
  bw.H = cam->height;
  bw.grid_x = grid_x_of(cam->width, b.tw);
  bw.tile_w = b.tw;
  bw.T = T;
  bw.bucket = opt->bucket_size;
  bw.abs_grad = opt->abs_grad;
  bw.NB = NB;
  bw.ranges = gw.ranges;
  bw.splat = gw.splat;
  bw.bg = cam->bg;
  bw.boff = gw.boff;
  bw.maxc = gw.maxc;
  bw.bucket_ranges = pw.bucket_ranges;
  bw.bucket_tile = pw.bucket_tile;
  bw.tfinal = T_final;
  bw.contrib = last_contributor;
  bw.pixc = pw.pixc;
  bw.acc = pw.acc;
  bw.actoff = gw.actoff;
  bw.dimg = dL_dimage;
  bw.raw_lin = raw_lin;
  bw.raw_con = raw_con;
  bw.views = (const ViewRec*)gw.views;
  bw.V = b.V;
  if (fast) {
    bw.vals = pw.v0;
    FGS_TRY(active_bucket_scan(gw.maxc, T, opt->bucket_size, gw.bcount, gw.actoff, gw.scratch, s),
            "active buckets");
    if (zero) FGS_TRY(f.join(0, 7), "join");
    FGS_TRY(launch_backward_fast(bw, s), "raster backward");
  } else {
    bw.vals = plain_sorted_in_alt(T) ? pw.v1 : pw.v0;
    if (zero) FGS_TRY(f.join(0, 7), "join");
    FGS_TRY(launch_backward_plain(bw, s), "raster backward");
  }
  GaussBwdArgs q;
  q.G = make_gauss(g);
  q.cam = make_cam(cam, g, b.tw);
  if (fast) {
    // the camera the forward used: the view table in the workspace
    const ViewRec* v0 = (const ViewRec*)gw.views;
    q.cam.view = v0->view;
    q.cam.proj = v0->proj;
    q.cam.campos = v0->campos;
  }
  q.P = g->P;
  q.raw_lin = raw_lin;
  q.raw_con = raw_con;
  q.radii = radii;
  q.splat = gw.splat;
  q.dmean2D = out->dmean2D;
  q.dconic = out->dconic;
  q.dopacity = out->dopacity;
  q.dcolor = out->dcolor;
  q.dmean3D = out->dmean3D;
  q.dcov3D = out->dcov3D;
  q.ddc = out->ddc;
  q.dsh = q.cam.max_coeffs > 0 ? out->dsh : nullptr;
  q.dscale = out->dscale;
  q.drot = out->drot;
  q.views = fast ? (const ViewRec*)gw.views : nullptr;
  q.V = b.V;
  q.sh_dm = b.V == 1 ? raw_lin + (size_t)PV * RAW_LIN_STRIDE : nullptr;
  q.shjac = b.V == 1 && fast ? gw.shjac : nullptr;
  q.accum = out->xyz_grad_accum;
  q.accum_abs = out->xyz_grad_accum_abs;
  q.max_radii = out->max_radii;
  q.denom = out->denom;
  FGS_TRY(launch_gauss_bwd(q, s), "preprocess backward");
  return FGS_OK;
}

int fgs_fma_probe(float* out, int blocks, int iters, void* stream) {
  if (!out || blocks < 1 || iters < 1) return fail(FGS_ERR_ARG, "bad probe arguments");
  FGS_TRY(launch_fma_probe(out, blocks, iters, (hipStream_t)stream), "fma probe");
  return FGS_OK;
}

int fgs_rcp_selftest(unsigned long long* mismatches, void* stream) {
  if (!mismatches) return fail(FGS_ERR_ARG, "mismatches is NULL");
  FGS_TRY(launch_rcp_selftest(mismatches, (hipStream_t)stream), "rcp selftest");
  return FGS_OK;
}

int fgs_batch_densify_stats(int P, int n_views, const int32_t* radii, const float* dmean2D,
                            float* xyz_grad_accum, float* xyz_grad_accum_abs, float* max_radii,
                            float* denom, void* stream) {
  if (P < 0 || n_views < 1 || (P > 0 && (!radii || !dmean2D)))
    return fail(FGS_ERR_ARG, "bad stats arguments");
  FGS_TRY(launch_densify_stats(P, n_views, radii, dmean2D, xyz_grad_accum, xyz_grad_accum_abs,
                               max_radii, denom, (hipStream_t)stream),
          "densify stats");
  return FGS_OK;
}

int fgs_batch_frame_view_get(void* gauss_ws, void* pair_ws, int P, const fgs_batch* batch,
                             int64_t R, int64_t NB, const fgs_options* opt, fgs_frame_view* v) {
  Batch b;
  int rc;
  if ((rc = check_opts(opt)) || (rc = make_batch(batch, b, opt, false))) return rc;
  if (!v) return fail(FGS_ERR_ARG, "view is NULL");
  const int T = b.T;
  const GaussWs gw = carve_gauss(gauss_ws, (int64_t)P * b.V, T, b.V == 1);
  const bool fast = opt->schedule == FGS_SCHEDULE_FAST;
  const PairWs pw = carve_pair(pair_ws, R, NB, b.HW, T, fast, keeps_ckpt(opt));
  *v = fgs_frame_view{};
  v->geom = gw.geom;
  v->splat = gw.splat;
  v->pair_count = gw.pcount;
  v->tile_count = nullptr;
  v->ranges = gw.ranges;
  v->bucket_off = fast ? gw.boff + 1 : nullptr;
  v->max_contrib = gw.maxc;
  v->counts = gw.counts;
  if (pair_ws) {
    if (fast) {
      v->sorted_keys = nullptr;
      v->sorted_vals = pw.v0;
      v->bucket_ranges = pw.bucket_ranges;
      v->bucket_tile = pw.bucket_tile;
      v->acc_t = nullptr;
      v->acc_c = pw.acc;
    } else {
      const bool alt = plain_sorted_in_alt(T);
      v->sorted_keys = alt ? pw.k1 : pw.k0;
      v->sorted_vals = alt ? pw.v1 : pw.v0;
    }
    v->pixel_colors = pw.pixc;
  }
  if (fast) {
    v->tile_features = gw.feat;
    v->tile_pred = gw.pred;
    v->work_items = gw.items;
    v->item_ticks = gw.cycles;
  }
  return FGS_OK;
}

// Single-view entry points: a batch of one view.
int fgs_preprocess(const fgs_camera* cam, const fgs_gaussians* g, const fgs_options* opt,
                   void* gauss_ws, size_t gauss_ws_bytes, int32_t* radii, void* stream) {
  const fgs_batch b{1, cam};
  return fgs_batch_preprocess(&b, g, opt, gauss_ws, gauss_ws_bytes, radii, stream);
}

int fgs_read_counts(const void* gauss_ws, int P, int width, int height, const fgs_options* opt,
                    int64_t counts_host[2], void* stream) {
  if (!gauss_ws || !counts_host) return fail(FGS_ERR_ARG, "NULL argument");
  const GaussWs w = carve_gauss(const_cast<void*>(gauss_ws), P, tiles_of(width, height, tile_w_of(opt)), true);
  // [0] pairs, [1] bucket-table capacity, up to the depth sort's fault flag
  uint32_t c[DSORT_FAULT + 1];
  hipStream_t s = (hipStream_t)stream;
  FGS_TRY(hipMemcpyAsync(c, w.counts, sizeof(c), hipMemcpyDeviceToHost, s), "read counts");
  FGS_TRY(hipStreamSynchronize(s), "read counts sync");
  if (opt && opt->schedule == FGS_SCHEDULE_FAST && c[DSORT_FAULT])
    return fail(FGS_ERR_HIP, "a look-back wait of the preprocess exceeded its bound");
  if (c[COUNT_LIST_FAULT])
    return fail(FGS_ERR_HIP, "a look-back wait of the list sort exceeded its bound");
  counts_host[0] = c[0];
  counts_host[1] = c[1];
  return FGS_OK;
}

int fgs_render(const fgs_camera* cam, const fgs_gaussians* g, const fgs_options* opt,
               void* gauss_ws, size_t gauss_ws_bytes, void* pair_ws, size_t pair_ws_bytes,
               int64_t R, int64_t NB, float* out_color, float* T_final,
               int32_t* last_contributor, void* stream) {
  const fgs_batch b{1, cam};
  return fgs_batch_render(&b, g, opt, gauss_ws, gauss_ws_bytes, pair_ws, pair_ws_bytes, R, NB,
                          out_color, T_final, last_contributor, stream);
}

int fgs_backward(const fgs_camera* cam, const fgs_gaussians* g, const fgs_options* opt,
                 void* gauss_ws, void* pair_ws, int64_t R, int64_t NB, const int32_t* radii,
                 const float* T_final, const int32_t* last_contributor, const float* dL_dimage,
                 void* grad_ws, const fgs_grads* out, void* stream) {
  const fgs_batch b{1, cam};
  return fgs_batch_backward(&b, g, opt, gauss_ws, pair_ws, R, NB, radii, T_final,
                            last_contributor, dL_dimage, grad_ws, out, stream);
}

int fgs_densify_stats(int P, const int32_t* radii, const float* dmean2D, float* xyz_grad_accum,
                      float* xyz_grad_accum_abs, float* max_radii, float* denom, void* stream) {
  return fgs_batch_densify_stats(P, 1, radii, dmean2D, xyz_grad_accum, xyz_grad_accum_abs,
                                 max_radii, denom, stream);
}

int fgs_batch_bucket_tables(void* gauss_ws, void* pair_ws, int P, const fgs_batch* batch,
                            int64_t R, int64_t NB, const fgs_options* opt, void* stream) {
  Batch b;
  int rc;
  if ((rc = check_opts(opt)) || (rc = make_batch(batch, b, opt, false))) return rc;
  if (opt->schedule != FGS_SCHEDULE_FAST) return fail(FGS_ERR_ARG, "bucket tables need the FAST schedule");
  if (!gauss_ws || !pair_ws) return fail(FGS_ERR_ARG, "NULL workspace");
  if (P < 0 || R < 0 || NB < 0) return fail(FGS_ERR_ARG, "bad frame size");
  if (P == 0 || NB == 0) return FGS_OK;
  const int T = b.T;
  const GaussWs gw = carve_gauss(gauss_ws, (int64_t)P * b.V, T, b.V == 1);
  const PairWs pw = carve_pair(pair_ws, R, NB, b.HW, T, true, keeps_ckpt(opt));
  FGS_TRY(launch_bucket_tables(T, NB, gw.ranges, gw.boff, opt->bucket_size, pw.bucket_ranges,
                               pw.bucket_tile, (hipStream_t)stream),
          "bucket tables");
  return FGS_OK;
}

int fgs_bucket_tables(void* gauss_ws, void* pair_ws, int P, int width, int height, int64_t R,
                      int64_t NB, const fgs_options* opt, void* stream) {
  fgs_camera c{};
  c.width = width;
  c.height = height;
  c.tanfovx = c.tanfovy = 1.0f;
  const fgs_batch b{1, &c};
  return fgs_batch_bucket_tables(gauss_ws, pair_ws, P, &b, R, NB, opt, stream);
}

int fgs_frame_view_get(void* gauss_ws, void* pair_ws, int P, int width, int height, int64_t R,
                       int64_t NB, const fgs_options* opt, fgs_frame_view* v) {
  // only the sizes of the camera matter here
  fgs_camera c{};
  c.width = width;
  c.height = height;
  c.tanfovx = c.tanfovy = 1.0f;
  const fgs_batch b{1, &c};
  return fgs_batch_frame_view_get(gauss_ws, pair_ws, P, &b, R, NB, opt, v);
}

int fgs_render_simp(void* gauss_ws, void* pair_ws, int P, int width, int height, int64_t R,
                    int64_t NB, const fgs_options* opt, float* accum_weights,
                    int32_t* accum_count, float* accum_max_count, void* stream) {
  int rc;
  if ((rc = check_opts(opt)) || (rc = check_whole_lists(opt))) return rc;
  if (P < 0 || width <= 0 || height <= 0 || R < 0) return fail(FGS_ERR_ARG, "bad frame size");
  if (P == 0) return FGS_OK;
  if (!gauss_ws || !pair_ws || !accum_weights || !accum_count || !accum_max_count)
    return fail(FGS_ERR_ARG, "NULL argument");
  const int T = tiles_of(width, height, tile_w_of(opt));
  const GaussWs gw = carve_gauss(gauss_ws, P, T, true);
  const bool fast = opt->schedule == FGS_SCHEDULE_FAST;
  const PairWs pw = carve_pair(pair_ws, R, NB, (int64_t)width * height, T, fast, keeps_ckpt(opt));
  const uint32_t* vals = fast ? pw.v0 : (plain_sorted_in_alt(T) ? pw.v1 : pw.v0);
  hipStream_t s = (hipStream_t)stream;
  FGS_TRY(fill_f32(accum_weights, 0.0f, P, s), "zero weights");
  FGS_TRY(fill_u32((uint32_t*)accum_count, 0u, P, s), "zero counts");
  FGS_TRY(fill_f32(accum_max_count, 0.0f, P, s), "zero max counts");
  FGS_TRY(launch_render_simp(width, height, tile_w_of(opt), T, gw.ranges, vals, gw.splat, accum_weights,
                             accum_count, accum_max_count, s),
          "render_simp");
  return FGS_OK;
}

int fgs_render_depth(void* gauss_ws, void* pair_ws, int P, int width, int height, int64_t R,
                     int64_t NB, const fgs_options* opt, const float* means3D,
                     const float* scales, const float* rotations, const float* projmatrix,
                     const float* campos, float* out_pts, float* out_depth, float* accum_alpha,
                     int32_t* gidx, float* discriminants, void* stream) {
  int rc;
  if ((rc = check_opts(opt)) || (rc = check_whole_lists(opt))) return rc;
  if (P < 0 || width <= 0 || height <= 0 || R < 0) return fail(FGS_ERR_ARG, "bad frame size");
  if (!gauss_ws || !pair_ws || !projmatrix || !campos || !out_pts || !out_depth ||
      !accum_alpha || !gidx || !discriminants)
    return fail(FGS_ERR_ARG, "NULL argument");
  if (P > 0 && (!means3D || !scales || !rotations))
    return fail(FGS_ERR_ARG, "render_depth needs means3D, scales and rotations");
  const int T = tiles_of(width, height, tile_w_of(opt));
  const GaussWs gw = carve_gauss(gauss_ws, P, T, true);
  const bool fast = opt->schedule == FGS_SCHEDULE_FAST;
  const PairWs pw = carve_pair(pair_ws, R, NB, (int64_t)width * height, T, fast, keeps_ckpt(opt));
  const uint32_t* vals = fast ? pw.v0 : (plain_sorted_in_alt(T) ? pw.v1 : pw.v0);
  FGS_TRY(launch_render_depth(width, height, tile_w_of(opt), T, gw.ranges, vals, gw.splat, means3D, scales,
                              rotations, projmatrix, campos, out_pts, out_depth, accum_alpha,
                              gidx, discriminants, (hipStream_t)stream),
          "render_depth");
  return FGS_OK;
}

int fgs_sparse_adam(const fgs_adam_tensor* tensors, int n_tensors, const uint8_t* visible, int N,
                    float beta1, float beta2, void* stream) {
  if (n_tensors < 0 || n_tensors > ADAM_MAX_TENSORS || N < 0 || (n_tensors > 0 && !tensors))
    return fail(FGS_ERR_ARG, "bad Adam tensor list (at most 8 tensors)");
  if (N == 0 || n_tensors == 0) return FGS_OK;
  if (!visible) return fail(FGS_ERR_ARG, "visible is NULL");
  AdamArgs a{};
  a.n = n_tensors;
  a.N = N;
  a.visible = visible;
  a.b1 = beta1;
  a.b2 = beta2;
  for (int t = 0; t < n_tensors; ++t) {
    const fgs_adam_tensor& d = tensors[t];
    if (d.M < 1 || !d.param || !d.grad || !d.exp_avg || !d.exp_avg_sq)
      return fail(FGS_ERR_ARG, "bad Adam tensor");
    a.t[t] = AdamTensor{d.param, d.grad, d.exp_avg, d.exp_avg_sq, d.M, d.lr, d.eps, 0};
  }
  FGS_TRY(launch_sparse_adam(a, (hipStream_t)stream), "sparse Adam");
  return FGS_OK;
}

int fgs_ssim_forward(float C1, float C2, const float* img1, const float* img2, int B, int CH,
                     int H, int W, float* ssim_map, float* dm_dmu1, float* dm_dsigma1_sq,
                     float* dm_dsigma12, void* stream) {
  if (B < 0 || CH < 0 || H < 0 || W < 0) return fail(FGS_ERR_ARG, "bad SSIM shape");
  if ((int64_t)B * CH * H * W == 0) return FGS_OK;
  if (!img1 || !img2 || !ssim_map) return fail(FGS_ERR_ARG, "NULL SSIM argument");
  if ((dm_dmu1 == nullptr) != (dm_dsigma1_sq == nullptr) ||
      (dm_dmu1 == nullptr) != (dm_dsigma12 == nullptr))
    return fail(FGS_ERR_ARG, "give all three SSIM derivative maps or none");
  if (H > 65535 * 16 || B > 65535) return fail(FGS_ERR_ARG, "SSIM image too large");
  FGS_TRY(launch_ssim_fwd(B, CH, H, W, C1, C2, img1, img2, ssim_map, dm_dmu1, dm_dsigma1_sq,
                          dm_dsigma12, (hipStream_t)stream),
          "SSIM forward");
  return FGS_OK;
}

int fgs_ssim_backward(const float* img1, const float* img2, const float* dL_dmap,
                      const float* dm_dmu1, const float* dm_dsigma1_sq, const float* dm_dsigma12,
                      int B, int CH, int H, int W, float* dL_dimg1, void* stream) {
  if (B < 0 || CH < 0 || H < 0 || W < 0) return fail(FGS_ERR_ARG, "bad SSIM shape");
  if ((int64_t)B * CH * H * W == 0) return FGS_OK;
  if (!img1 || !img2 || !dL_dmap || !dm_dmu1 || !dm_dsigma1_sq || !dm_dsigma12 || !dL_dimg1)
    return fail(FGS_ERR_ARG, "NULL SSIM argument");
  if (H > 65535 * 16 || B > 65535) return fail(FGS_ERR_ARG, "SSIM image too large");
  FGS_TRY(launch_ssim_bwd(B, CH, H, W, img1, img2, dL_dmap, dm_dmu1, dm_dsigma1_sq, dm_dsigma12,
                          dL_dimg1, (hipStream_t)stream),
          "SSIM backward");
  return FGS_OK;
}

int fgs_mlp_forward(const float* params, const float* x, int64_t n, float* y, int bf16,
                    void* stream) {
  if (n < 0 || (n > 0 && (!params || !x || !y))) return fail(FGS_ERR_ARG, "bad MLP arguments");
  FGS_TRY(launch_mlp_infer(params, x, n, y, bf16 != 0, (hipStream_t)stream), "MLP forward");
  return FGS_OK;
}

size_t fgs_mlp_grad_ws_size(int64_t n) { return mlp_grad_ws_bytes(n); }

int fgs_mlp_loss_grad(const float* params, const float* x, const float* target, int64_t n,
                      float split_weight, float* grad, float* loss, void* ws, size_t ws_bytes,
                      void* stream) {
  if (n < 0 || !params || !grad || (n > 0 && (!x || !target)))
    return fail(FGS_ERR_ARG, "bad MLP arguments");
  if (n > 0 && (!ws || ws_bytes < mlp_grad_ws_bytes(n)))
    return fail(FGS_ERR_CAPACITY, "MLP workspace too small");
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) {
    FGS_TRY(fill_f32(grad, 0.0f, MLP_NPARAMS, s), "MLP zero grad");
    if (loss) FGS_TRY(fill_f32(loss, 0.0f, 1, s), "MLP zero loss");
    return FGS_OK;
  }
  FGS_TRY(launch_mlp_loss_grad(params, x, target, n, split_weight, grad, loss, ws, s), "MLP grad");
  return FGS_OK;
}

int fgs_adam_update(float* params, const float* grad, float* exp_avg, float* exp_avg_sq,
                    int64_t n, float lr, float beta1, float beta2, float eps, int step,
                    void* stream) {
  if (n < 0 || step < 1 || (n > 0 && (!params || !grad || !exp_avg || !exp_avg_sq)))
    return fail(FGS_ERR_ARG, "bad Adam arguments");
  FGS_TRY(launch_adam(params, grad, exp_avg, exp_avg_sq, n, lr, beta1, beta2, eps, step,
                      (hipStream_t)stream),
          "Adam");
  return FGS_OK;
}

}  // extern "C"
