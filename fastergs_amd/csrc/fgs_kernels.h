// Host-side launchers shared between the .hip translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fgs_layout.h"
#include "fgs_math.h"

namespace fgs {

// ------------------------------------------------------------------ views
// The FAST schedule renders a batch of V views in one launch per stage.
// Per-Gaussian records are indexed view-major (record v * P + i for
// Gaussian i in view v); the tiles of all views form one index space (view
// v owns tiles [tile_off, tile_off + grid_x * grid_y)), and per-pixel
// arrays are the views' images concatenated (view v from pixel pix_off;
// colour planes at 3 * pix_off). A single view is a batch of one.
constexpr int MAX_VIEWS = 32;  // ViewSetup travels as kernel arguments (< 4 KB)
struct ViewRec {
  float view[16];    // world -> view, column-major
  float proj[16];    // world -> clip, column-major
  float campos[4];
  float bg[4];
  float tanfovx, tanfovy, focal_x, focal_y;
  int W, H, grid_x, grid_y;
  uint32_t tile_off, pix_off;
  int tile_w, tile_h;  // tile shape (tile_w * tile_h = TILE_PIX)
};
struct ViewSrc {     // host-side description; matrices are device pointers
  const float* view;
  const float* proj;
  const float* campos;
  const float* bg;
  float tanfovx, tanfovy, focal_x, focal_y;
  int W, H;
  uint32_t tile_off, pix_off;
  int tile_w, tile_h;
};
struct ViewSetup {
  ViewSrc v[MAX_VIEWS];
  int n;
};
// Copies the camera data into the device table (kernel arguments only, so
// a captured graph re-reads the camera buffers on every replay) and clears
// the frame's counters (zero_n words at zero).
hipError_t launch_views_setup(const ViewSetup& src, ViewRec* out, uint32_t* zero, int64_t zero_n,
                              hipStream_t s);

// Camera of view r for the shared per-Gaussian math.
__device__ __forceinline__ CamRef<float> view_cam(const CamRef<float>& base, const ViewRec& r) {
  CamRef<float> c = base;
  c.view = r.view;
  c.proj = r.proj;
  c.campos = r.campos;
  c.tanfovx = r.tanfovx;
  c.tanfovy = r.tanfovy;
  c.focal_x = r.focal_x;
  c.focal_y = r.focal_y;
  c.W = r.W;
  c.H = r.H;
  c.grid_x = r.grid_x;
  c.grid_y = r.grid_y;
  c.tile_w = r.tile_w;
  c.tile_h = r.tile_h;
  return c;
}
// Camera of view source s (caller's buffers) for the shared per-Gaussian math.
__device__ __forceinline__ CamRef<float> src_cam(const CamRef<float>& base, const ViewSrc& s) {
  CamRef<float> c = base;
  c.view = s.view;
  c.proj = s.proj;
  c.campos = s.campos;
  c.tanfovx = s.tanfovx;
  c.tanfovy = s.tanfovy;
  c.focal_x = s.focal_x;
  c.focal_y = s.focal_y;
  c.W = s.W;
  c.H = s.H;
  c.grid_x = grid_dim(s.W, s.tile_w);
  c.grid_y = grid_dim(s.H, s.tile_h);
  c.tile_w = s.tile_w;
  c.tile_h = s.tile_h;
  return c;
}
// View owning global tile t (binary search over the views' first tiles).
__device__ __forceinline__ int view_of_tile(const ViewRec* views, int V, uint32_t t) {
  int lo = 0, hi = V;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (views[mid].tile_off <= t) lo = mid; else hi = mid;
  }
  return lo;
}

// scan_sort.hip
hipError_t scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, void* ws,
                    hipStream_t s);
// Pair offsets of the FAST preprocess by depth rank r, from the depth
// sort's keys (depth bits << 32 | record) and values (the record's pair
// count, or PAIRS_LARGE for a large rect: its count is read from count[] and
// its rank written to lrank[record]): off[r] = exclusive prefix over ranks,
// counts[0] = the total, counts[1] = ceil(total / bucket) + T. vals is
// overwritten with the resolved counts.
constexpr uint32_t PAIRS_LARGE = 0x80000000u;
hipError_t scan_pair_offsets(const uint64_t* keys, uint32_t* vals, const uint32_t* count,
                             uint32_t* lrank, int64_t n, uint32_t* off, uint32_t* counts,
                             int bucket, int T, void* ws, hipStream_t s);
// ids = i * LARGE_BANDS + b for every index i < n and b < flag[i] (flags
// 0..LARGE_BANDS), ascending; *count = their number (on the device).
hipError_t expand_flags(const uint32_t* flag, int64_t n, uint32_t* ids, uint32_t* count, void* ws,
                        hipStream_t s);
// Stable LSD radix sort of (key, value) pairs on key bits [begin_bit,
// end_bit) (keys of 16, 32 or 64 bits; vals may be NULL for key-only sorts).
// sync: when given, SORT_SYNC_WORDS words that are zero on entry, used for
// the onesweep state instead of clearing it in a launch of its own.
// Passes alternate between the arrays and their *_alt twins; in_alt tells
// where the keys ended up. last_vals: when given, the final pass writes the
// values there. n_dev: when given, the element count is min(*n_dev, n) read
// on the device (n is then the capacity the grids are sized for); no host
// synchronisation either way. The onesweep passes bound their look-back
// waits (a wait cannot exceed the bound when tiles are claimed in launch
// order); a wait that does sets a fault word to 1: fault_word when given
// (the caller clears it), else the last word of the sort's own state
// (sort_fault_word).
enum SortKind {
  SORT_PAIRS = 0,
  SORT_DEPTH = 1,
  SORT_TILES = 2,
  SORT_TILES_REDO = 3,  // the redo build of prefix-list frames
  SORT_TILES_PREFIX = 4  // the prefix lists
};
template <typename K>
hipError_t sort_kv(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, int64_t n,
                   int begin_bit, int end_bit, void* ws, hipStream_t s, bool& in_alt,
                   const uint32_t* n_dev, uint32_t* last_vals, SortKind kind,
                   uint32_t* sync = nullptr, uint32_t* fault_word = nullptr);
// The same sort with three-launch passes only (histogram, scan, stable
// scatter per digit; no look-back waits), for any n. Its workspace is
// sort_classic_ws_bytes(n), every word of it written before it is read.
template <typename K>
hipError_t sort_kv_passes(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, int64_t n,
                          int begin_bit, int end_bit, void* ws, hipStream_t s, bool& in_alt);
// fill.hip-style helpers (render.hip)
hipError_t fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
hipError_t fill_f32(float* p, float v, int64_t n, hipStream_t s);
// 8 x iters FMAs per thread of blocks x 256 threads; out: blocks x 256 floats.
hipError_t launch_fma_probe(float* out, int blocks, int iters, hipStream_t s);

// preprocess.hip
struct PreArgs {
  GaussRef<float> G;
  CamRef<float> cam;
  float* geom;
  float* splat;
  uint32_t* dkey;
  uint32_t* pcount;
  int32_t* radii;
  int P;
  int split_tiles;
  uint32_t* lflag;     // fast: row bands of a large rect (walked by whole waves), 0 if small
  uint64_t* dsort_k;   // fast: depth sort keys, depth bits << 32 | record
  uint32_t* dsort_v;   // fast: depth sort values, small-rect pair count or PAIRS_LARGE
  uint32_t* counts;
  const ViewRec* views;  // fast: V views (records are V x P)
  int V;
  float* shjac;        // fast, one view: the colour pass stores each Gaussian's SH Jacobian
                       // (P x 12 floats: rows J[ch][0..2], 0) for the backward; nullptr: not
};
hipError_t launch_preprocess_plain(const PreArgs& a, hipStream_t s);
// Geometry pass of the FAST schedule; also writes the view table from vs
// and clears zero_n (< 2^31) words at zero. With P = 0 only the table and
// the clearing are done.
hipError_t launch_preprocess_fast(const PreArgs& a, const ViewSetup& vs, ViewRec* table,
                                  uint32_t* zero, int zero_n, hipStream_t s);
// SH colour of every valid record (after launch_preprocess_fast; only the
// compositing pass reads its output). No-op with precomputed colours.
hipError_t launch_sh_color(const PreArgs& a, hipStream_t s);
// Pair emission writes only slots below cap (the pair capacity); a frame
// whose pair count exceeds it is detected on the host from the count.
hipError_t launch_duplicate_plain(int P, const float* geom, const uint32_t* dkey,
                                  const uint32_t* pcount, const uint32_t* poff, int grid_x,
                                  int tw, uint32_t cap, uint64_t* keys, uint32_t* vals,
                                  hipStream_t s);
// FAST emission over V views of P Gaussians (records v * P + i) as (tile
// id, record) into two arrays: tile ids are global, 16-bit when t16 (at
// most 65536 tiles), else 32-bit.
// poff: pair offsets by depth rank; emission reads order (the sorted depth
// keys, record in the low 32 bits) for the small rects and lrank (the rank
// of each large record) for the large ones.
// Emission of the records of depth rank < rank_limit only (prefix lists);
// gate: when given, the launch does nothing unless gate[0] != 0.
hipError_t launch_large_tiles(bool emit, int P, int V, const ViewRec* views, const float* geom,
                              const uint32_t* large_ids, const uint32_t* counts, uint32_t* pcount,
                              const uint32_t* poff, const uint32_t* lrank, uint32_t* lband,
                              uint32_t cap, void* tkeys, uint32_t* recs, bool t16, hipStream_t s,
                              uint32_t rank_limit = 0xffffffffu, const uint32_t* gate = nullptr);
// Small-rect emission of the ranks [0, limit) (limit <= P * V). mode:
// EMIT_WHOLE (every rank; sets counts[COUNT_PREFIX_N], [COUNT_REDO],
// [COUNT_REDO_N] to 0), EMIT_PREFIX (limit < P * V: counts[COUNT_PREFIX_N] =
// poff[limit], the others 0) or EMIT_REDO (the launch does nothing unless
// counts[COUNT_REDO] is set; writes no counts).
enum EmitMode { EMIT_WHOLE = 0, EMIT_PREFIX = 1, EMIT_REDO = 2 };
hipError_t launch_emit_small(int P, int V, const ViewRec* views, const float* geom, int split,
                             const uint64_t* order, const uint32_t* poff, uint32_t cap,
                             void* tkeys, uint32_t* recs, bool t16, hipStream_t s,
                             int64_t limit, uint32_t* counts, EmitMode mode);

// render.hip
// Range kernels take the pair count as (capacity R, device count n_dev).
hipError_t launch_identify_ranges(int64_t R, const uint32_t* n_dev, const uint64_t* keys,
                                  uint32_t* ranges, int T, hipStream_t s);
hipError_t launch_bucket_tables(int T, int64_t cap, const uint32_t* ranges, const uint32_t* boff,
                                int bucket, uint32_t* bucket_ranges, uint32_t* bucket_tile,
                                hipStream_t s);
// Ranges from the sorted tile ids, plus per-tile bucket counts and cleared
// furthest-contributor slots (the FAST schedule's bucket_counts).
// Ranges, bucket counts and (scanned by the last workgroup) the bucket
// offsets boff; arrive: a zero counter word the launch leaves at zero.
// gate: when given, the launch does nothing unless gate[0] != 0.
hipError_t launch_tile_ranges(int64_t R, const uint32_t* n_dev, const void* tiles, bool t16,
                              uint32_t* ranges, int T, int bucket, uint32_t* bcount,
                              uint32_t* maxc, uint32_t* boff, uint32_t* arrive, hipStream_t s,
                              const uint32_t* gate = nullptr);
// Exclusive prefix of the backward's active buckets ceil(maxc[t] / bucket)
// into out[0..T), the total into out[T] (one launch up to 8192 tiles, else
// the counts into `counts` and a scan).
hipError_t active_bucket_scan(const uint32_t* maxc, int T, int bucket, uint32_t* counts,
                              uint32_t* out, void* ws, hipStream_t s);
hipError_t launch_active_counts(int T, const uint32_t* maxc, int bucket, uint32_t* out,
                                hipStream_t s);
struct CompArgs {
  int W, H, grid_x, T, bucket;
  int tile_w;  // tile shape of the PLAIN kernel (FAST reads it from the view records)
  // the frame's counts words: when a sort's look-back bound was exceeded
  // (counts[COUNT_LIST_FAULT] or counts[DSORT_FAULT] set) the lists are not
  // valid, and images and T_final are written NaN
  const uint32_t* counts;
  const uint32_t* ranges;
  const uint32_t* vals;
  const float* splat;
  const float* bg;
  const uint32_t* boff;  // exclusive bucket offsets (fast)
  float* out;
  float* tfinal;
  int32_t* contrib;
  float* pixc;
  uint32_t* maxc;
  float* acc;            // fast checkpoints
  // fast: optional work list (tile << 2 | part: 0 whole tile, 1 / 2 upper /
  // lower 16x8 band) of n_items[0] <= 2T items; or, with slots set, two grid
  // slots per tile in tile order, tiles split when split_pred[2t + 1] > 0
  // (if given) or their list has >= split_len (> 0) entries; optional
  // per-workgroup duration log
  const uint32_t* items;
  const uint32_t* n_items;
  int slots;
  int split_len;
  const float* split_pred;
  uint32_t* cycles;
  // fast: the work-queue counter of the persistent form (zero at launch, left
  // zero), or NULL for one workgroup per grid index
  uint32_t* queue;
  const ViewRec* views;  // fast: tiles / pixels of V views (W, H, grid_x, bg per view)
  int V;
  // fast, prefix lists: redo = counts + COUNT_REDO; a workgroup with a pixel
  // that its list leaves open (not stopped at the end of a whole bucket)
  // sets redo[0] = 1 and redo[1] = counts[0], and status[0] = 1 when given
  uint32_t* redo;
  uint32_t* status;
  // fast: when given, the launch does nothing unless gate[0] != 0
  const uint32_t* gate;
  // fast: zero_n floats (a multiple of 4, 16-byte aligned) to zero, spread
  // over the grid's threads (the backward's raw sums), or NULL
  float* zero;
  int64_t zero_n;
};
hipError_t launch_composite_plain(const CompArgs& a, hipStream_t s);
hipError_t launch_composite_fast(const CompArgs& a, hipStream_t s);

hipError_t launch_render_simp(int W, int H, int tw, int T, const uint32_t* ranges,
                              const uint32_t* vals,
                              const float* splat, float* wsum, int32_t* wcount, float* maxcount,
                              hipStream_t s);
hipError_t launch_render_depth(int W, int H, int tw, int T, const uint32_t* ranges,
                               const uint32_t* vals,
                               const float* splat, const float* means3D, const float* scales,
                               const float* rotations, const float* proj, const float* campos,
                               float* out_pts, float* out_depth, float* accum_alpha,
                               int32_t* gidx, float* disc, hipStream_t s);

// sched.hip
hipError_t launch_tile_features(int T, const uint32_t* ranges, const uint32_t* vals,
                                const float* geom, const float* splat, float* feat,
                                hipStream_t s);
// Tile features (into feat) and the cost model's bf16 prediction (into
// pred) in one launch: what launch_tile_features + launch_mlp_infer(bf16)
// compute, to float rounding of the feature sums.
hipError_t launch_tile_predict(int T, const uint32_t* ranges, const uint32_t* vals,
                               const float* geom, const float* splat, const float* params,
                               float* feat, float* pred, hipStream_t s);
hipError_t launch_mlp_infer(const float* params, const float* x, int64_t n, float* y, bool bf16,
                            hipStream_t s);
size_t mlp_grad_ws_bytes(int64_t n);
hipError_t launch_mlp_loss_grad(const float* params, const float* x, const float* target,
                                int64_t n, float split_weight, float* grad, float* loss, void* ws,
                                hipStream_t s);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr,
                       float b1, float b2, float eps, int step, hipStream_t s);
hipError_t launch_sched_items(int T, const uint32_t* ranges, int bucket, const float* pred,
                              uint32_t* items, uint32_t* n_items, hipStream_t s);

// optim.hip: visibility-masked Adam over up to ADAM_MAX_TENSORS tensors
constexpr int ADAM_MAX_TENSORS = 8;
struct AdamTensor {
  float* param;
  const float* grad;
  float* exp_avg;
  float* exp_avg_sq;
  int M;      // parameters per Gaussian
  float lr;   // bias-compensated step size
  float eps;
  int vec;    // set by the launcher: all four arrays 16-byte aligned
};
struct AdamArgs {
  AdamTensor t[ADAM_MAX_TENSORS];
  int64_t block_off[ADAM_MAX_TENSORS];  // first workgroup of each tensor (launcher)
  int n;
  int N;
  const uint8_t* visible;
  float b1, b2;
};
hipError_t launch_sparse_adam(AdamArgs& a, hipStream_t s);

// ssim.hip
hipError_t launch_ssim_fwd(int B, int CH, int H, int W, float C1, float C2, const float* img1,
                           const float* img2, float* map, float* dmu1, float* ds1, float* ds12,
                           hipStream_t s);
hipError_t launch_ssim_bwd(int B, int CH, int H, int W, const float* img1, const float* img2,
                           const float* dmap, const float* dmu1, const float* ds1,
                           const float* ds12, float* dimg1, hipStream_t s);

// backward.hip
struct BwdArgs {
  int W, H, grid_x, T, bucket, abs_grad;
  int tile_w;  // tile shape of the PLAIN kernel (FAST reads it from the view records)
  int64_t NB;
  const uint32_t* ranges;
  const uint32_t* vals;
  const float* splat;
  const float* bg;
  const uint32_t* boff;
  const uint32_t* maxc;
  const uint32_t* bucket_ranges;
  const uint32_t* bucket_tile;
  const float* tfinal;
  const int32_t* contrib;
  const float* pixc;
  const float* acc;
  const uint32_t* actoff;  // fast: exclusive prefix of active buckets per tile, [T] = total
  const float* dimg;
  // raw sums of the P (V x P) records: float rows of RAW_LIN_STRIDE (every
  // field but the conic ones) and double rows of RAW_CON_STRIDE (conic)
  float* raw_lin;
  double* raw_con;
  const ViewRec* views;  // fast: V views
  int V;
};
hipError_t launch_backward_plain(const BwdArgs& a, hipStream_t s);
// Counts the floats in [2^-7, 1] whose backward reciprocal is not the IEEE
// quotient (bad: 1 counter, cleared by the launch).
hipError_t launch_rcp_selftest(unsigned long long* bad, hipStream_t s);
hipError_t launch_backward_fast(const BwdArgs& a, hipStream_t s);
// Per-Gaussian adjoint. With views (V >= 1 of them, records v * P + i):
// dmean2D / dconic are per view (V x P rows), every other gradient is the
// sum over the views (fixed order). Without views: the single camera cam.
struct GaussBwdArgs {
  GaussRef<float> G;
  CamRef<float> cam;
  int P;
  const float* raw_lin;  // as in BwdArgs
  const double* raw_con;
  const int32_t* radii;
  const float* splat;
  float *dmean2D, *dconic, *dopacity, *dcolor, *dmean3D, *dcov3D, *ddc, *dsh, *dscale, *drot;
  const ViewRec* views;
  int V;
  float* sh_dm;  // one view: P x 4 scratch for the split SH pass (nullptr: fused kernel)
  // one view: the SH Jacobians the forward's colour pass stored (P x 12
  // floats: rows J[ch][0..2] and, in the first row's fourth word, the
  // colour clamp bits), used instead of the coefficients (nullptr: none)
  const float* shjac;
  // optional densification statistics added by the adjoint pass (one view)
  float *accum, *accum_abs, *max_radii, *denom;
};
hipError_t launch_gauss_bwd(const GaussBwdArgs& a, hipStream_t s);
// Densification statistics of V views (radii / dmean2D rows v * P + i) added
// in view order into the P-sized statistics.
hipError_t launch_densify_stats(int P, int V, const int32_t* radii, const float* dmean2D,
                                float* accum, float* accum_abs, float* max_radii, float* denom,
                                hipStream_t s);

}  // namespace fgs
