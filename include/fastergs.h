/* C ABI of the gfx950 Gaussian-splatting rasterizer (libfastergs.so).
 *
 * This is the boundary that replaces the reference backend's torch binding
 * (FasterGS gs_submodules/diff-gaussian-rasterization_faster/csrc/pybind.cpp,
 * the `_C.ops` module of :371-416). Every pointer is a device pointer unless
 * stated otherwise; no torch types cross it. Entry points never allocate:
 * callers pass workspaces sized by the *_ws_size queries, and every call
 * takes the HIP stream to run on. Calls on distinct workspaces are
 * independent (no device globals). All functions return FGS_OK or a negative
 * fgs_status; fgs_last_error() gives a message for the calling thread.
 *
 * Forward sequence (mirrors _RasterizeGaussians.forward,
 * diff_gaussian_rasterization_faster/__init__.py:80-155):
 *   fgs_preprocess -> fgs_read_counts (the one host sync: pairs R, buckets NB)
 *   -> fgs_render.
 * fgs_render and fgs_backward take R and NB as capacities and read the
 * actual counts on the device, so with fixed capacities (e.g. from a
 * previous frame plus headroom) preprocess + render + backward contain no
 * host synchronisation and can be captured into one hipGraph; camera
 * matrices are read through device pointers at run time, so a replay after
 * rewriting them renders the new view. Such a frame is valid iff the counts
 * that fgs_read_counts returns afterwards are <= the capacities used
 * (pairs beyond the capacity are dropped, never written out of bounds).
 * Backward (__init__.py:159-187): fgs_backward, then optionally
 * fgs_densify_stats.
 */
#ifndef FASTERGS_H
#define FASTERGS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum fgs_status {
  FGS_OK = 0,
  FGS_ERR_ARG = -1,      /* invalid argument or shape */
  FGS_ERR_HIP = -2,      /* a HIP runtime call failed */
  FGS_ERR_CAPACITY = -3, /* workspace too small for the data */
} fgs_status;

/* Schedules. PLAIN: one thread per Gaussian, scan + duplicate, 64-bit key
 * LSD radix sort, per-tile sequential compositing, per-pixel reverse
 * backward. FAST: wave-packed / split preprocess, depth-presorted emission
 * with a tile-only radix sort, bucketed compositing with per-tile bucket
 * cutoff, bucket-checkpointed backward with wave reductions. */
enum { FGS_SCHEDULE_PLAIN = 0, FGS_SCHEDULE_FAST = 1 };

typedef struct fgs_camera {
  int width, height;
  float tanfovx, tanfovy;
  float scale_modifier;
  float mult; /* power multiplier of GaussianRasterizationSettings.mult */
  int sh_degree;
  const float* viewmatrix; /* 16, column-major world->view (3DGS world_view_transform) */
  const float* projmatrix; /* 16, column-major full projection */
  const float* campos;     /* 3 */
  const float* bg;         /* 3 */
} fgs_camera;

typedef struct fgs_gaussians {
  int P;
  int max_coeffs;             /* rest SH coefficients per Gaussian (shs.size(1)) */
  const float* means3D;       /* P x 3 */
  const float* opacities;     /* P (activated) */
  const float* scales;        /* P x 3 (activated) */
  const float* rotations;     /* P x 4 (w,x,y,z), used as given */
  const float* dc;            /* P x 3, required with shs */
  const float* shs;           /* P x max_coeffs x 3 or NULL */
  const float* colors_precomp;/* P x 3 or NULL (then SH is used) */
  const uint8_t* culling;     /* P bool or NULL */
} fgs_gaussians;

/* Work schedule of the FAST compositing pass (outputs never depend on it).
 * RASTER: one workgroup per tile in tile order. HEURISTIC: tile order, tiles
 * whose list has >= split_len entries are composited as two 16x8 row bands
 * by two workgroups (two grid slots per tile, decided in the kernel).
 * LEARNED: a cost model (MLP 8 -> 64 -> 64 -> 2, bf16 inference) predicts
 * each tile's log2 duration and a split flag from 8 features of its list;
 * the split flags are applied in tile order like HEURISTIC, or with order
 * set the items are dispatched longest-predicted first from a work list. */
enum { FGS_TILES_RASTER = 0, FGS_TILES_HEURISTIC = 1, FGS_TILES_LEARNED = 2 };
#define FGS_MLP_IN 8
#define FGS_MLP_HIDDEN 64
#define FGS_MLP_OUT 2
/* W1[64][8], b1[64], W2[64][64], b2[64], W3[2][64], b3[2] (fp32) */
#define FGS_MLP_NPARAMS 4866
typedef struct fgs_tile_sched {
  int mode;                /* FGS_TILES_* */
  int split_len;           /* HEURISTIC split threshold in list entries (<= 0: never split) */
  const float* mlp_params; /* LEARNED: FGS_MLP_NPARAMS floats (device) */
  int features;            /* compute the tile features even when not LEARNED */
  int log_cycles;          /* log each workgroup's duration (100 MHz ticks) */
  int order;               /* LEARNED: dispatch longest-predicted first */
  /* A work list for fgs_render to dispatch (any mode; the outputs do not
     depend on it): n_items[0] <= 2 T entries of (tile << 2 | part), part 0 a
     whole tile, 1 / 2 its upper / lower 16x8 row band, every tile of the
     frame exactly once (whole or as both bands). Device memory; NULL: the
     schedule decides as above. */
  uint32_t* items;
  uint32_t* n_items;
  /* LEARNED with items: the tile features (T x 8 floats, device) of the
     previous frame of the same size; fgs_preprocess builds the work list
     from them (cost model, longest-predicted first, split flags) into
     items / n_items on a side stream beside the preprocess, so no
     prediction runs on the frame's critical path (on a camera trajectory
     consecutive views are alike). NULL: none. */
  const float* prev_features;
  /* fgs_render writes this frame's tile features (T x 8 floats, device)
     here, computed on a side stream beside compositing (for the next
     frame's prev_features). NULL: none. Side streams are joined before
     the entry points return. */
  float* features_out;
} fgs_tile_sched;

typedef struct fgs_options {
  int schedule;     /* FGS_SCHEDULE_* */
  int bucket_size;  /* 16, 32, 64, 128 or 256 */
  int split_tiles;  /* fast preprocess: rects with >= this many tiles are split */
  int abs_grad;     /* backward accumulates |dL/dmean2D| terms (needed for xyz_gradient_accum_abs) */
  const fgs_tile_sched* tile_sched; /* FAST compositing schedule; NULL = RASTER */
  int forward_only; /* the frame will not be differentiated: compositing stores no bucket
                       checkpoints or per-pixel colours and the pair workspace holds none
                       (pass the same value to every call on the frame; fgs_backward then
                       returns FGS_ERR_ARG) */
  int composite_queue; /* FAST compositing as a persistent grid (the workgroups resident at
                          once) taking work items from a device counter, instead of one
                          workgroup per item; the outputs are the same */
  int tile_w;       /* tile shape: tiles of 256 pixels, tile_w x (256 / tile_w) with tile_w 8,
                       16 or 32; 0 = 16, the reference's 16 x 16 (BLOCK_X x BLOCK_Y). Every
                       call on a frame takes the same value. The per-tile lists, and with them
                       the FAST bucket boundaries (where the T < 1e-4 cut-off is tested) and
                       render_simp's per-tile stop, are defined on the tile; the PLAIN images
                       do not depend on it. T (tiles) = ceil(W / tile_w) * ceil(H / tile_h). */
  int list_prefix;  /* FAST, no tile_sched: > 0 renders the frame first from prefix lists, the
                       per-tile lists of its list_prefix nearest records (depth ranks below
                       list_prefix, < P x views; a batch's depth order interleaves its views'
                       records and keeps each view's own order): emission, the list sort and
                       the ranges handle their pairs only. The lists are in depth order, so a prefix list is the
                       front of the tile's whole list, and a pixel that stops at the end of a
                       whole bucket within it has its final values. When some pixel of some tile
                       is left open, the whole lists are built and every tile is composited
                       again, on the device (no host wait; launches that find nothing to do
                       return at once). The images, T_final, last contributors, checkpoints
                       and gradients never depend on it. The frame's lists and ranges
                       (fgs_frame_view_get) are the prefix lists when they sufficed: to rebuild
                       whole lists, call fgs_render again on the frame with list_prefix 0 (the
                       outputs keep their values). fgs_render_simp and fgs_render_depth, which
                       read whole lists, return FGS_ERR_ARG for options with list_prefix > 0.
                       0 = whole lists. */
  void* zero_grad_ws; /* fgs_render (FAST): when non-NULL, the raw sums of the gradient
                         workspace the frame's fgs_backward will take
                         (fgs_batch_grad_ws_size(P, views) bytes) are zeroed by compositing,
                         whose time is VALU issue, instead of by a fill at the start of the
                         backward. No counterpart in the reference (which clears its
                         gradient buffers in the backward). */
  int grad_ws_zeroed; /* fgs_backward: the grad_ws it takes was zeroed by this frame's
                         fgs_render (zero_grad_ws) and not used since; it then skips its own
                         zeroing */
  uint32_t* list_status; /* with list_prefix: device-accessible word (device memory, or mapped
                            pinned host memory read after the frame completes) that compositing
                            sets to 1 when the prefix lists did not suffice; never cleared by
                            the library. NULL: none. */
} fgs_options;

/* Device pointers into the two workspaces, for inspection by tests and
 * for callers that want the intermediate products. Arrays not produced by
 * the chosen schedule are NULL. */
typedef struct fgs_frame_view {
  float* geom;            /* P x 8: x, y, conic a, b, c, power, rect min (x | y<<16), rect max */
  float* splat;           /* P x 12: x, y, ca, cb, cc, log2 op, depth, 0, r, g, b, clamp bits */
  uint32_t* pair_count;   /* P */
  uint32_t* tile_count;   /* T */
  uint32_t* ranges;       /* T x 2 */
  uint32_t* bucket_off;   /* T, inclusive prefix of per-tile bucket counts */
  uint32_t* max_contrib;  /* T */
  uint32_t* counts;       /* [0] pairs R, [1] buckets NB */
  uint64_t* sorted_keys;  /* R, tile<<32 | depth bits (PLAIN) */
  uint32_t* sorted_vals;  /* R, per-tile Gaussian lists */
  uint32_t* bucket_ranges;/* NB x 2 */
  uint32_t* bucket_tile;  /* NB */
  float* acc_t;           /* NB x 256 transmittance after each bucket (FAST) */
  float* acc_c;           /* NB x 256 x 4 colour after each bucket (FAST) */
  float* pixel_colors;    /* H x W x 4 */
  float* tile_features;   /* T x 8 (FAST, when computed) */
  float* tile_pred;       /* T x 2 cost-model outputs (LEARNED) */
  uint32_t* work_items;   /* LEARNED with order: counts[4] items, tile << 2 | part
                             (0 whole, 1 / 2 upper / lower band) */
  uint32_t* item_ticks;   /* per-workgroup duration, 100 MHz ticks (log_cycles): by
                             work item, by grid slot (2 per tile) or by tile (RASTER) */
} fgs_frame_view;

/* Workspace sizes in bytes. The gauss workspace depends on P and the image;
 * the pair workspace additionally on R and NB from fgs_read_counts. */
size_t fgs_gauss_ws_size(int P, int width, int height, const fgs_options* opt);
size_t fgs_pair_ws_size(int P, int width, int height, int64_t R, int64_t NB,
                        const fgs_options* opt);
/* Gradient workspace of the backward. It starts with the raw per-Gaussian
 * sums, readable after fgs_backward: P rows of 4 doubles (conic sums 3, 1
 * unused; added order-independently) then P rows of 8 floats (colour 3,
 * mean2D 2, power 1, |mean2D| 2); then P x 4 floats of scratch. */
size_t fgs_grad_ws_size(int P);

/* Replaces faster::preprocess (preprocess.cu:928-962, pybind.cpp:26-76):
 * projection, covariances, SH colour, tile rects, exact tile/ellipse pair
 * counts and the pair offsets. Writes every entry of radii (P int32). */
int fgs_preprocess(const fgs_camera* cam, const fgs_gaussians* g, const fgs_options* opt,
                   void* gauss_ws, size_t gauss_ws_bytes, int32_t* radii, void* stream);

/* Copies [R, NB] to host memory (synchronises the stream). Replaces the
 * reference's curr_offset.cpu() and bucket_sum copies (__init__.py:117,
 * render.cu:424). */
int fgs_read_counts(const void* gauss_ws, int P, int width, int height, const fgs_options* opt,
                    int64_t counts_host[2], void* stream);

/* Replaces sort_gaussian + render_16x16 (sort.cu:28-44, render.cu:393-451):
 * pair emission, sort, tile ranges, bucket tables and compositing. Writes
 * out_color (3 x H x W), T_final (H x W), last_contributor (H x W int32).
 * R, NB: pair and bucket-table capacities (>= fgs_read_counts' values for
 * a complete frame); pair_ws must be fgs_pair_ws_size(..., R, NB, ...). */
int fgs_render(const fgs_camera* cam, const fgs_gaussians* g, const fgs_options* opt,
               void* gauss_ws, size_t gauss_ws_bytes, void* pair_ws, size_t pair_ws_bytes,
               int64_t R, int64_t NB, float* out_color, float* T_final,
               int32_t* last_contributor, void* stream);

/* Replaces faster::render_backward + preprocess_backward (backward.cu:922-977,
 * pybind.cpp:211-339). dL_dimage is 3 x H x W. Outputs are fully written
 * (no pre-zeroing needed): dmean2D P x 4, dconic P x 4, dopacity P,
 * dcolor P x 3, dmean3D P x 3, dcov3D P x 6, ddc P x 3, dsh P x max_coeffs x 3,
 * dscale P x 3, drot P x 4; dconic, dcolor and dcov3D may be NULL (not
 * written; the autograd path of the drop-in API returns none of them unless
 * colours are precomputed). grad_ws is fgs_grad_ws_size(P) bytes. */
typedef struct fgs_grads {
  float *dmean2D, *dconic, *dopacity, *dcolor, *dmean3D, *dcov3D, *ddc, *dsh, *dscale, *drot;
  /* Optional (NULL: left alone): the frame's densification statistics are
     added to these P-sized arrays in the same call, exactly as a following
     fgs_densify_stats(P, radii, dmean2D, ...) would add them (one view: in
     the per-Gaussian adjoint pass itself, which then needs no second read
     of dmean2D and radii). */
  float *xyz_grad_accum, *xyz_grad_accum_abs, *max_radii, *denom;
} fgs_grads;
int fgs_backward(const fgs_camera* cam, const fgs_gaussians* g, const fgs_options* opt,
                 void* gauss_ws, void* pair_ws, int64_t R, int64_t NB, const int32_t* radii,
                 const float* T_final, const int32_t* last_contributor, const float* dL_dimage,
                 void* grad_ws, const fgs_grads* out, void* stream);

/* Replaces faster::add_densification_stats (backward.cu:892-918,
 * pybind.cpp:324-335). Any of the four outputs may be NULL. */
int fgs_densify_stats(int P, const int32_t* radii, const float* dmean2D, float* xyz_grad_accum,
                      float* xyz_grad_accum_abs, float* max_radii, float* denom, void* stream);

/* Replaces _C.ops.render_simp (pybind.cpp:180-208, render_simp.cu:36-311)
 * on a rendered frame's lists (either schedule, one view): per Gaussian the
 * sum of blending weights alpha*T over the pixels (accum_weights, P float),
 * the number of pixels with alpha >= 1/255 (accum_count, P int32) and the
 * number of pixels for which it has the largest weight (accum_max_count, P
 * float). Pixels do not stop on their own; a tile stops after a pair of
 * splats once all its 256 pixels have T < 1e-4. Outputs are overwritten. */
int fgs_render_simp(void* gauss_ws, void* pair_ws, int P, int width, int height, int64_t R,
                    int64_t NB, const fgs_options* opt, float* accum_weights,
                    int32_t* accum_count, float* accum_max_count, void* stream);

/* Replaces _C.ops.render_depth (pybind.cpp:134-172, render_depth.cu:92-380)
 * on a rendered frame's lists (either schedule, one view). Per pixel the
 * camera ray through the pixel centre (inverse of projmatrix, the 16 floats
 * of fgs_camera.projmatrix; campos 3 floats) walks the tile's list front to
 * back until T (1 - alpha) < 1e-4; for the splat of largest weight alpha*T
 * among those whose 3-sigma ellipsoid (means3D, scales, rotations as given,
 * P x 3 / P x 3 / P x 4) has its chord midpoint in front of the camera:
 * out_pts (3 x H x W) the midpoint, out_depth (H x W) its ray parameter
 * over the ray's length, gidx (H x W int32) the Gaussian, discriminants
 * (H x W) the quadratic's discriminant; accum_alpha (H x W) the pixel's
 * final T. Every pixel is written (zeros and T = 1 where nothing was hit). */
int fgs_render_depth(void* gauss_ws, void* pair_ws, int P, int width, int height, int64_t R,
                     int64_t NB, const fgs_options* opt, const float* means3D,
                     const float* scales, const float* rotations, const float* projmatrix,
                     const float* campos, float* out_pts, float* out_depth, float* accum_alpha,
                     int32_t* gidx, float* discriminants, void* stream);

/* Bucket -> (list range, tile) tables of a rendered FAST frame, built on
 * demand into the pair workspace and read through the frame view (replaces
 * the tables of perBucketRange, FGS/csrc/cuda_rasterizer/render.cu:25-49,
 * which the reference's backward reads; here every pass indexes buckets
 * through the per-tile bucket offsets, so no pass of a training step needs
 * them). R and NB as given to fgs_render. */
int fgs_bucket_tables(void* gauss_ws, void* pair_ws, int P, int width, int height, int64_t R,
                      int64_t NB, const fgs_options* opt, void* stream);

/* Views of the workspaces' arrays (host-side pointer arithmetic only). */
int fgs_frame_view_get(void* gauss_ws, void* pair_ws, int P, int width, int height, int64_t R,
                       int64_t NB, const fgs_options* opt, fgs_frame_view* view);

/* Stage-level entry points used by tests and benchmarks. */
/* Stable LSD radix sort of (u64 key, u32 value) pairs on key bits
 * [begin_bit, end_bit); result lands in keys/vals (the *_alt buffers are
 * scratch). Replaces cub::DeviceRadixSort::SortPairs (sort.cu:35-39). */
size_t fgs_sort_ws_size(int64_t n);
int fgs_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                   int64_t n, int begin_bit, int end_bit, void* ws, size_t ws_bytes,
                   void* stream);
/* fgs_sort_pairs without a workspace argument: the same result, with the
 * histograms and scan partials in a static device array of the library
 * (8 MiB, one per device, held until the process ends) instead of a caller
 * buffer. No counterpart in the reference: for callers that cannot allocate
 * (code under graph capture, tight serving loops). n <= 2^24 (else
 * FGS_ERR_CAPACITY). It runs the three-launch passes only (no waits that
 * could time out); nothing in it allocates. Every call on one device shares
 * that device's array: calls that overlap on one device (on different
 * streams or host threads) race. Call it with the stream's device reachable
 * (the array's address is looked up on that device on its first call). */
int fgs_sort_pairs_noalloc(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                           int64_t n, int begin_bit, int end_bit, void* stream);
/* Status of the last fgs_sort_pairs on workspace ws with n pairs
 * (synchronises the stream): FGS_ERR_HIP if one of its bounded look-back
 * waits gave up (the result is then not sorted), else FGS_OK. The frame-level
 * sorts report the same condition through fgs_read_counts, and their images
 * are written NaN. */
int fgs_sort_check(const void* ws, int64_t n, void* stream);
/* Exclusive prefix sum of n uint32 (in place allowed); total written to
 * *total_dev. Replaces cub::DeviceScan (render.cu:420-422). */
size_t fgs_scan_ws_size(int64_t n);
int fgs_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev, void* ws,
                 size_t ws_bytes, void* stream);

/* Self-check of the reverse walk's reciprocal: counts (into *mismatches,
 * device memory, overwritten) the floats x in [2^-7, 1] for which the
 * backward's 1/x differs from the IEEE quotient; 0 on a correct device. */
int fgs_rcp_selftest(unsigned long long* mismatches, void* stream);

/* Measurement probe of the attainable non-packed FP32 FMA rate: blocks x
 * 256 threads each run 8 independent chains of iters fused multiply-adds
 * (16 x iters flops per thread); out receives blocks x 256 floats. Time it
 * with events on the stream. */
int fgs_fma_probe(float* out, int blocks, int iters, void* stream);

/* Cost-model kernels. fgs_mlp_forward: y (n x 2) from x (n x 8), bf16 != 0
 * rounds weights and activations to bfloat16 (fp32 accumulation).
 * fgs_mlp_loss_grad: loss = mean over rows of (y0 - t0)^2 + split_weight *
 * BCE(sigmoid(y1), t1) with targets t (n x 2); writes the full parameter
 * gradient (FGS_MLP_NPARAMS) and the loss (1 float, may be NULL), with
 * fixed-order (deterministic) reductions. fgs_adam_update: one Adam step
 * (torch.optim.Adam rule, bias-corrected, step >= 1) over n parameters. */
int fgs_mlp_forward(const float* params, const float* x, int64_t n, float* y, int bf16,
                    void* stream);
size_t fgs_mlp_grad_ws_size(int64_t n);
int fgs_mlp_loss_grad(const float* params, const float* x, const float* target, int64_t n,
                      float split_weight, float* grad, float* loss, void* ws, size_t ws_bytes,
                      void* stream);
int fgs_adam_update(float* params, const float* grad, float* exp_avg, float* exp_avg_sq,
                    int64_t n, float lr, float beta1, float beta2, float eps, int step,
                    void* stream);

/* Batches of views (FAST schedule): n_views <= 32 cameras (same sh_degree,
 * scale_modifier and mult; any image sizes) render the same Gaussians with
 * one launch per stage -- the multi-view training step. Per-view arrays are
 * concatenated in view order:
 *   radii, dmean2D, dconic        n_views x P rows
 *   out_color and dL_dimage      view v: 3 x H_v x W_v from 3 * pix_off(v)
 *   T_final, last_contributor    view v: H_v x W_v from pix_off(v)
 * where pix_off(v) = sum over u < v of H_u * W_u. Every other gradient
 * (dopacity, dcolor, dmean3D, dcov3D, ddc, dsh, dscale, drot) is the sum
 * over the views, added in view order; fgs_batch_densify_stats adds the
 * views' statistics in view order. Each single-view entry point above is
 * the batch of its one camera. P x n_views must stay below 2^29 (FGS_ERR_ARG
 * otherwise). */
typedef struct fgs_batch {
  int n_views;
  const fgs_camera* cameras; /* host array of n_views cameras */
} fgs_batch;
size_t fgs_batch_gauss_ws_size(int P, const fgs_batch* batch, const fgs_options* opt);
size_t fgs_batch_pair_ws_size(int P, const fgs_batch* batch, int64_t R, int64_t NB,
                              const fgs_options* opt);
size_t fgs_batch_grad_ws_size(int P, int n_views);
int fgs_batch_preprocess(const fgs_batch* batch, const fgs_gaussians* g, const fgs_options* opt,
                         void* gauss_ws, size_t gauss_ws_bytes, int32_t* radii, void* stream);
int fgs_batch_read_counts(const void* gauss_ws, int P, const fgs_batch* batch,
                          const fgs_options* opt, int64_t counts_host[2], void* stream);
int fgs_batch_render(const fgs_batch* batch, const fgs_gaussians* g, const fgs_options* opt,
                     void* gauss_ws, size_t gauss_ws_bytes, void* pair_ws, size_t pair_ws_bytes,
                     int64_t R, int64_t NB, float* out_color, float* T_final,
                     int32_t* last_contributor, void* stream);
int fgs_batch_backward(const fgs_batch* batch, const fgs_gaussians* g, const fgs_options* opt,
                       void* gauss_ws, void* pair_ws, int64_t R, int64_t NB, const int32_t* radii,
                       const float* T_final, const int32_t* last_contributor,
                       const float* dL_dimage, void* grad_ws, const fgs_grads* out, void* stream);
int fgs_batch_densify_stats(int P, int n_views, const int32_t* radii, const float* dmean2D,
                            float* xyz_grad_accum, float* xyz_grad_accum_abs, float* max_radii,
                            float* denom, void* stream);
/* fgs_bucket_tables of a batch of views. */
int fgs_batch_bucket_tables(void* gauss_ws, void* pair_ws, int P, const fgs_batch* batch,
                            int64_t R, int64_t NB, const fgs_options* opt, void* stream);
int fgs_batch_frame_view_get(void* gauss_ws, void* pair_ws, int P, const fgs_batch* batch,
                             int64_t R, int64_t NB, const fgs_options* opt, fgs_frame_view* view);

/* Replaces _C.ops.adamUpdate (pybind.cpp:341-366, adam.cu:9-67) as used by
 * SparseGaussianAdam.step (__init__.py:267-298), for up to 8 parameter
 * tensors in one launch: for every Gaussian g < N with visible[g] != 0 and
 * each of its M parameters, m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
 * p += -lr m / (sqrt(v) + eps); other Gaussians are left untouched. lr is
 * the caller's bias-compensated step size. */
typedef struct fgs_adam_tensor {
  float* param;       /* N x M */
  const float* grad;  /* N x M */
  float* exp_avg;     /* N x M */
  float* exp_avg_sq;  /* N x M */
  int M;
  float lr;
  float eps;
} fgs_adam_tensor;
int fgs_sparse_adam(const fgs_adam_tensor* tensors, int n_tensors, const uint8_t* visible, int N,
                    float beta1, float beta2, void* stream);

/* Fused SSIM of the training loss (fused-ssim: ssim.cu:186-441). Images are
 * B x CH x H x W float32. fgs_ssim_forward writes the SSIM map and, when
 * dm_dmu1 is not NULL (training), the partial derivatives dm/dmu1,
 * dm/dsigma1^2, dm/dsigma12 (all B x CH x H x W). fgs_ssim_backward turns
 * dL/dmap into dL/dimg1. 11-tap Gaussian window (sigma 1.5), zero padding. */
int fgs_ssim_forward(float C1, float C2, const float* img1, const float* img2, int B, int CH,
                     int H, int W, float* ssim_map, float* dm_dmu1, float* dm_dsigma1_sq,
                     float* dm_dsigma12, void* stream);
int fgs_ssim_backward(const float* img1, const float* img2, const float* dL_dmap,
                      const float* dm_dmu1, const float* dm_dsigma1_sq, const float* dm_dsigma12,
                      int B, int CH, int H, int W, float* dL_dimg1, void* stream);

/* Per-kernel timing for measurement runs: while enabled, kernel launches are
 * bracketed by HIP events on their stream; fgs_profile_read synchronises and
 * writes "name total_ms launches\n" lines (aggregated since the last enable)
 * into buf and returns the full length. Not thread-safe; off by default. */
int fgs_profile_enable(int on);
int fgs_profile_read(char* buf, size_t cap);

const char* fgs_last_error(void);
const char* fgs_build_info(void);

#ifdef __cplusplus
}
#endif
#endif /* FASTERGS_H */
