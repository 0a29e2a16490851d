// Workspace carving shared by the C-ABI entry points (host code).
//
// HBM layout (all arrays 256-byte aligned, structure of arrays per stage):
//   gauss workspace (depends on P and the tile grid T):
//     geom   P x 32 B  (x, y, conic a, b | conic c, power, rect min, rect max)
//     splat  P x 48 B  (x, y, ca, cb | cc, log2 op, depth, - | r, g, b, clamp)
//     dkey / pcount / poff   P x 4 B each (+1 for the offsets' total)
//     depth-sort keys/values P x (8+8+4+4) B (fast schedule)
//     tile histogram, ranges, bucket offsets, per-tile max contributor  T x 20 B
//     tile schedule: features T x 32 B, predictions T x 8 B, work items and
//     their timing log 2T x 8 B
//     counters and scan/sort scratch
//   pair workspace (depends on pairs R, buckets NB, pixels HW):
//     keys (2 x R x 8 B), values (2 x R x 4 B), sort histogram scratch,
//     bucket ranges / owner tile (NB x 12 B), per-bucket checkpoints
//     (NB x 256 x (4 + 16) B, fast schedule), pixel colours (HW x 16 B).
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace fgs {

constexpr size_t ALIGN = 256;
// Per-view camera table at the start of the gauss workspace (32 views of at
// most VIEW_REC_BYTES each; the record type lives with the kernels).
constexpr int VIEW_TABLE_VIEWS = 32;
constexpr size_t VIEW_REC_BYTES = 256;
inline size_t align_up(size_t v) { return (v + ALIGN - 1) & ~(ALIGN - 1); }

// Radix sort / scan tiling (must match scan_sort.hip).
constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

inline size_t scan_ws_bytes(int64_t n) {
  const size_t nb = (size_t)((n + SCAN_TILE - 1) / SCAN_TILE);
  return align_up((nb + 1) * sizeof(uint32_t));
}
struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(void* b) : base((char*)b) {}
  template <typename T>
  T* take(size_t count) {
    T* p = base ? (T*)(base + off) : nullptr;
    off += align_up(count * sizeof(T));
    return p;
  }
};

// Onesweep radix sort (inputs up to ONESWEEP_MAX_KEYS): digit histograms of
// up to SORT_MAX_PASSES passes, tile counters + fault flag, and two status
// arrays of 256 words per SORT_TILE keys (alternating between passes).
constexpr int SORT_MAX_PASSES = 8;
// Zero-initialised state of one onesweep sort (digit histograms, tile
// counters, fault flag) that a caller may clear ahead of the sort.
constexpr int SORT_SYNC_WORDS = SORT_MAX_PASSES * 256 + SORT_MAX_PASSES + 1;
constexpr int64_t ONESWEEP_MAX_KEYS = int64_t(1) << 22;
struct SortWs {
  uint32_t* hist;
  uint32_t* counters;
  uint32_t* status[2];
};
template <typename C>
inline SortWs carve_sort(C& c, int64_t nb) {
  SortWs w;
  w.hist = c.template take<uint32_t>(SORT_MAX_PASSES * 256);
  w.counters = c.template take<uint32_t>(SORT_MAX_PASSES + 1);
  w.status[0] = c.template take<uint32_t>((size_t)nb * 256);
  w.status[1] = c.template take<uint32_t>((size_t)nb * 256);
  return w;
}

// Smallest tile (keys per workgroup) any radix pass uses: the workspace
// holds per-tile rows for that many tiles.
constexpr int SORT_MIN_TILE = 2048;
inline size_t sort_ws_bytes(int64_t n) {
  const size_t nb = (size_t)((n + SORT_MIN_TILE - 1) / SORT_MIN_TILE);
  const size_t rows = nb > 0 ? nb : 1;
  // classic passes: digit-major histogram + its scan
  const size_t hist = 256 * rows;
  const size_t classic = align_up(hist * sizeof(uint32_t)) + scan_ws_bytes((int64_t)hist);
  // onesweep passes
  const size_t onesweep = align_up(SORT_MAX_PASSES * 256 * sizeof(uint32_t)) +
                          align_up((SORT_MAX_PASSES + 1) * sizeof(uint32_t)) +
                          2 * align_up(rows * 256 * sizeof(uint32_t));
  return (classic > onesweep ? classic : onesweep) + ALIGN;
}

// Workspace of the three-launch passes alone (SORT_TILE-key tiles):
// digit-major histogram rows and the scan partials over them.
inline size_t sort_classic_ws_bytes(int64_t n) {
  const size_t nb = (size_t)((n + SORT_TILE - 1) / SORT_TILE);
  const size_t hist = 256 * (nb > 0 ? nb : 1);
  return align_up(hist * sizeof(uint32_t)) + scan_ws_bytes((int64_t)hist);
}

inline SortWs carve_sort(void* ws, int64_t nb) {
  Carver c(ws);
  return carve_sort(c, nb);
}

// Large rects are walked in up to LARGE_BANDS row bands of about
// LARGE_BAND_TILES tiles each, one wave per band.
constexpr int LARGE_BANDS = 8;
constexpr int LARGE_BAND_TILES = 512;

// Counters of the persistent compositing grid's work queue: one per XCD
// group of workgroups, each in its own 256-byte line, after the depth sort's
// state (zeroed with it at the start of a frame, left zero by compositing).
constexpr int COMPOSITE_QUEUE_GROUPS = 8;
constexpr int COMPOSITE_QUEUE_STRIDE = 64;
constexpr int COMPOSITE_QUEUE_OFF = (16 + SORT_SYNC_WORDS + 63) / 64 * 64;
// State of the redo list build's sort (prefix lists), zeroed by the prefix
// emission of every render.
constexpr int REDO_SYNC_OFF = COMPOSITE_QUEUE_OFF + COMPOSITE_QUEUE_GROUPS * COMPOSITE_QUEUE_STRIDE;
constexpr int GAUSS_SYNC_WORDS = REDO_SYNC_OFF + SORT_SYNC_WORDS;
// counts words the last-workgroup launches count arrivals in
constexpr int COUNT_ARRIVE_RANGES = 6;
// counts word set when a look-back wait of the render's list sort (tile sort
// or PLAIN pair sort) exceeded its bound
constexpr int COUNT_LIST_FAULT = 5;
// Prefix lists (FAST, one view): [7] pair count of the lists of the nearest
// records (0 when the frame's lists are whole), [8] set by compositing when
// some tile's prefix list did not finish it (the whole lists are then built
// and every tile composited again), [9] the pair count of that second list
// build: counts[0] when [8] is set, else 0.
constexpr int COUNT_PREFIX_N = 7;
constexpr int COUNT_REDO = 8;
constexpr int COUNT_REDO_N = 9;
// The depth sort's fault flag (last word of its state, counted from
// counts), set when a look-back wait exceeds its bound.
constexpr int DSORT_FAULT = 16 + SORT_SYNC_WORDS - 1;

struct GaussWs {
  float* geom;
  float* splat;
  uint32_t* dkey;
  uint32_t* pcount;
  uint32_t* poff;      // pair offsets: by record (PLAIN), by depth rank (FAST)
  uint64_t* dsort_k0;  // fast: depth keys (depth bits << 32 | record), sorted in place:
  uint64_t* dsort_k1;  //   the records by depth rank
  uint32_t* dsort_v0;  // fast: small-rect pair count or PAIRS_LARGE, sorted with the keys,
  uint32_t* dsort_v1;  //   then the resolved counts by depth rank
  uint32_t* lflag;     // fast: row bands of a large rect (walked by whole waves), 0 if small
  uint32_t* large_ids; // fast: row bands of large rects (record * LARGE_BANDS + band), in order
  uint32_t* lband;     // fast: hits of each listed band
  uint32_t* ranges;    // T x 2
  uint32_t* boff;      // T + 1: exclusive prefix of per-tile bucket counts, [T] = NB
  uint32_t* bcount;    // T + 1 scratch (bucket counts, then active bucket counts)
  uint32_t* actoff;    // T + 1: exclusive prefix of per-tile active buckets, [T] = total
  uint32_t* maxc;      // T
  uint32_t* lrank;     // fast: depth rank of each large-rect record
  uint32_t* counts;    // [0] R, [1] NB capacity, [2] NB, [3] large rects, [4] work items,
                       // [5] list-sort fault (COUNT_LIST_FAULT), [6] arrival counter
                       // (COUNT_ARRIVE_RANGES); compositing's work-queue
                       // counters from COMPOSITE_QUEUE_OFF
  uint32_t* dsort_sync;  // fast: depth sort state (SORT_SYNC_WORDS), right after counts
  float* shjac;        // fast, one view: P x 12 SH Jacobians (colour pass -> backward)
  float* feat;         // T x 8 tile features (scheduled compositing)
  float* pred;         // T x 2 cost-model outputs (predicted log2 cost, split logit)
  uint32_t* items;     // 2T compositing work items (tile << 2 | part)
  uint32_t* cycles;    // 2T per-item timing log (100 MHz ticks)
  void* views;         // view table (FAST)
  void* cscratch;      // fast: large-rect compaction (runs beside the depth sort)
  void* scratch;
  size_t scratch_bytes;
  size_t total;
};

// P: records (views x Gaussians), T: tiles of all views; jac: room for the
// SH Jacobians (single views).
inline GaussWs carve_gauss(void* base, int64_t P, int64_t T, bool jac) {
  Carver c(base);
  GaussWs w;
  w.views = c.take<char>(VIEW_TABLE_VIEWS * VIEW_REC_BYTES);
  const size_t p = (size_t)(P > 0 ? P : 1);
  const size_t t = (size_t)(T > 0 ? T : 1);
  w.geom = c.take<float>(8 * p);
  w.splat = c.take<float>(12 * p);
  w.dkey = c.take<uint32_t>(p);
  w.pcount = c.take<uint32_t>(p);
  w.poff = c.take<uint32_t>(p + 1);
  w.dsort_k0 = c.take<uint64_t>(p);
  w.dsort_k1 = c.take<uint64_t>(p);
  w.dsort_v0 = c.take<uint32_t>(p);
  w.dsort_v1 = c.take<uint32_t>(p);
  w.lflag = c.take<uint32_t>(p);
  w.large_ids = c.take<uint32_t>((size_t)LARGE_BANDS * p);
  w.lband = c.take<uint32_t>((size_t)LARGE_BANDS * p);
  w.ranges = c.take<uint32_t>(2 * t);
  w.boff = c.take<uint32_t>(t + 1);
  w.bcount = c.take<uint32_t>(t + 1);
  w.actoff = c.take<uint32_t>(t + 1);
  w.maxc = c.take<uint32_t>(t);
  w.lrank = c.take<uint32_t>(p);
  // counts and the depth sort's state form one region, cleared together at
  // the start of a frame (GAUSS_SYNC_WORDS)
  w.counts = c.take<uint32_t>(GAUSS_SYNC_WORDS);
  w.dsort_sync = w.counts + 16;
  w.shjac = jac ? c.take<float>(12 * p) : nullptr;
  w.feat = c.take<float>(8 * t);
  w.pred = c.take<float>(2 * t);
  w.items = c.take<uint32_t>(2 * t);
  w.cycles = c.take<uint32_t>(2 * t);
  size_t sb = sort_ws_bytes((int64_t)p);
  size_t s2 = scan_ws_bytes((int64_t)p + 1);
  size_t s3 = scan_ws_bytes((int64_t)t + 1);
  if (s2 > sb) sb = s2;
  if (s3 > sb) sb = s3;
  w.scratch_bytes = sb;
  w.scratch = c.take<char>(sb);
  w.cscratch = c.take<char>(scan_ws_bytes((int64_t)p));
  w.total = c.off;
  return w;
}

struct PairWs {
  uint64_t* k0;  // plain: keys (tile << 32 | depth bits)
  uint64_t* k1;
  uint32_t* v0;  // plain: values; fast: the sorted per-tile Gaussian lists
  uint32_t* v1;
  void* t0;      // fast: tile id of each pair (16-bit up to 65536 tiles, else 32-bit)
  void* t1;
  uint32_t* r0;  // fast: record of each pair, emission order
  uint32_t* r1;
  bool tile16;   // fast: tile ids are 16-bit
  void* scratch;
  size_t scratch_bytes;
  uint32_t* bucket_ranges;
  uint32_t* bucket_tile;
  float* acc;    // fast: NB x 256 x (C0, C1, C2, T) after each bucket
  float* pixc;   // HW x (C0, C1, C2, 0)
  size_t total;
};

// T: tiles of all views (the width of the fast schedule's tile ids).
// ckpt: the bucket checkpoints (FAST) and per-pixel colours the backward
// reads; a forward-only frame has neither.
inline PairWs carve_pair(void* base, int64_t R, int64_t NB, int64_t HW, int64_t T, bool fast,
                         bool ckpt = true) {
  Carver c(base);
  PairWs w{};
  const size_t r = (size_t)(R > 0 ? R : 1);
  const size_t nb = (size_t)(NB > 0 ? NB : 1);
  if (fast) {
    w.tile16 = T <= 65536;
    const size_t tb = w.tile16 ? 2 : 4;
    w.t0 = c.take<char>(r * tb);
    w.t1 = c.take<char>(r * tb);
    w.r0 = c.take<uint32_t>(r);
    w.r1 = c.take<uint32_t>(r);
    w.v0 = c.take<uint32_t>(r);
  } else {
    w.k0 = c.take<uint64_t>(r);
    w.k1 = c.take<uint64_t>(r);
    w.v0 = c.take<uint32_t>(r);
    w.v1 = c.take<uint32_t>(r);
  }
  w.scratch_bytes = sort_ws_bytes((int64_t)r);
  w.scratch = c.take<char>(w.scratch_bytes);
  if (fast) {
    w.bucket_ranges = c.take<uint32_t>(2 * nb);
    w.bucket_tile = c.take<uint32_t>(nb);
    w.acc = ckpt ? c.take<float>(nb * 256 * 4) : nullptr;
  } else {
    w.bucket_ranges = nullptr;
    w.bucket_tile = nullptr;
    w.acc = nullptr;
  }
  w.pixc = ckpt ? c.take<float>((size_t)(HW > 0 ? HW : 1) * 4) : nullptr;
  w.total = c.off;
  return w;
}

}  // namespace fgs
