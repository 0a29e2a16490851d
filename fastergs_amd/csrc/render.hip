// Tile ranges, bucket tables and compositing.
//
// identifyTileRanges: FGS/csrc/cuda_rasterizer/render.cu:54-79.
// perTileBucketCount / perBucketRange: render.cu:13-49.
// Compositing semantics: render.cu:99-123 (alpha, blend), :130-275 (bucket
// from T=1), :286-391 (merge with the T < 1e-4 cutoff after a bucket).
//
// PLAIN compositing: one 256-thread workgroup (4 wave64) per 16x16 tile, one
// pixel per thread, one running (T, C) per pixel over the whole list, splats
// staged 256 at a time in LDS, the T < 1e-4 test at every bucket boundary;
// the workgroup leaves as soon as all of its pixels have stopped.
#include <hip/hip_runtime.h>

#include "fgs_device.h"
#include "fgs_kernels.h"
#include "fgs_prof.h"
#include "fgs_layout.h"

namespace fgs {

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f32(float* p, float v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}
hipError_t fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s) {
  if (n > 0) k_fill_u32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, v, n);
  return hipGetLastError();
}
// Attainable non-packed FP32 FMA rate: 8 independent FMA chains per thread,
// `iters` steps each (for measurement runs; out gets one value per thread so
// the chains are not removed).
__global__ __launch_bounds__(256) void k_fma_probe(float* __restrict__ out, int iters, float a,
                                                   float b) {
  float x0 = (float)threadIdx.x, x1 = x0 + 1.f, x2 = x0 + 2.f, x3 = x0 + 3.f;
  float x4 = x0 + 4.f, x5 = x0 + 5.f, x6 = x0 + 6.f, x7 = x0 + 7.f;
  for (int i = 0; i < iters; ++i) {
    x0 = fmaf(x0, a, b); x1 = fmaf(x1, a, b); x2 = fmaf(x2, a, b); x3 = fmaf(x3, a, b);
    x4 = fmaf(x4, a, b); x5 = fmaf(x5, a, b); x6 = fmaf(x6, a, b); x7 = fmaf(x7, a, b);
  }
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = ((x0 + x1) + (x2 + x3)) + ((x4 + x5) + (x6 + x7));
}
hipError_t launch_fma_probe(float* out, int blocks, int iters, hipStream_t s) {
  if (blocks > 0) k_fma_probe<<<blocks, 256, 0, s>>>(out, iters, 0.999f, 0.001f);
  return hipGetLastError();
}

hipError_t fill_f32(float* p, float v, int64_t n, hipStream_t s) {
  if (n > 0) k_fill_f32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, v, n);
  return hipGetLastError();
}

__global__ void k_identify_ranges(int64_t R_cap, const uint32_t* __restrict__ n_dev,
                                  const uint64_t* __restrict__ keys,
                                  uint32_t* __restrict__ ranges) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t R = dev_count(n_dev, R_cap);
  if (idx >= R) return;
  const uint32_t cur = (uint32_t)(keys[idx] >> 32);
  if (idx == 0 || (uint32_t)(keys[idx - 1] >> 32) != cur) ranges[2 * cur] = (uint32_t)idx;
  if (idx == R - 1 || (uint32_t)(keys[idx + 1] >> 32) != cur) ranges[2 * cur + 1] = (uint32_t)(idx + 1);
}

hipError_t launch_identify_ranges(int64_t R, const uint32_t* n_dev, const uint64_t* keys,
                                  uint32_t* ranges, int T, hipStream_t s) {
  hipError_t e = fill_u32(ranges, 0u, 2 * (int64_t)T, s);
  if (e != hipSuccess) return e;
  if (R > 0)
    k_identify_ranges<<<(unsigned)((R + 255) / 256), 256, 0, s>>>(R, n_dev, keys, ranges);
  return hipGetLastError();
}


// boff = exclusive prefix of bucket counts (boff[t] = first bucket of tile
// t, boff[T] = NB). One thread per bucket: its tile is the last t with
// boff[t] <= b (binary search over the T + 1 offsets).
__global__ void k_bucket_tables(int T, int64_t cap, const uint32_t* __restrict__ ranges,
                                const uint32_t* __restrict__ boff, int bucket,
                                uint32_t* __restrict__ bucket_ranges,
                                uint32_t* __restrict__ bucket_tile) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= cap || b >= (int64_t)boff[T]) return;
  int lo = 0, hi = T;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if ((int64_t)boff[mid] <= b) lo = mid; else hi = mid;
  }
  const uint32_t s = ranges[2 * lo] + (uint32_t)(b - boff[lo]) * (uint32_t)bucket;
  const uint32_t e = ranges[2 * lo + 1];
  bucket_ranges[2 * b] = s;
  bucket_ranges[2 * b + 1] = min(s + (uint32_t)bucket, e);
  bucket_tile[b] = (uint32_t)lo;
}
hipError_t launch_bucket_tables(int T, int64_t cap, const uint32_t* ranges, const uint32_t* boff,
                                int bucket, uint32_t* bucket_ranges, uint32_t* bucket_tile,
                                hipStream_t s) {
  if (T > 0 && cap > 0)
    k_bucket_tables<<<(unsigned)((cap + 255) / 256), 256, 0, s>>>(T, cap, ranges, boff, bucket,
                                                                  bucket_ranges, bucket_tile);
  return hipGetLastError();
}

// Ranges from the sorted tile ids (16- or 32-bit):
// ranges[t] = [lower_bound(t), lower_bound(t + 1)).
// Lower bound of v in the sorted a[0, n) by a group of 8 lanes (lane j of
// the group = j, group bit offset g8 in the wave's ballots): each step the 8
// lanes probe the 8 points that cut [lo, hi) into 9 parts and the interval
// shrinks to the part where "a[i] < v" turns false, so 21M ids take 8
// dependent loads instead of the 25 of a binary search. Invariant: a[i] < v
// for i < lo and a[i] >= v for i >= hi.
template <typename TK>
__device__ __forceinline__ uint32_t lower_bound_group8(const TK* a, uint32_t n, uint32_t v,
                                                       int j, int g8) {
  uint32_t lo = 0, hi = n;
  while (__ballot(hi - lo > 8u)) {
    const uint32_t len = hi - lo;
    const bool open = len > 8u;
    const uint32_t q = lo + (uint32_t)(((uint64_t)(j + 1) * len) / 9u);
    const bool pred = open && (uint32_t)a[q] < v;
    const uint32_t m = (uint32_t)__popcll((__ballot(pred) >> g8) & 0xffull);
    if (open) {
      const uint32_t nlo = m > 0 ? lo + (uint32_t)(((uint64_t)m * len) / 9u) + 1u : lo;
      const uint32_t nhi = m < 8 ? lo + (uint32_t)(((uint64_t)(m + 1) * len) / 9u) : hi;
      lo = nlo;
      hi = nhi;
    }
  }
  const uint32_t pos = lo + (uint32_t)j;
  const bool pred = pos < hi && (uint32_t)a[pos] < v;
  return lo + (uint32_t)__popcll((__ballot(pred) >> g8) & 0xffull);
}
// Also writes each tile's bucket count ceil(len / bucket) and clears its
// furthest-contributor slot, which compositing combines with atomicMax when
// a tile is split into row bands; the last workgroup to finish scans the
// bucket counts into boff (T + 1 entries, total at [T]). Workgroup b's 32
// groups of 8 lanes look up the lower bounds of tiles 31 b .. 31 b + 31 and
// it writes the ranges of the first 31 of them.
constexpr int RANGES_PER_BLOCK = 31;
template <typename TK>
__global__ __launch_bounds__(256) void k_tile_ranges(int64_t R_cap,
                                                     const uint32_t* __restrict__ n_dev,
                                                     const TK* __restrict__ tiles,
                                                     uint32_t* __restrict__ ranges, int T,
                                                     int bucket, uint32_t* __restrict__ bcount,
                                                     uint32_t* __restrict__ maxc,
                                                     uint32_t* __restrict__ boff,
                                                     uint32_t* __restrict__ arrive,
                                                     const uint32_t* __restrict__ gate) {
  if (gate && gate[0] == 0u) return;
  __shared__ uint32_t s_lo[32];
  const int grp = threadIdx.x >> 3, j = threadIdx.x & 7, g8 = (threadIdx.x & 56);
  const int t = blockIdx.x * RANGES_PER_BLOCK + grp;
  const uint32_t R = (uint32_t)dev_count(n_dev, R_cap);
  // every lane of the wave takes part in the ballots; groups past T search
  // for T as well (their result is not used)
  const uint32_t lb = lower_bound_group8(tiles, R, (uint32_t)min(t, T), j, g8);
  if (j == 0) s_lo[grp] = lb;
  __syncthreads();
  if (threadIdx.x < RANGES_PER_BLOCK) {
    const int tt = blockIdx.x * RANGES_PER_BLOCK + threadIdx.x;
    if (tt < T) {
      const uint32_t lo = s_lo[threadIdx.x], hi = s_lo[threadIdx.x + 1];
      // empty tiles keep [0, 0) like identifyTileRanges (render.cu:406-411)
      ranges[2 * tt] = lo < hi ? lo : 0u;
      ranges[2 * tt + 1] = lo < hi ? hi : 0u;
      publish_u32(&bcount[tt], (hi - lo + bucket - 1) / bucket);
      maxc[tt] = 0u;
    }
  }
  if (last_workgroup(arrive))
    block_scan_into<256>(T, [&](int i) { return read_published_u32(&bcount[i]); }, boff);
}
hipError_t launch_tile_ranges(int64_t R, const uint32_t* n_dev, const void* tiles, bool t16,
                              uint32_t* ranges, int T, int bucket, uint32_t* bcount,
                              uint32_t* maxc, uint32_t* boff, uint32_t* arrive, hipStream_t s,
                              const uint32_t* gate) {
  if (T <= 0) return hipSuccess;
  const unsigned grid = (unsigned)((T + RANGES_PER_BLOCK - 1) / RANGES_PER_BLOCK);
  const char* name = gate ? "tile_ranges_redo" : "tile_ranges";
  if (t16)
    FGS_TIMED(name, s,
              (k_tile_ranges<uint16_t><<<grid, 256, 0, s>>>(R, n_dev, (const uint16_t*)tiles,
                                                            ranges, T, bucket, bcount, maxc,
                                                            boff, arrive, gate)));
  else
    FGS_TIMED(name, s,
              (k_tile_ranges<uint32_t><<<grid, 256, 0, s>>>(R, n_dev, (const uint32_t*)tiles,
                                                            ranges, T, bucket, bcount, maxc,
                                                            boff, arrive, gate)));
  return hipGetLastError();
}

// Buckets of each tile that some pixel still needs in the backward:
// ceil(max_contrib / B) (backward.cu:721-722 skips the others).
__global__ void k_active_counts(int T, const uint32_t* __restrict__ maxc, int bucket,
                                uint32_t* __restrict__ out) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < T) out[t] = (maxc[t] + bucket - 1) / bucket;
}
hipError_t launch_active_counts(int T, const uint32_t* maxc, int bucket, uint32_t* out,
                                hipStream_t s) {
  if (T > 0) k_active_counts<<<(T + 255) / 256, 256, 0, s>>>(T, maxc, bucket, out);
  return hipGetLastError();
}


// True when a sort of this frame gave up a look-back wait (the lists are
// then not valid; compositing writes NaN so that the frame cannot pass as an
// image, and fgs_read_counts reports the error).
__device__ __forceinline__ bool lists_faulted(const uint32_t* counts) {
  return counts && (counts[COUNT_LIST_FAULT] | counts[DSORT_FAULT]) != 0u;
}

// ------------------------------------------------------ PLAIN compositing
__global__ __launch_bounds__(256) void k_composite_plain(CompArgs a) {
  __shared__ float4 sA[256];
  __shared__ float4 sB[256];
  __shared__ float sC[256];
  __shared__ int wmax[4];
  const int tile = blockIdx.x;
  const int tx = tile % a.grid_x, ty = tile / a.grid_x;
  const int t = threadIdx.x;
  const int tw = a.tile_w, twl = tile_w_log2(tw);
  const int px = tx * tw + (t & (tw - 1)), py = ty * tile_h_of(tw) + (t >> twl);
  const bool inside = px < a.W && py < a.H;
  const uint32_t start = a.ranges[2 * tile], end = a.ranges[2 * tile + 1];
  const float fx = (float)px, fy = (float)py;
  float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
  int cnt = 0;
  bool done = !inside;
  const int B = a.bucket;
  for (uint32_t base = start; base < end; base += 256) {
    if (__syncthreads_count(!done) == 0) break;
    const uint32_t k = base + t;
    if (k < end) {
      const float4* s = (const float4*)(a.splat + 12 * (size_t)a.vals[k]);
      const float4 s1 = s[1], s2 = s[2];
      sA[t] = s[0];
      sB[t] = make_float4(s1.x, s1.y, s2.x, s2.y);  // cc, log2 op, r, g
      sC[t] = s2.z;
    }
    __syncthreads();
    const int n = (int)min(256u, end - base);
    if (!done) {
      for (int j = 0; j < n; j += B) {
        const int e = min(j + B, n);
        for (int q = j; q < e; ++q) {
          const float4 A = sA[q], Bv = sB[q];
          const float al = splat_alpha(A.x, A.y, A.z, A.w, Bv.x, Bv.y, fx, fy);
          blend(al, Bv.z, Bv.w, sC[q], T, C0, C1, C2);
        }
        cnt += e - j;
        if (T < T_CUT_F) {
          done = true;
          break;
        }
      }
    }
  }
  if (lists_faulted(a.counts)) T = C0 = __builtin_nanf("");
  if (inside) {
    const size_t HW = (size_t)a.W * a.H;
    const size_t g = (size_t)py * a.W + px;
    a.out[g] = fma_(T, a.bg[0], C0);
    a.out[HW + g] = fma_(T, a.bg[1], C1);
    a.out[2 * HW + g] = fma_(T, a.bg[2], C2);
    a.tfinal[g] = T;
    a.contrib[g] = cnt;
    if (a.pixc) ((float4*)a.pixc)[g] = make_float4(C0, C1, C2, 0.0f);
  }
  const int m = wave_max_i(cnt);
  if (lane_id() == 0) wmax[t / WAVE] = m;
  __syncthreads();
  if (t == 0) a.maxc[tile] = (uint32_t)max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
}

hipError_t launch_composite_plain(const CompArgs& a, hipStream_t s) {
  if (a.T > 0) FGS_TIMED("composite_plain", s, (k_composite_plain<<<a.T, 256, 0, s>>>(a)));
  return hipGetLastError();
}

// ------------------------------------------------------- FAST compositing
// One 256-thread workgroup (4 wave64) per work item: a whole 16x16 tile, or
// one of its two 16x8 row bands when the tile schedule splits it (pixels
// are independent, so a band does exactly the per-pixel work of the whole
// tile for its rows, with the upper 128 threads staging only). One pixel
// per thread; splats staged 256 at a time in LDS. Every bucket is blended
// from T = 1 (render.cu:130-275) and merged into the pixel's running (T, C)
// in bucket order (render.cu:329-362); the checkpoint after each bucket is
// stored for the backward (unless the frame is forward-only), and a pixel
// stops after the first bucket that leaves T < 1e-4. A wave whose 64 pixels
// have all stopped skips the rest of the list, and the workgroup leaves once
// all its pixels have: buckets past the cut-off (which the reference renders
// and then discards) are not evaluated. The outputs do not depend on the
// item order or the splitting.
// Work item of grid index i: false when there is nothing to do (past the
// list, or the unused second slot of a tile that is not split).
__device__ __forceinline__ bool composite_item_of(const CompArgs& a, uint32_t i, uint32_t& item,
                                                  int& log_idx) {
  if (a.items) {
    if (i >= a.n_items[0]) return false;
    item = a.items[i];
    log_idx = (int)i;
    return (item >> 2) < (uint32_t)a.T && (item & 3u) != 3u;  // else not an item of this frame
  }
  if (a.slots) {
    // two grid slots per tile: the upper and lower band of a split tile, or
    // the whole tile in slot 0. Workgroup b runs on XCD b % 8, so the slots
    // are laid out [all first slots | all second slots]: interleaving them
    // would leave the mostly idle second slots to half of the XCDs.
    const int slot = (int)i >= a.T ? 1 : 0;
    const int tl = (int)i - slot * a.T;
    const uint32_t len = a.ranges[2 * tl + 1] - a.ranges[2 * tl];
    const bool split = a.split_pred ? (a.split_pred[2 * tl + 1] > 0.0f && len > (uint32_t)a.bucket)
                                    : (a.split_len > 0 && len >= (uint32_t)a.split_len);
    log_idx = slot * a.T + tl;
    if (!split && slot) {
      if (a.cycles && threadIdx.x == 0) a.cycles[log_idx] = 0u;
      return false;
    }
    item = ((uint32_t)tl << 2) | (split ? 1u + (uint32_t)slot : 0u);
    return true;
  }
  item = i << 2;
  log_idx = (int)i;
  return true;
}

struct CompShared {
  float4 sA[256];
  float4 sB[256];
  float sC[256];
  int wmax[4];
};

// One work item (tile << 2 | part) by the whole workgroup.
__device__ __forceinline__ void composite_item(const CompArgs& a, CompShared& sh, uint32_t item,
                                               int log_idx, uint64_t t_start) {
  const int tile = (int)(item >> 2), part = (int)(item & 3u);
  // the tile's view: its size, tile grid and shape, background and image
  // offsets
  const ViewRec& vr = a.views[a.V > 1 ? view_of_tile(a.views, a.V, (uint32_t)tile) : 0];
  const int W = vr.W, H = vr.H, gx = vr.grid_x, tw = vr.tile_w, th = vr.tile_h;
  const int twl = tile_w_log2(tw);
  const bool band = part != 0;
  const int row0 = part == 2 ? th / 2 : 0;
  const int nrows = band ? th / 2 : th;
  const int lt = tile - (int)vr.tile_off;
  const int tx = lt % gx, ty = lt / gx;
  const int t = threadIdx.x;
  const int w = t / WAVE, lane = t & (WAVE - 1);
  // this thread's pixel (a band's upper 128 threads have none)
  const int mrow = row0 + (t >> twl);
  const int lp = mrow * tw + (t & (tw - 1));
  const int px = tx * tw + (t & (tw - 1)), py = ty * th + mrow;
  const bool mine = (t >> twl) < nrows;
  const bool inside = mine && px < W && py < H;
  const float fx = (float)px, fy = (float)py;
  const uint32_t start = a.ranges[2 * tile], end = a.ranges[2 * tile + 1];
  const int B = a.bucket;
  const uint32_t b0 = a.boff[tile];
  float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
  int cnt = 0;
  bool done = !inside;
  for (uint32_t base = start; base < end; base += 256) {
    if (__syncthreads_count(!done) == 0) break;
    const uint32_t k = base + (uint32_t)t;
    if (k < end) {
      const float4* r = (const float4*)(a.splat + 12 * (size_t)a.vals[k]);
      const float4 r1 = r[1], r2 = r[2];
      sh.sA[t] = r[0];
      sh.sB[t] = make_float4(r1.x, r1.y, r2.x, r2.y);  // cc, log2 op, r, g
      sh.sC[t] = r2.z;
    }
    __syncthreads();
    if (!done) {
      const int n = (int)min(256u, end - base);
      // bucket index of the batch's first bucket (256 is a multiple of B)
      const uint32_t bq0 = (base - start) / (uint32_t)B;
      for (int j = 0, q = 0; j < n; j += B, ++q) {
        const int e = min(j + B, n);
        float Tb = 1.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
        for (int i = j; i < e; ++i) {
          const float4 A = sh.sA[i], Bv = sh.sB[i];
          blend(splat_alpha(A.x, A.y, A.z, A.w, Bv.x, Bv.y, fx, fy), Bv.z, Bv.w, sh.sC[i], Tb, c0,
                c1, c2);
        }
        merge(Tb, c0, c1, c2, T, C0, C1, C2);
        if (a.acc)
          ((float4*)a.acc)[(size_t)(b0 + bq0 + (uint32_t)q) * TILE_PIX + lp] =
              make_float4(C0, C1, C2, T);
        cnt += e - j;
        if (T < T_CUT_F) {
          done = true;
          break;
        }
      }
    }
  }
  if (a.redo) {
    // Prefix lists: the list holds the front of the tile's whole list. A
    // pixel that stopped at the end of a whole bucket has its final values;
    // one still open at the end of the list, or stopped after a bucket the
    // list cut short, needs the rest of the whole list.
    const bool open = inside && (!done || cnt % B != 0);
    if (__syncthreads_or(open) && t == 0) {
      a.redo[0] = 1u;
      a.redo[1] = a.counts[0];
      if (a.status) *a.status = 1u;
    }
  }
  if (lists_faulted(a.counts)) T = C0 = __builtin_nanf("");
  if (inside) {
    const size_t HW = (size_t)W * H;
    const size_t g = (size_t)vr.pix_off + (size_t)py * W + px;
    float* out = a.out + 3 * (size_t)vr.pix_off + (size_t)py * W + px;
    out[0] = fma_(T, vr.bg[0], C0);
    out[HW] = fma_(T, vr.bg[1], C1);
    out[2 * HW] = fma_(T, vr.bg[2], C2);
    a.tfinal[g] = T;
    a.contrib[g] = cnt;
    if (a.pixc) ((float4*)a.pixc)[g] = make_float4(C0, C1, C2, 0.0f);
  }
  const int m = wave_max_i(mine ? cnt : 0);
  if (lane == 0) sh.wmax[w] = m;
  __syncthreads();
  if (t == 0) {
    const uint32_t mx = (uint32_t)max(max(sh.wmax[0], sh.wmax[1]), max(sh.wmax[2], sh.wmax[3]));
    // the two bands of a split tile combine into the entry k_tile_ranges cleared
    if (band)
      atomicMax(&a.maxc[tile], mx);
    else
      a.maxc[tile] = mx;
    if (a.cycles) a.cycles[log_idx] = (uint32_t)(wall_clock64() - t_start);
  }
}

// QUEUE = false: workgroup b does grid index b. QUEUE = true: a persistent
// grid (as many workgroups as are resident at once) works through the
// frame's grid indices from a work queue. Grid index i belongs to group
// i % 8, served by the workgroups b with b % 8 equal to it (those the
// dispatcher places on one XCD) through that group's own counter in a
// separate 256-byte line: a single counter takes every take of the chip in
// turn and its atomics queue up (measured: 3x slower compositing). A
// workgroup does its own grid index first, then takes further ones from its
// group's counter, each take issued before the current item so that its
// latency hides behind the item. Every workgroup makes exactly one failing
// take (past its group's items), so the workgroup that receives the last
// value handed out resets the counter to zero for the next launch.
template <bool QUEUE>
__global__ __launch_bounds__(256) void k_composite_fast(CompArgs a) {
  __shared__ CompShared sh;
  if (a.gate && a.gate[0] == 0u) return;  // whole grid: the prefix lists sufficed
  if (a.zero) {
    const int64_t n4 = a.zero_n / 4, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n4; j += stride)
      ((float4*)a.zero)[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  if (!QUEUE) {
    const uint64_t t_start = a.cycles ? wall_clock64() : 0;
    uint32_t item;
    int log_idx;
    if (composite_item_of(a, blockIdx.x, item, log_idx)) composite_item(a, sh, item, log_idx, t_start);
    return;
  }
  __shared__ uint32_t s_next;
  const uint32_t n = a.items ? min(a.n_items[0], 2u * (uint32_t)a.T)
                             : (a.slots ? 2u * (uint32_t)a.T : (uint32_t)a.T);
  const uint32_t c = blockIdx.x & 7u, G = gridDim.x;
  const uint32_t Gc = (G - c + 7u) / 8u;             // workgroups of group c (>= 1)
  const uint32_t nc = n > c ? (n - c + 7u) / 8u : 0u;  // grid indices of group c
  uint32_t* ctr = a.queue + COMPOSITE_QUEUE_STRIDE * c;
  uint32_t k = blockIdx.x >> 3;  // k-th index of the group: grid index 8 k + c
  bool have = k < nc;
  for (;;) {
    uint32_t take = 0u;
    if (threadIdx.x == 0) take = atomicAdd(ctr, 1u);
    if (have) {
      const uint64_t t_start = a.cycles ? wall_clock64() : 0;
      uint32_t item;
      int log_idx;
      if (composite_item_of(a, 8u * k + c, item, log_idx))
        composite_item(a, sh, item, log_idx, t_start);
    }
    if (threadIdx.x == 0) s_next = take;
    __syncthreads();
    const uint32_t j = s_next;
    __syncthreads();  // read by every thread before thread 0 stores the next one
    k = Gc + j;
    if (k >= nc) {  // this workgroup's one failing take
      if (threadIdx.x == 0 && j == (nc > Gc ? nc - Gc : 0u) + Gc - 1u) atomicExch(ctr, 0u);
      return;  // whole workgroup
    }
    have = true;
  }
}

// Resident 256-thread workgroups of the queue form on this device.
static int composite_queue_grid() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  if (!cached[dev]) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_composite_fast<true>, 256, 0) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    cached[dev] = per_cu * cus;
  }
  return cached[dev];
}

hipError_t launch_composite_fast(const CompArgs& a, hipStream_t s) {
  if (a.T <= 0) return hipSuccess;
  if (a.bucket > 256 || 256 % a.bucket != 0) return hipErrorInvalidValue;
  const unsigned grid = (a.items || a.slots) ? 2u * (unsigned)a.T : (unsigned)a.T;
  if (a.queue) {
    const int resident = composite_queue_grid();
    if (resident <= 0) return hipErrorInvalidValue;
    const unsigned g = grid < (unsigned)resident ? grid : (unsigned)resident;
    FGS_TIMED(a.gate ? "composite_fast_redo" : "composite_fast", s,
              (k_composite_fast<true><<<g, 256, 0, s>>>(a)));
  } else {
    FGS_TIMED(a.gate ? "composite_fast_redo" : "composite_fast", s,
              (k_composite_fast<false><<<grid, 256, 0, s>>>(a)));
  }
  return hipGetLastError();
}

// ----------------------------------------------------------- render_simp
// Importance accumulation over a rendered frame's lists (render_simp.cu:36-301):
// per tile, pixels blend front to back without stopping on their own; after
// each pair of splats (a leading single splat when the list length is odd)
// the tile stops once all 256 of its pixels -- those outside the image keep
// T = 1 -- have T < 1e-4. Per splat the weights alpha*T and the number of
// pixels with alpha >= 1/255 are summed; per pixel the splat of largest
// weight (the first of equals) gets +1 at the end.
//
// A tile with pixels outside the image therefore never stops and walks its
// whole list, serially per pixel: such tiles (the last tile row and column
// when the image is not a multiple of the tile) are cut into pieces of 32
// pixels that run first in the grid, each with its alphas computed by three
// waves and the T chains run by a fourth (simp_piece); the full tiles follow,
// one workgroup each, walking in chunks until the stop (simp_tile).
constexpr int SIMP_CHUNK = 32;
constexpr int SIMP_PIECE = 32;

// Work items of one render_simp launch. Tiles with pixels outside the image
// never stop (those pixels keep T = 1), so their whole lists are walked:
// their pixels inside the image are cut into pieces of SIMP_PIECE
// (row-major over the tile's inside rectangle), handled first in the grid
// by simp_piece. The tiles of the full-tile region follow, one per
// workgroup.
struct SimpGrid {
  int W, H, gx, gy;
  int tw, th;   // tile shape
  int hr, wr;   // inside rows of the last tile row / columns of the last column (0: full)
  int gxf;      // full-width tiles of the last row
  int nbf, nc, nr;  // pieces per last-row tile, of the corner tile, per last-column tile
  int n_pieces;
  int gxi, gyi;     // the full-tile region
};
inline SimpGrid simp_grid(int W, int H, int tw) {
  SimpGrid g;
  g.W = W;
  g.H = H;
  g.tw = tw;
  g.th = tile_h_of(tw);
  g.gx = grid_dim(W, g.tw);
  g.gy = grid_dim(H, g.th);
  g.hr = H % g.th;
  g.wr = W % g.tw;
  g.gxf = g.wr ? g.gx - 1 : g.gx;
  g.nbf = g.hr ? (g.tw * g.hr + SIMP_PIECE - 1) / SIMP_PIECE : 0;
  g.nc = (g.hr && g.wr) ? (g.wr * g.hr + SIMP_PIECE - 1) / SIMP_PIECE : 0;
  g.nr = g.wr ? (g.wr * g.th + SIMP_PIECE - 1) / SIMP_PIECE : 0;
  g.gxi = g.gxf;
  g.gyi = g.hr ? g.gy - 1 : g.gy;
  g.n_pieces = g.gxf * g.nbf + g.nc + g.gyi * g.nr;
  return g;
}

// First position at which the tile checks its stop at or after position
// pos: pair ends are the odd positions of an even-length list and the even
// positions >= 2 of an odd-length one (whose position 0 is the single).
__device__ __forceinline__ uint32_t simp_pair_end(uint32_t pos, uint32_t len) {
  return (len & 1u) ? max(2u, pos + (pos & 1u)) : (pos | 1u);
}

// Blend of one staged chunk for the calling lane's pixel (inside the image):
// the weights into w[k], T and the running maximum updated (kmax: chunk
// index of a new maximum, else -1), and the wave's count of pixels hit by
// splat k into lane 2k of cnt. Entries k >= n of a partial chunk (FULL
// false) hold stale values: they are read anyway, so the body has no branch,
// and masked out of the hits. Splats are read from LDS four ahead.
template <bool FULL>
__device__ __forceinline__ void simp_chunk(const float4* __restrict__ sA,
                                           const float2* __restrict__ sB, uint32_t n, float fx,
                                           float fy, float& T, float& wmax, int& kmax,
                                           float (&w)[SIMP_CHUNK], float& cnt) {
  constexpr int D = 4;
  float4 A[D];
  float2 Bv[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    A[d] = sA[d];
    Bv[d] = sB[d];
  }
#pragma unroll
  for (int k = 0; k < SIMP_CHUNK; ++k) {
    const float4 Ak = A[k % D];
    const float2 Bk = Bv[k % D];
    if (k + D < SIMP_CHUNK) {
      A[k % D] = sA[k + D];
      Bv[k % D] = sB[k + D];
    }
    const float ar = alpha_capped(splat_power(Ak.x, Ak.y, Ak.z, Ak.w, Bk.x, Bk.y, fx, fy));
    const bool above = !(ar < 1.0f / 255.0f);
    const bool valid = FULL || (uint32_t)k < n;
    const uint64_t hitm = valid ? __builtin_amdgcn_ballot_w64(above) : 0ull;
    const bool hit = above && valid;
    const float a = hit ? ar : 0.0f;
    w[k] = a * T;
    T = T - w[k];
    if (wmax < w[k]) {
      wmax = w[k];
      kmax = k;
    }
    const int hits = (int)__builtin_popcountll(hitm);
    asm("v_writelane_b32 %0, %1, %2" : "+v"(cnt) : "s"(hits), "i"(2 * k));
  }
}

// One atomic per distinct maximum in the wave (neighbouring pixels mostly
// share it; the counts are integers, so the float sums stay exact).
__device__ __forceinline__ void simp_max_flush(bool has, uint32_t id, float* maxcount) {
  const int lane = lane_id();
  uint64_t pending = __builtin_amdgcn_ballot_w64(has);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)id, leader);
    const uint64_t same = __builtin_amdgcn_ballot_w64(has && id == k) & pending;
    if (lane == leader) atomic_add_f(&maxcount[k], (float)__popcll(same));
    pending &= ~same;
  }
}

// A tile of the full-tile region: one pixel per thread, the list walked in
// chunks of SIMP_CHUNK splats with one barrier per chunk. The chunk's splats
// are staged in LDS (double buffered; the next chunk's loads are in flight
// while the current one is blended), each lane keeps its pixel's weights of
// the chunk in registers, and one transposing wave sum per chunk gives the
// per-splat totals; the hit counts come from ballots. T only decreases, so a
// wave whose 64 pixels are all below 1e-4 at a chunk end finds the first
// pair end at which that held by replaying the chunk's T chain (exactly,
// from the stored weights) in the lanes that crossed in it; the tile's stop
// is the largest of the four waves' stops, and only splats up to it are
// added to memory. Past the stop every weight is < 1e-4 while each pixel's
// largest weight is >= 1/255 (its first contributor meets T = 1), so the
// per-pixel maxima need no such cut.
__device__ __forceinline__ void simp_tile(int tile, const SimpGrid& g,
                                          const uint32_t* __restrict__ ranges,
                                          const uint32_t* __restrict__ vals,
                                          const float* __restrict__ splat, float* __restrict__ wsum,
                                          int32_t* __restrict__ wcount,
                                          float* __restrict__ maxcount) {
  constexpr int CH = SIMP_CHUNK;
  constexpr uint32_t NONE = 0xffffffffu;
  __shared__ float4 sA[2][CH];        // gx, gy, ca, cb
  __shared__ float2 sB[2][CH];        // cc, log2 opacity
  __shared__ float sW[2][4][CH];      // per wave and splat: weight sum
  __shared__ uint32_t sC[2][4][CH];   // per wave and splat: pixels hit
  __shared__ uint32_t sStop[2][4];    // per wave: its stop position, or NONE
  const int tx = tile % g.gx, ty = tile / g.gx;
  const int t = threadIdx.x;
  const int wave = t >> 6, lane = t & (WAVE - 1);
  const float fx = (float)(tx * g.tw + (t & (g.tw - 1))),
              fy = (float)(ty * g.th + (t >> tile_w_log2(g.tw)));
  const uint32_t start = ranges[2 * tile], len = ranges[2 * tile + 1] - start;
  // staging: thread t < CH handles position p0 + t of each chunk; ids are
  // loaded two chunks ahead and records one chunk ahead, so no load waits
  // on another within a chunk
  const bool stager = t < CH;
  auto load_id = [&](uint32_t p0) -> uint32_t {
    return stager && p0 + (uint32_t)t < len ? vals[start + p0 + t] : 0u;
  };
  float4 nA = make_float4(0.f, 0.f, 0.f, 0.f);
  float2 nB = make_float2(0.f, 0.f);
  auto load_rec = [&](uint32_t id, uint32_t p0) {
    if (stager && p0 + (uint32_t)t < len) {
      const float4* r = (const float4*)(splat + 12 * (size_t)id);
      const float4 r1 = r[1];
      nA = r[0];
      nB = make_float2(r1.x, r1.y);
    }
  };
  uint32_t id0 = load_id(0), id1 = load_id(CH);  // ids of chunks c and c + 1
  load_rec(id0, 0);
  if (stager) {
    sA[0][t] = nA;
    sB[0][t] = nB;
  }
  __syncthreads();
  float T = 1.0f, wmax = 0.0f;
  uint32_t imax = 0;       // list position of the largest weight
  uint32_t my_stop = NONE;  // this wave's stop position, once known
  uint32_t c = 0;
  for (uint32_t p0 = 0; p0 < len; p0 += CH, ++c) {
    const int b = (int)(c & 1u);
    const uint32_t n = min((uint32_t)CH, len - p0);
    load_rec(id1, p0 + CH);                  // records of chunk c + 1
    const uint32_t id2 = load_id(p0 + 2 * CH);  // ids of chunk c + 2
    float w[CH];
    float cnt = 0.0f;  // lane 2k: pixels of this wave that splat k hits
    const float T0 = T;
    int kmax = -1;
    if (n == (uint32_t)CH)
      simp_chunk<true>(sA[b], sB[b], n, fx, fy, T, wmax, kmax, w, cnt);
    else
      simp_chunk<false>(sA[b], sB[b], n, fx, fy, T, wmax, kmax, w, cnt);
    if (kmax >= 0) imax = p0 + (uint32_t)kmax;
    if (my_stop == NONE && __builtin_amdgcn_ballot_w64(T < T_CUT_F) == ~0ull) {
      // every pixel is below the cut now; the last one to get there crossed
      // in this chunk: replay its T chain to the crossing position
      int k0 = -1;
      if (T0 >= T_CUT_F) {
        float Tr = T0;
#pragma unroll
        for (int k = 0; k < CH; ++k) {
          Tr = Tr - w[k];
          if (k0 < 0 && Tr < T_CUT_F) k0 = k;
        }
      }
      my_stop = simp_pair_end(p0 + (uint32_t)max(wave_max_i(k0), 0), len);
    }
    const float ws = wave_transpose_sum32(w);
    if ((lane & 1) == 0) {
      sW[b][wave][lane >> 1] = ws;
      sC[b][wave][lane >> 1] = (uint32_t)__builtin_bit_cast(int, cnt);
    }
    if (lane == 0) sStop[b][wave] = my_stop;
    if (stager && p0 + CH + (uint32_t)t < len) {
      sA[b ^ 1][t] = nA;
      sB[b ^ 1][t] = nB;
    }
    __syncthreads();
    const uint32_t stop = max(max(sStop[b][0], sStop[b][1]), max(sStop[b][2], sStop[b][3]));
    if ((uint32_t)t < n && p0 + (uint32_t)t <= stop) {
      const uint32_t hits = (sC[b][0][t] + sC[b][1][t]) + (sC[b][2][t] + sC[b][3][t]);
      if (hits != 0u) {
        atomic_add_f(&wsum[id0], (sW[b][0][t] + sW[b][1][t]) + (sW[b][2][t] + sW[b][3][t]));
        atomicAdd(&wcount[id0], (int32_t)hits);
      }
    }
    if (stop < p0 + n) break;
    id0 = id1;
    id1 = id2;
  }
  const bool has = wmax > 0.0f;
  simp_max_flush(has, has ? vals[start + imax] : 0u, maxcount);
}

// A piece of a tile that never stops: up to SIMP_PIECE of its pixels inside
// the image, the whole list walked. Wave 0 runs the pixels' T chains (lanes
// below SIMP_PIECE) over alphas that waves 1-3 computed one chunk ahead into
// LDS; a producing pass evaluates two splats at once (lanes 0-31 the first,
// 32-63 the second, one pixel per lane), pass g of a chunk (splats 2g, 2g +
// 1) by wave 1 + g % 3, which also counts the hits. Wave 1 stages the splats
// two chunks ahead (a ring of three), its loads one chunk earlier. Wave 0
// sums the weights of a chunk with one transposing wave sum and adds them
// to memory with the staged counts. A lone wave issues a dependent VALU
// instruction only every few cycles, so the piece is sized for the chain
// wave to be the slowest of the four.
template <bool FULL>
__device__ __forceinline__ void simp_piece_chain(const float (*__restrict__ al)[SIMP_PIECE],
                                                 uint32_t n, int p, bool inside, float& T,
                                                 float& wmax, int& kmax,
                                                 float (&w)[SIMP_CHUNK]) {
#pragma unroll
  for (int k = 0; k < SIMP_CHUNK; ++k) {
    const float av = al[k][p];  // stale for k >= n: masked
    const float a = inside && (FULL || (uint32_t)k < n) ? av : 0.0f;
    w[k] = a * T;
    T = T - w[k];
    if (wmax < w[k]) {
      wmax = w[k];
      kmax = k;
    }
  }
}

__device__ __forceinline__ void simp_piece(int piece, const SimpGrid& g,
                                           const uint32_t* __restrict__ ranges,
                                           const uint32_t* __restrict__ vals,
                                           const float* __restrict__ splat,
                                           float* __restrict__ wsum, int32_t* __restrict__ wcount,
                                           float* __restrict__ maxcount) {
  constexpr int CH = SIMP_CHUNK, PX = SIMP_PIECE;
  constexpr int PASSES = CH / 2;               // two splats per producing pass
  constexpr int PER = (PASSES + 2) / 3;        // passes of wave 1 (the most)
  // alpha (0 if not hit) per chunk splat and piece pixel; rows CH.. take the
  // unused last passes of waves 2 and 3, so the stores need no branch
  __shared__ float sAl[2][CH + 4][PX];
  __shared__ uint32_t sCnt[2][CH];    // pixels of the piece hit per chunk splat
  __shared__ float4 qA[3][CH];        // staged splats: gx, gy, ca, cb
  __shared__ float2 qB[3][CH];        // cc, log2 opacity
  __shared__ uint32_t qId[3][CH];
  // the piece: tile, index within it, the tile's inside rectangle
  int tile, q, iw, ih;
  {
    int b = piece;
    if (b < g.gxf * g.nbf) {
      tile = (g.gy - 1) * g.gx + b / g.nbf;
      q = b % g.nbf;
      iw = g.tw;
      ih = g.hr;
    } else if ((b -= g.gxf * g.nbf) < g.nc) {
      tile = (g.gy - 1) * g.gx + g.gx - 1;
      q = b;
      iw = g.wr;
      ih = g.hr;
    } else {
      b -= g.nc;
      tile = (b / g.nr) * g.gx + g.gx - 1;
      q = b % g.nr;
      iw = g.wr;
      ih = g.th;
    }
  }
  const int tx = tile % g.gx, ty = tile / g.gx;
  const int t = threadIdx.x;
  const int wave = t >> 6, lane = t & (WAVE - 1);
  // this lane's pixel: p of the piece (both halves of a producing wave map
  // lanes l and l + 32 to the same pixel; in the chain wave only lanes < 32)
  const int p = lane & (PX - 1);
  const int i = PX * q + p;
  const bool in_piece = i < iw * ih;
  const bool inside = in_piece && (wave != 0 || lane < PX);
  const float fx = (float)(tx * g.tw + i % iw), fy = (float)(ty * g.th + i / iw);
  const uint32_t start = ranges[2 * tile], len = ranges[2 * tile + 1] - start;
  const uint32_t nch = (len + CH - 1) / CH;
  // stager: wave 1, lanes < CH, position c CH + lane of chunk c. Step c
  // stages chunk c + 2 from records loaded in step c - 1, loads the records
  // of chunk c + 3 from ids loaded in step c - 1, and the ids of chunk c + 4.
  const bool stager = wave == 1 && lane < CH;
  auto in_list = [&](uint32_t c) { return stager && c < nch && c * CH + (uint32_t)lane < len; };
  auto load_id = [&](uint32_t c) -> uint32_t {
    return in_list(c) ? vals[start + c * CH + (uint32_t)lane] : 0u;
  };
  struct Rec {
    float4 A;
    float2 B;
    uint32_t id;
  };
  auto load_rec = [&](uint32_t c, uint32_t id) -> Rec {
    Rec r{make_float4(0.f, 0.f, 0.f, 0.f), make_float2(0.f, 0.f), id};
    if (in_list(c)) {
      const float4* pr = (const float4*)(splat + 12 * (size_t)id);
      const float4 p1 = pr[1];
      r.A = pr[0];
      r.B = make_float2(p1.x, p1.y);
    }
    return r;
  };
  auto stage = [&](uint32_t c, const Rec& r) {
    if (in_list(c)) {
      qA[c % 3][lane] = r.A;
      qB[c % 3][lane] = r.B;
      qId[c % 3][lane] = r.id;
    }
  };
  // alphas of chunk c; entries k >= n are computed from stale splats and
  // never read. Hit counts go to lane 2m + half of `hc` for the wave's m-th
  // pass and are stored once per chunk.
  const int half = lane >> 5;
  const uint64_t inside_mask = __builtin_amdgcn_ballot_w64(inside);
  auto produce = [&](uint32_t c) {
    if (c >= nch) return;
    const uint32_t n = min((uint32_t)CH, len - c * CH);
    const int r = (int)(c % 3), bb = (int)(c & 1u);
    const int j0 = wave - 1;
    float hc = 0.0f;
    float ar[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int k = min(2 * (j0 + 3 * m) + half, CH - 1);  // rows >= CH read splat CH - 1
      const float4 A = qA[r][k];
      const float2 Bv = qB[r][k];
      ar[m] = alpha_capped(splat_power(A.x, A.y, A.z, A.w, Bv.x, Bv.y, fx, fy));
    }
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int k = 2 * (j0 + 3 * m) + half;  // < CH + 4
      const bool above = !(ar[m] < 1.0f / 255.0f);
      sAl[bb][k][p] = above && inside ? ar[m] : 0.0f;
      const uint64_t hitm = __builtin_amdgcn_ballot_w64(above) & inside_mask;
      const int lo = (int)__builtin_popcount((uint32_t)hitm);
      const int hi = (int)__builtin_popcount((uint32_t)(hitm >> 32));
      asm("v_writelane_b32 %0, %1, %2" : "+v"(hc) : "s"(lo), "i"(2 * m));
      asm("v_writelane_b32 %0, %1, %2" : "+v"(hc) : "s"(hi), "i"(2 * m + 1));
    }
    const int k = 2 * (j0 + 3 * (lane >> 1)) + (lane & 1);
    if (lane < 2 * PER && k < (int)n) sCnt[bb][k] = (uint32_t)__builtin_bit_cast(int, hc);
  };
  Rec R;       // records of chunk c + 2 at the start of step c
  uint32_t I;  // ids of chunk c + 3 at the start of step c
  {
    const uint32_t i0 = load_id(0), i1 = load_id(1), i2 = load_id(2);
    I = load_id(3);
    const Rec r0 = load_rec(0, i0), r1 = load_rec(1, i1);
    R = load_rec(2, i2);
    stage(0, r0);
    stage(1, r1);
  }
  __syncthreads();
  if (wave != 0) produce(0);
  __syncthreads();
  float T = 1.0f, wmax = 0.0f;
  uint32_t imax = 0;
  for (uint32_t c = 0; c < nch; ++c) {
    if (wave == 0) {
      const uint32_t p0 = c * CH;
      const uint32_t n = min((uint32_t)CH, len - p0);
      const int bb = (int)(c & 1u);
      float w[CH];
      int kmax = -1;
      if (n == (uint32_t)CH)
        simp_piece_chain<true>(sAl[bb], n, p, inside, T, wmax, kmax, w);
      else
        simp_piece_chain<false>(sAl[bb], n, p, inside, T, wmax, kmax, w);
      if (kmax >= 0) imax = p0 + (uint32_t)kmax;
      const float ws = wave_transpose_sum32(w);
      const uint32_t k = (uint32_t)lane >> 1;
      if ((lane & 1) == 0 && k < n) {
        const uint32_t hits = sCnt[bb][k];
        if (hits != 0u) {
          const uint32_t id = qId[c % 3][k];
          atomic_add_f(&wsum[id], ws);
          atomicAdd(&wcount[id], (int32_t)hits);
        }
      }
    } else {
      produce(c + 1);
      stage(c + 2, R);
      R = load_rec(c + 3, I);
      I = load_id(c + 4);
    }
    __syncthreads();
  }
  if (wave == 0) {
    const bool has = inside && wmax > 0.0f;
    simp_max_flush(has, has ? vals[start + imax] : 0u, maxcount);
  }
}

__global__ __launch_bounds__(256) void k_render_simp(SimpGrid g,
                                                     const uint32_t* __restrict__ ranges,
                                                     const uint32_t* __restrict__ vals,
                                                     const float* __restrict__ splat,
                                                     float* __restrict__ wsum,
                                                     int32_t* __restrict__ wcount,
                                                     float* __restrict__ maxcount) {
  const int b = (int)blockIdx.x;
  if (b < g.n_pieces) {
    simp_piece(b, g, ranges, vals, splat, wsum, wcount, maxcount);
  } else {
    const int i = b - g.n_pieces;
    simp_tile((i / g.gxi) * g.gx + i % g.gxi, g, ranges, vals, splat, wsum, wcount, maxcount);
  }
}

hipError_t launch_render_simp(int W, int H, int tw, int T, const uint32_t* ranges,
                              const uint32_t* vals, const float* splat, float* wsum,
                              int32_t* wcount, float* maxcount, hipStream_t s) {
  if (T <= 0) return hipSuccess;
  const SimpGrid g = simp_grid(W, H, tw);
  const int grid = g.n_pieces + g.gxi * g.gyi;
  if (grid > 0)
    FGS_TIMED("render_simp", s,
              (k_render_simp<<<grid, 256, 0, s>>>(g, ranges, vals, splat, wsum, wcount, maxcount)));
  return hipGetLastError();
}

// ----------------------------------------------------------- render_depth
// Per-pixel depth over a rendered frame's lists (render_depth.cu:92-290):
// one 256-thread workgroup per tile, one pixel per thread. Each thread
// builds its pixel's camera ray (inverse projection, once per thread);
// splats are staged 256 at a time in LDS together with the per-Gaussian
// part of the ray/ellipsoid quadratic (rotation, rotated camera offset,
// 3 sigma, constant term), which depends on the Gaussian and the camera
// only. A pixel stops when T (1 - alpha) < 1e-4; the workgroup stops once
// all its pixels have stopped. Every pixel inside the image is written.
__global__ __launch_bounds__(256) void k_render_depth(
    int W, int H, int grid_x, int tw, const uint32_t* __restrict__ ranges,
    const uint32_t* __restrict__ vals, const float* __restrict__ splat,
    const float* __restrict__ means3D, const float* __restrict__ scales,
    const float* __restrict__ rotations, const float* __restrict__ proj,
    const float* __restrict__ campos, float* __restrict__ out_pts, float* __restrict__ out_depth,
    float* __restrict__ accum_alpha, int32_t* __restrict__ gidx, float* __restrict__ disc) {
  __shared__ float4 sA[256];
  __shared__ float2 sB[256];
  __shared__ uint32_t sId[256];
  __shared__ Ellipsoid<float> sE[256];
  const int tile = blockIdx.x;
  const int tx = tile % grid_x, ty = tile / grid_x;
  const int t = threadIdx.x;
  const int px = tx * tw + (t & (tw - 1)), py = ty * tile_h_of(tw) + (t >> tile_w_log2(tw));
  const bool inside = px < W && py < H;
  const float fx = (float)px, fy = (float)py;
  const float o[3] = {campos[0], campos[1], campos[2]};
  float nd[3], len;
  {
    float m[16], inv[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = proj[k];
    inverse4(m, inv);
    depth_ray(inv, o, px, py, W, H, nd, len);
  }
  const uint32_t start = ranges[2 * tile], end = ranges[2 * tile + 1];
  const uint32_t len_list = end - start;
  DepthPix<float> D;
  D.done = !inside;
  for (uint32_t base = 0; base < len_list; base += 256) {
    if (__syncthreads_count(D.done) == 256) break;
    const uint32_t n = min(256u, len_list - base);
    if (t < (int)n) {
      const uint32_t id = vals[start + base + t];
      const float4* r = (const float4*)(splat + 12 * (size_t)id);
      const float4 r1 = r[1];
      sA[t] = r[0];
      sB[t] = make_float2(r1.x, r1.y);
      sId[t] = id;
      const float mean[3] = {means3D[3 * (size_t)id], means3D[3 * (size_t)id + 1],
                             means3D[3 * (size_t)id + 2]};
      const float sc[3] = {scales[3 * (size_t)id], scales[3 * (size_t)id + 1],
                           scales[3 * (size_t)id + 2]};
      const float q[4] = {rotations[4 * (size_t)id], rotations[4 * (size_t)id + 1],
                          rotations[4 * (size_t)id + 2], rotations[4 * (size_t)id + 3]};
      Ellipsoid<float> E;
      ellipsoid_setup(mean, sc, q, o, E);
      sE[t] = E;
    }
    __syncthreads();
    for (uint32_t j = 0; j < n && !D.done; ++j) {
      const float4 A = sA[j];
      const float2 B = sB[j];
      const float a = splat_alpha(A.x, A.y, A.z, A.w, B.x, B.y, fx, fy);
      if (a == 0.0f) continue;  // alpha < 1/255
      depth_pix_step(D, a, sId[j], sE[j], o, nd, len);
    }
  }
  if (inside) {
    const size_t HW = (size_t)W * H, pix = (size_t)py * W + px;
    out_pts[pix] = D.pt[0];
    out_pts[HW + pix] = D.pt[1];
    out_pts[2 * HW + pix] = D.pt[2];
    out_depth[pix] = D.depth;
    accum_alpha[pix] = D.T;
    gidx[pix] = D.idx;
    disc[pix] = D.disc;
  }
}

hipError_t launch_render_depth(int W, int H, int tw, int T, const uint32_t* ranges,
                               const uint32_t* vals,
                               const float* splat, const float* means3D, const float* scales,
                               const float* rotations, const float* proj, const float* campos,
                               float* out_pts, float* out_depth, float* accum_alpha,
                               int32_t* gidx, float* disc, hipStream_t s) {
  if (T > 0)
    FGS_TIMED("render_depth", s,
              (k_render_depth<<<T, 256, 0, s>>>(W, H, grid_dim(W, tw), tw, ranges, vals,
                                                 splat, means3D, scales, rotations, proj, campos,
                                                 out_pts, out_depth, accum_alpha, gidx, disc)));
  return hipGetLastError();
}

}  // namespace fgs
