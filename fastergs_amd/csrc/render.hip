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

// Ranges from the sorted tile ids (16- or 32-bit) by binary search, one
// thread per tile: ranges[t] = [lower_bound(t), lower_bound(t + 1)).
template <typename TK>
__device__ __forceinline__ uint32_t lower_bound_tile(const TK* a, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((uint32_t)a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}
// Also writes each tile's bucket count ceil(len / bucket) and clears its
// furthest-contributor slot, which compositing combines with atomicMax when
// a tile is split into row bands; the last workgroup to finish scans the
// bucket counts into boff (T + 1 entries, total at [T]). Workgroup b looks
// up the lower bounds of tiles 255 b .. 255 b + 255 (one search per
// thread) and writes the ranges of the first 255 of them.
constexpr int RANGES_PER_BLOCK = 255;
template <typename TK>
__global__ __launch_bounds__(256) void k_tile_ranges(int64_t R_cap,
                                                     const uint32_t* __restrict__ n_dev,
                                                     const TK* __restrict__ tiles,
                                                     uint32_t* __restrict__ ranges, int T,
                                                     int bucket, uint32_t* __restrict__ bcount,
                                                     uint32_t* __restrict__ maxc,
                                                     uint32_t* __restrict__ boff,
                                                     uint32_t* __restrict__ arrive) {
  __shared__ uint32_t s_lo[256];
  const int t = blockIdx.x * RANGES_PER_BLOCK + threadIdx.x;
  const uint32_t R = (uint32_t)dev_count(n_dev, R_cap);
  if (t <= T) s_lo[threadIdx.x] = lower_bound_tile(tiles, R, (uint32_t)t);
  __syncthreads();
  if (threadIdx.x < RANGES_PER_BLOCK && t < T) {
    const uint32_t lo = s_lo[threadIdx.x], hi = s_lo[threadIdx.x + 1];
    // empty tiles keep [0, 0) like identifyTileRanges (render.cu:406-411)
    ranges[2 * t] = lo < hi ? lo : 0u;
    ranges[2 * t + 1] = lo < hi ? hi : 0u;
    publish_u32(&bcount[t], (hi - lo + bucket - 1) / bucket);
    maxc[t] = 0u;
  }
  if (last_workgroup(arrive))
    block_scan_into<256>(T, [&](int i) { return read_published_u32(&bcount[i]); }, boff);
}
hipError_t launch_tile_ranges(int64_t R, const uint32_t* n_dev, const void* tiles, bool t16,
                              uint32_t* ranges, int T, int bucket, uint32_t* bcount,
                              uint32_t* maxc, uint32_t* boff, uint32_t* arrive, hipStream_t s) {
  if (T <= 0) return hipSuccess;
  const unsigned grid = (unsigned)((T + RANGES_PER_BLOCK - 1) / RANGES_PER_BLOCK);
  if (t16)
    FGS_TIMED("tile_ranges", s,
              (k_tile_ranges<uint16_t><<<grid, 256, 0, s>>>(R, n_dev, (const uint16_t*)tiles,
                                                            ranges, T, bucket, bcount, maxc,
                                                            boff, arrive)));
  else
    FGS_TIMED("tile_ranges", s,
              (k_tile_ranges<uint32_t><<<grid, 256, 0, s>>>(R, n_dev, (const uint32_t*)tiles,
                                                            ranges, T, bucket, bcount, maxc,
                                                            boff, arrive)));
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
  const int px = tx * TILE_W + (t & 15), py = ty * TILE_H + (t >> 4);
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
// tile for its rows). The item's buckets are composited in rounds; every
// bucket from T = 1 (render.cu:130-275), its splats staged in LDS. The
// partials (T_b, C_b) of a round's buckets are then merged in order per
// pixel (render.cu:329-362), the checkpoint after each bucket is stored for
// the backward (unless the frame is forward-only), and a pixel stops after
// the first bucket that leaves T < 1e-4. The workgroup stops at the first
// round in which every pixel of the item has stopped, so buckets past the
// cutoff (which the reference renders and then discards) are not evaluated.
//
// Round widths ramp up, 1, 2, then 4 buckets for a whole tile (2, then 4 for
// a band): in a round of K buckets, 4 / K waves share each bucket, each
// taking a part of the item's pixels (1, 2 or 4 per lane), so a tile whose
// pixels all stop within the first buckets evaluates little beyond them,
// while later rounds keep four buckets in flight. Every pixel's partial of a
// bucket is the same sequential blend wherever it is computed, so outputs do
// not depend on the round widths, the item order or the splitting.

// Sequential blend of n staged splats from T = 1 for the NPX pixels of a
// lane: (x0 + (k & 1), y0 + (k >> 1)) for k < NPX; partials into out[lp0 +
// (k & 1) + TILE_W (k >> 1)].
template <int NPX>
__device__ __forceinline__ void bucket_pass(const float4* sA, const float4* sB, const float* sC,
                                            int n, float fx0, float fy0, float4* out, int lp0) {
  const float fx1 = fx0 + 1.0f, fy1 = fy0 + 1.0f;
  float T0 = 1.f, T1 = 1.f, T2 = 1.f, T3 = 1.f;
  float c00 = 0.f, c01 = 0.f, c02 = 0.f, c10 = 0.f, c11 = 0.f, c12 = 0.f;
  float c20 = 0.f, c21 = 0.f, c22 = 0.f, c30 = 0.f, c31 = 0.f, c32 = 0.f;
  for (int k = 0; k < n; ++k) {
    const float4 A = sA[k], Bv = sB[k];
    const float cb = sC[k];
    blend(splat_alpha(A.x, A.y, A.z, A.w, Bv.x, Bv.y, fx0, fy0), Bv.z, Bv.w, cb, T0, c00, c01, c02);
    if (NPX >= 2)
      blend(splat_alpha(A.x, A.y, A.z, A.w, Bv.x, Bv.y, fx1, fy0), Bv.z, Bv.w, cb, T1, c10, c11, c12);
    if (NPX == 4) {
      blend(splat_alpha(A.x, A.y, A.z, A.w, Bv.x, Bv.y, fx0, fy1), Bv.z, Bv.w, cb, T2, c20, c21, c22);
      blend(splat_alpha(A.x, A.y, A.z, A.w, Bv.x, Bv.y, fx1, fy1), Bv.z, Bv.w, cb, T3, c30, c31, c32);
    }
  }
  out[lp0] = make_float4(c00, c01, c02, T0);
  if (NPX >= 2) out[lp0 + 1] = make_float4(c10, c11, c12, T1);
  if (NPX == 4) {
    out[lp0 + TILE_W] = make_float4(c20, c21, c22, T2);
    out[lp0 + TILE_W + 1] = make_float4(c30, c31, c32, T3);
  }
}

template <int BMAX>
__global__ __launch_bounds__(256) void k_composite_fast(CompArgs a) {
  __shared__ float4 sA[4][BMAX];
  __shared__ float4 sB[4][BMAX];
  __shared__ float sC[4][BMAX];
  __shared__ float4 sP[4][TILE_PIX];
  __shared__ int wmax[4];
  const uint64_t t_start = a.cycles ? wall_clock64() : 0;
  uint32_t item;
  int log_idx;  // where the duration is logged: by work item, by (slot, tile) or by tile
  if (a.items) {
    if (blockIdx.x >= a.n_items[0]) return;  // whole workgroup
    item = a.items[blockIdx.x];
    log_idx = (int)blockIdx.x;
  } else if (a.slots) {
    // two grid slots per tile: the upper and lower band of a split tile, or
    // the whole tile in slot 0. Workgroup b runs on XCD b % 8, so the slots
    // are laid out [all first slots | all second slots]: interleaving them
    // would leave the mostly idle second slots to half of the XCDs.
    const int slot = (int)blockIdx.x >= a.T ? 1 : 0;
    const int tl = (int)blockIdx.x - slot * a.T;

    const uint32_t len = a.ranges[2 * tl + 1] - a.ranges[2 * tl];
    const bool split = a.split_pred ? (a.split_pred[2 * tl + 1] > 0.0f && len > (uint32_t)a.bucket)
                                    : (a.split_len > 0 && len >= (uint32_t)a.split_len);
    log_idx = slot * a.T + tl;
    if (!split && slot) {
      if (a.cycles && threadIdx.x == 0) a.cycles[log_idx] = 0u;
      return;  // whole workgroup
    }
    item = ((uint32_t)tl << 2) | (split ? 1u + (uint32_t)slot : 0u);
  } else {
    item = (uint32_t)blockIdx.x << 2;
    log_idx = (int)blockIdx.x;
  }
  const int tile = (int)(item >> 2), part = (int)(item & 3u);
  const bool band = part != 0;
  const int row0 = part == 2 ? TILE_H / 2 : 0;
  const int nrows = band ? TILE_H / 2 : TILE_H;
  // the tile's view: its size, tile grid, background and image offsets
  const ViewRec& vr = a.views[a.V > 1 ? view_of_tile(a.views, a.V, (uint32_t)tile) : 0];
  const int W = vr.W, H = vr.H, gx = vr.grid_x;
  const int lt = tile - (int)vr.tile_off;
  const int tx = lt % gx, ty = lt / gx;
  const int t = threadIdx.x;
  const int w = t / WAVE, lane = t & (WAVE - 1);
  // merge / output pixel of this thread (lp: index inside the tile)
  const int mrow = row0 + (t >> 4);
  const int lp = mrow * TILE_W + (t & 15);
  const int px = tx * TILE_W + (t & 15), py = ty * TILE_H + mrow;
  const bool merger = (t >> 4) < nrows;
  const bool inside = merger && px < W && py < H;
  const uint32_t start = a.ranges[2 * tile], end = a.ranges[2 * tile + 1];
  const int B = a.bucket;
  const int nb = (int)((end - start + B - 1) / B);
  const uint32_t b0 = a.boff[tile];
  float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
  int cnt = 0;
  bool done = !inside;
  int first = 0;  // first bucket of the round
  for (int round = 0; first < nb; ++round) {
    if (__syncthreads_count(!done) == 0) break;
    // buckets this round (K) and waves per bucket (G = 4 / K)
    const int K = band ? (round == 0 ? 2 : 4) : (round == 0 ? 1 : (round == 1 ? 2 : 4));
    const int G = 4 / K;
    const int q = w / G, sub = w % G;  // this wave's bucket in the round, its share
    const int bl = first + q;
    const uint32_t s = start + (uint32_t)bl * B;
    const int n = bl < nb ? (int)min((uint32_t)B, end - s) : 0;
    for (int k = sub * WAVE + lane; k < n; k += G * WAVE) {
      const float4* r = (const float4*)(a.splat + 12 * (size_t)a.vals[s + k]);
      const float4 r1 = r[1], r2 = r[2];
      sA[q][k] = r[0];
      sB[q][k] = make_float4(r1.x, r1.y, r2.x, r2.y);  // cc, log2 op, r, g
      sC[q][k] = r2.z;
    }
    __syncthreads();
    if (n > 0) {
      // this wave's pixels: 256 / G of the tile (128 / G of a band)
      const int pxl = band ? 128 / G : 256 / G;
      if (pxl == 64) {  // one per lane: 4 rows of 16
        const int qx = lane & 15, qy = (band ? row0 : 0) + 4 * sub + (lane >> 4);
        bucket_pass<1>(sA[q], sB[q], sC[q], n, (float)(tx * TILE_W + qx),
                       (float)(ty * TILE_H + qy), sP[q], qy * TILE_W + qx);
      } else if (pxl == 128) {  // two per lane: 8 rows of 16 in 2x1 pairs
        const int qx = 2 * (lane & 7), qy = (band ? row0 : 8 * sub) + (lane >> 3);
        bucket_pass<2>(sA[q], sB[q], sC[q], n, (float)(tx * TILE_W + qx),
                       (float)(ty * TILE_H + qy), sP[q], qy * TILE_W + qx);
      } else {  // four per lane: the tile in 2x2 quads
        const int qx = 2 * (lane & 7), qy = 2 * (lane >> 3);
        bucket_pass<4>(sA[q], sB[q], sC[q], n, (float)(tx * TILE_W + qx),
                       (float)(ty * TILE_H + qy), sP[q], qy * TILE_W + qx);
      }
    }
    __syncthreads();
    if (!done) {
      for (int j = 0; j < K; ++j) {
        const int bq = first + j;
        if (bq >= nb) break;
        const float4 v = sP[j][lp];
        merge(v.w, v.x, v.y, v.z, T, C0, C1, C2);
        if (a.acc)
          ((float4*)a.acc)[(size_t)(b0 + bq) * TILE_PIX + lp] = make_float4(C0, C1, C2, T);
        cnt += (int)min((uint32_t)B, end - (start + (uint32_t)bq * B));
        if (T < T_CUT_F) {
          done = true;
          break;
        }
      }
    }
    first += K;
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
  const int m = wave_max_i(cnt);
  if (lane == 0) wmax[w] = m;
  __syncthreads();
  if (t == 0) {
    const uint32_t mx = (uint32_t)max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
    // the two bands of a split tile combine into the entry k_tile_ranges cleared
    if (band)
      atomicMax(&a.maxc[tile], mx);
    else
      a.maxc[tile] = mx;
    if (a.cycles) a.cycles[log_idx] = (uint32_t)(wall_clock64() - t_start);
  }
}

hipError_t launch_composite_fast(const CompArgs& a, hipStream_t s) {
  if (a.T <= 0) return hipSuccess;
  const unsigned grid = (a.items || a.slots) ? 2u * (unsigned)a.T : (unsigned)a.T;
  if (a.bucket <= 64)
    FGS_TIMED("composite_fast", s, (k_composite_fast<64><<<grid, 256, 0, s>>>(a)));
  else if (a.bucket <= 128)
    FGS_TIMED("composite_fast", s, (k_composite_fast<128><<<grid, 256, 0, s>>>(a)));
  else
    FGS_TIMED("composite_fast", s, (k_composite_fast<256><<<grid, 256, 0, s>>>(a)));
  return hipGetLastError();
}

// ----------------------------------------------------------- render_simp
// Importance accumulation over a rendered frame's lists (render_simp.cu:36-301):
// one 256-thread workgroup per tile, one pixel per thread, splats staged 256
// at a time in LDS. Pixels blend front to back without stopping on their
// own; after each pair of splats (a leading single splat when the list
// length is odd) the tile stops once all 256 pixels of the tile -- those
// outside the image keep T = 1 -- have T < 1e-4. Per splat the weights
// alpha*T and the number of pixels with alpha >= 1/255 are summed; per pixel
// the splat of largest weight (the first of equals) gets +1 at the end.
//
// No barrier per pair: T only decreases, so each wave notes the first pair
// end at which all its 64 pixels are below 1e-4, and the tile's stop point
// is the largest of the four. Waves run through the whole batch; the
// per-splat sums (DPP wave sum, ballot count) go to LDS with LDS atomics and
// are flushed to memory once per batch, only for splats up to the stop
// point. Past the stop point every weight is < 1e-4 while each pixel's
// largest weight is >= 1/255 (its first contributor meets T = 1), so the
// per-pixel maxima need no such cut.
__global__ __launch_bounds__(256) void k_render_simp(int W, int H, int grid_x,
                                                     const uint32_t* __restrict__ ranges,
                                                     const uint32_t* __restrict__ vals,
                                                     const float* __restrict__ splat,
                                                     float* __restrict__ wsum,
                                                     int32_t* __restrict__ wcount,
                                                     float* __restrict__ maxcount) {
  __shared__ float4 sA[256];
  __shared__ float2 sB[256];
  __shared__ uint32_t sId[256];
  __shared__ float sW[256];
  __shared__ uint32_t sC[256];
  __shared__ uint32_t sStop[4];
  const int tile = blockIdx.x;
  const int tx = tile % grid_x, ty = tile / grid_x;
  const int t = threadIdx.x;
  const int wave = t >> 6;
  const int px = tx * TILE_W + (t & 15), py = ty * TILE_H + (t >> 4);
  const bool inside = px < W && py < H;
  const float fx = (float)px, fy = (float)py;
  const uint32_t start = ranges[2 * tile], end = ranges[2 * tile + 1];
  const uint32_t len = end - start;
  float T = 1.0f, wmax = 0.0f;
  uint32_t imax = 0;
  bool flag = false;
  const int lane = lane_id();
  uint32_t wave_stop = 0xffffffffu;  // first pair end with all 64 pixels below the cut
  for (uint32_t base = 0; base < len; base += 256) {
    const uint32_t n = min(256u, len - base);
    if (t < (int)n) {
      const uint32_t id = vals[start + base + t];
      const float4* r = (const float4*)(splat + 12 * (size_t)id);
      const float4 r1 = r[1];
      sA[t] = r[0];
      sB[t] = make_float2(r1.x, r1.y);
      sId[t] = id;
    }
    sW[t] = 0.0f;
    sC[t] = 0u;
    __syncthreads();
    for (uint32_t j = 0; j < n; ++j) {
      float w = 0.0f;
      bool hit = false;
      if (inside) {
        const float4 A = sA[j];
        const float2 B = sB[j];
        const float a = splat_alpha(A.x, A.y, A.z, A.w, B.x, B.y, fx, fy);
        if (a != 0.0f) {  // alpha >= 1/255
          w = a * T;
          T = T - w;
          hit = true;
          if (wmax < w) {
            wmax = w;
            imax = sId[j];
            flag = true;
          }
        }
      }
      const uint64_t hits = __ballot(hit);
      if (hits) {
        const float ws = wave_sum(w);
        if (lane == 0) {
          atomicAdd(&sW[j], ws);
          atomicAdd(&sC[j], (uint32_t)__popcll(hits));
        }
      }
      // pairs follow the leading single of an odd-length list: position pos
      // ends a pair when it is odd (even length) or even and >= 2 (odd length)
      const uint32_t pos = base + j;
      const bool pair_end = (len & 1u) ? (pos >= 2u && (pos & 1u) == 0u) : ((pos & 1u) == 1u);
      if (pair_end && wave_stop == 0xffffffffu && __ballot(T < T_CUT_F) == ~0ull) wave_stop = pos;
    }
    if (lane == 0) sStop[wave] = wave_stop;
    __syncthreads();
    const uint32_t stop = max(max(sStop[0], sStop[1]), max(sStop[2], sStop[3]));
    if (t < (int)n && base + t <= stop && sC[t] != 0u) {
      atomic_add_f(&wsum[sId[t]], sW[t]);
      atomicAdd(&wcount[sId[t]], (int32_t)sC[t]);
    }
    if (stop != 0xffffffffu) break;
    __syncthreads();
  }
  // one atomic per distinct maximum in the wave (neighbouring pixels mostly
  // share it; the counts are integers, so the float sums stay exact)
  const bool has = inside && flag;
  uint64_t pending = __ballot(has);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)imax, leader);
    const uint64_t same = __ballot(has && imax == k) & pending;
    if (lane == leader) atomic_add_f(&maxcount[k], (float)__popcll(same));
    pending &= ~same;
  }
}

hipError_t launch_render_simp(int W, int H, int T, const uint32_t* ranges, const uint32_t* vals,
                              const float* splat, float* wsum, int32_t* wcount, float* maxcount,
                              hipStream_t s) {
  if (T > 0)
    FGS_TIMED("render_simp", s,
              (k_render_simp<<<T, 256, 0, s>>>(W, H, (W + TILE_W - 1) / TILE_W, ranges, vals, splat,
                                                wsum, wcount, maxcount)));
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
    int W, int H, int grid_x, const uint32_t* __restrict__ ranges,
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
  const int px = tx * TILE_W + (t & 15), py = ty * TILE_H + (t >> 4);
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

hipError_t launch_render_depth(int W, int H, int T, const uint32_t* ranges, const uint32_t* vals,
                               const float* splat, const float* means3D, const float* scales,
                               const float* rotations, const float* proj, const float* campos,
                               float* out_pts, float* out_depth, float* accum_alpha,
                               int32_t* gidx, float* disc, hipStream_t s) {
  if (T > 0)
    FGS_TIMED("render_depth", s,
              (k_render_depth<<<T, 256, 0, s>>>(W, H, (W + TILE_W - 1) / TILE_W, ranges, vals,
                                                 splat, means3D, scales, rotations, proj, campos,
                                                 out_pts, out_depth, accum_alpha, gidx, disc)));
  return hipGetLastError();
}

}  // namespace fgs
