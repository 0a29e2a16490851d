// Preprocess and pair emission kernels.
//
// Per Gaussian (reference: preprocessSmallCUDA, FGS/csrc/cuda_rasterizer/
// preprocess.cu:414-746): cull, project, 3D/2D covariance, conic, radius,
// opacity-aware half extents, tile rect, SH colour; then the exact
// tile/ellipse test (preprocess.cu:600-607) decides which tiles get a pair.
//
// Records written per Gaussian (structure of arrays, 16-byte aligned rows):
//   geom  8 floats: x, y, conic a, b, c, power, rect min, rect max (packed u16x2)
//   splat 12 floats: x, y, ca, cb | cc, log2 op, depth, 0 | r, g, b, clamp bits
//   (the colour is one aligned float4, written by the colour pass in one store)
// The splat row is what compositing and its backward read (48 B per
// Gaussian; the reference reads fields 0-9 of a 128 B row).
//
// PLAIN schedule: one thread per Gaussian does the math and then walks its
// whole rect twice (count, emit) -- large splats are stragglers.
// FAST schedule: the math is one thread per Gaussian; the tile walk is
// spread over the wave: rects below `split_tiles` tiles are packed 64 tiles
// per wave step across lanes, larger rects become tile-bounded tasks that
// whole waves take from a queue.
#include <hip/hip_runtime.h>

#include "fgs_device.h"
#include "fgs_kernels.h"
#include "fgs_prof.h"
#include "fgs_layout.h"

namespace fgs {

__device__ __forceinline__ void store_records(size_t i, const Proj<float>& o, float* geom,
                                              float* splat) {
  float4* g4 = (float4*)(geom + 8 * (size_t)i);
  const uint32_t rmin = (uint32_t)o.rect[0] | ((uint32_t)o.rect[1] << 16);
  const uint32_t rmax = (uint32_t)o.rect[2] | ((uint32_t)o.rect[3] << 16);
  g4[0] = make_float4(o.x, o.y, o.conic[0], o.conic[1]);
  g4[1] = make_float4(o.conic[2], o.power, u2f(rmin), u2f(rmax));
  float ca, cb, cc, w;
  log2_conic(o.conic, o.log2op, ca, cb, cc, w);
  float4* s4 = (float4*)(splat + 12 * (size_t)i);
  s4[0] = make_float4(o.x, o.y, ca, cb);
  s4[1] = make_float4(cc, w, o.depth, 0.0f);
  s4[2] = make_float4(o.rgb[0], o.rgb[1], o.rgb[2], u2f((uint32_t)o.clamp));
}

__device__ __forceinline__ void store_invalid(size_t i, float* geom, float* splat) {
  float4* g4 = (float4*)(geom + 8 * (size_t)i);
  g4[0] = make_float4(0.f, 0.f, 0.f, 0.f);
  g4[1] = make_float4(0.f, 0.f, 0.f, 0.f);
  float4* s4 = (float4*)(splat + 12 * (size_t)i);
  s4[1] = make_float4(0.f, 0.f, 0.f, 0.f);
  s4[2] = make_float4(0.f, 0.f, 0.f, 0.f);
}

struct GeomRow {
  float x, y, conic[3], power;
  int r0, r1, r2, r3;
};

__device__ __forceinline__ GeomRow load_geom(const float* geom, size_t i) {
  const float4* g4 = (const float4*)(geom + 8 * (size_t)i);
  const float4 a = g4[0], b = g4[1];
  GeomRow r;
  r.x = a.x;
  r.y = a.y;
  r.conic[0] = a.z;
  r.conic[1] = a.w;
  r.conic[2] = b.x;
  r.power = b.y;
  const uint32_t mn = f2u(b.z), mx = f2u(b.w);
  r.r0 = (int)(mn & 0xffffu);
  r.r1 = (int)(mn >> 16);
  r.r2 = (int)(mx & 0xffffu);
  r.r3 = (int)(mx >> 16);
  return r;
}

// ------------------------------------------------------------------ PLAIN
__global__ __launch_bounds__(256) void k_preprocess_plain(PreArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.P) return;
  Proj<float> o;
  preprocess_gauss(i, a.G, a.cam, o);
  a.radii[i] = o.valid ? o.radius : 0;
  if (!o.valid) {
    a.pcount[i] = 0;
    a.dkey[i] = 0xffffffffu;
    store_invalid(i, a.geom, a.splat);
    return;
  }
  uint32_t n = 0;
  for (int ty = o.rect[1]; ty < o.rect[3]; ++ty)
    for (int tx = o.rect[0]; tx < o.rect[2]; ++tx)
      n += tile_hit(tx, ty, a.cam.tile_w, a.cam.tile_h, o.x, o.y, o.conic, o.power) ? 1u : 0u;
  a.pcount[i] = n;
  a.dkey[i] = depth_bits(o.depth);
  store_records(i, o, a.geom, a.splat);
}

hipError_t launch_preprocess_plain(const PreArgs& a, hipStream_t s) {
  if (a.P > 0) FGS_TIMED("preprocess_plain", s, (k_preprocess_plain<<<(a.P + 255) / 256, 256, 0, s>>>(a)));
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_duplicate_plain(int P, const float* __restrict__ geom,
                                                         const uint32_t* __restrict__ dkey,
                                                         const uint32_t* __restrict__ pcount,
                                                         const uint32_t* __restrict__ poff,
                                                         int grid_x, int tw, int th,
                                                         uint32_t cap,
                                                         uint64_t* __restrict__ keys,
                                                         uint32_t* __restrict__ vals) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P || pcount[i] == 0) return;
  const GeomRow g = load_geom(geom, i);
  uint32_t k = poff[i];
  const uint64_t d = dkey[i];
  for (int ty = g.r1; ty < g.r3; ++ty)
    for (int tx = g.r0; tx < g.r2; ++tx)
      if (tile_hit(tx, ty, tw, th, g.x, g.y, g.conic, g.power)) {
        if (k < cap) {
          keys[k] = ((uint64_t)(uint32_t)(ty * grid_x + tx) << 32) | d;
          vals[k] = (uint32_t)i;
        }
        ++k;
      }
}

hipError_t launch_duplicate_plain(int P, const float* geom, const uint32_t* dkey,
                                  const uint32_t* pcount, const uint32_t* poff, int grid_x,
                                  int tw, uint32_t cap, uint64_t* keys, uint32_t* vals,
                                  hipStream_t s) {
  if (P > 0)
    FGS_TIMED("duplicate_plain", s,
              (k_duplicate_plain<<<(P + 255) / 256, 256, 0, s>>>(P, geom, dkey, pcount, poff, grid_x,
                                                                 tw, tile_h_of(tw), cap, keys,
                                                                 vals)));
  return hipGetLastError();
}

// ------------------------------------------------------------------- FAST
// Per-wave staging of the 64 Gaussians whose small rects are walked
// together: centre, conic, power, rect origin / width and the inclusive
// prefix of rect areas over the wave.
struct WaveTiles {
  float x[WAVE], y[WAVE], c0[WAVE], c1[WAVE], c2[WAVE], pw[WAVE];
  int r0[WAVE], r1[WAVE], w[WAVE];
  uint32_t incl[WAVE], area[WAVE], cnt[WAVE];
};
// Emission also stages each Gaussian's record, first pair slot, view tile
// offset and tile-grid width.
struct WaveEmit {
  uint32_t rec[WAVE], pos[WAVE], toff[WAVE];
  int gx[WAVE];
};

// Row bands of a large rect of wr x hr tiles: up to LARGE_BANDS bands of
// `rows` whole tile rows, about LARGE_BAND_TILES tiles each (none for an
// empty rect).
__device__ __forceinline__ void large_bands(int wr, int hr, int& rows, int& nbands) {
  if (wr <= 0 || hr <= 0) {
    rows = 1;
    nbands = 0;
    return;
  }
  const int area = wr * hr;
  int want = (area + LARGE_BAND_TILES - 1) / LARGE_BAND_TILES;
  want = want < 1 ? 1 : (want > LARGE_BANDS ? LARGE_BANDS : want);
  rows = (hr + want - 1) / want;
  nbands = (hr + rows - 1) / rows;
}

// One emitted pair: its (global) tile id, 16- or 32-bit, and its record.
__device__ __forceinline__ void store_pair(void* tkeys, uint32_t* recs, bool t16, uint32_t pos,
                                           uint32_t tile, uint32_t rec) {
  if (t16)
    ((uint16_t*)tkeys)[pos] = (uint16_t)tile;
  else
    ((uint32_t*)tkeys)[pos] = tile;
  recs[pos] = rec;
}

// Walk the concatenated small rects of one wave 64 candidate tiles per step
// (lane-packed); count (EMIT=false) or emit (EMIT=true, E staged) the pairs
// that pass the exact tile test on tw x th tiles. Within a Gaussian's block
// of the pair array the order is free: its pairs are in distinct tiles and
// the tile sort is stable.
template <bool EMIT>
__device__ __forceinline__ void walk_small(WaveTiles& L, const WaveEmit* E, uint32_t S,
                                           uint32_t cap, void* __restrict__ tkeys,
                                           uint32_t* __restrict__ recs, bool t16, int tw,
                                           int th) {
  const int lane = lane_id();
  for (uint32_t base = 0; base < S; base += WAVE) {
    const uint32_t f = base + lane;
    if (f < S) {
      int lo = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (L.incl[lo + step - 1] <= f) lo += step;
      const int j = lo;
      const int l = (int)(f - (L.incl[j] - L.area[j]));
      const int wj = L.w[j];
      const int q = (int)((float)l / (float)wj);  // exact for l < 2^24
      const int tx = L.r0[j] + (l - q * wj), ty = L.r1[j] + q;
      const float conic[3] = {L.c0[j], L.c1[j], L.c2[j]};
      if (tile_hit(tx, ty, tw, th, L.x[j], L.y[j], conic, L.pw[j])) {
        const uint32_t k = atomicAdd(&L.cnt[j], 1u);
        if (EMIT) {
          const uint32_t pos = E->pos[j] + k;
          if (pos < cap)
            store_pair(tkeys, recs, t16, pos, E->toff[j] + (uint32_t)(ty * E->gx[j] + tx),
                       E->rec[j]);
        }
      }
    }
  }
}

__device__ __forceinline__ void stage_small(WaveTiles& L, int lane, bool small, float x, float y,
                                            const float* conic, float power, int r0, int r1,
                                            int r2, int r3, uint32_t& S) {
  const uint32_t area = small ? (uint32_t)((r2 - r0) * (r3 - r1)) : 0u;
  L.x[lane] = x;
  L.y[lane] = y;
  L.c0[lane] = conic[0];
  L.c1[lane] = conic[1];
  L.c2[lane] = conic[2];
  L.pw[lane] = power;
  L.r0[lane] = r0;
  L.r1[lane] = r1;
  L.w[lane] = r2 - r0 > 0 ? r2 - r0 : 1;
  L.area[lane] = area;
  L.cnt[lane] = 0;
  const uint32_t inc = wave_incl_scan(area);
  L.incl[lane] = inc;
  S = (uint32_t)__shfl(inc, WAVE - 1, WAVE);
}

// Geometry pass of the fast preprocess: one thread per Gaussian for the math
// (colour comes from k_sh_color), then the wave's lane-packed tile walk.
// Only 3 KB of LDS per wave64 workgroup, so occupancy is set by registers.
constexpr int PRE_THREADS = WAVE;

// Grid (ceil(P / 64), V): blockIdx.y is the view, records v * P + i.
__device__ __forceinline__ void write_view(const ViewSetup& src, ViewRec* out, int v, int t);

// The cameras come from the kernel arguments (the caller's device
// buffers); workgroup (0, 0) also writes the view table the later passes
// read and clears the frame's counters (zero_n words).
__global__ __launch_bounds__(PRE_THREADS) void k_preprocess_fast(PreArgs a, ViewSetup vs,
                                                                 ViewRec* table, uint32_t* zero,
                                                                 int zero_n) {
  __shared__ WaveTiles L;
  const int v = blockIdx.y;
  if (blockIdx.x == 0 && blockIdx.y == 0) {
    for (int k = threadIdx.x; k < zero_n; k += PRE_THREADS) zero[k] = 0u;
    for (int u = 0; u < vs.n; ++u) write_view(vs, table, u, threadIdx.x);
  }
  const CamRef<float> cam = src_cam(a.cam, vs.v[v]);
  const int i = blockIdx.x * PRE_THREADS + threadIdx.x;
  const int lane = threadIdx.x;
  const size_t ri = (size_t)v * a.P + i;
  Proj<float> o;
  o.valid = 0;
  if (i < a.P) {
    // SH colour is evaluated by k_sh_color; precomputed colours are copied here
    preprocess_gauss<float>(i, a.G, cam, o, nullptr, nullptr, a.G.colors != nullptr);
    a.radii[ri] = o.valid ? o.radius : 0;
    const uint32_t dk = o.valid ? depth_bits(o.depth) : 0xffffffffu;
    a.dkey[ri] = dk;
    // sorted on the high half; ties keep record order (the sort is stable)
    a.dsort_k[ri] = ((uint64_t)dk << 32) | (uint64_t)ri;
    if (o.valid)
      store_records(ri, o, a.geom, a.splat);
    else
      store_invalid(ri, a.geom, a.splat);
  }
  const int area = o.valid ? (o.rect[2] - o.rect[0]) * (o.rect[3] - o.rect[1]) : 0;
  const bool small = o.valid && area < a.split_tiles;
  // large rects: their band counts here, expanded into one list entry per
  // band by a scan (no same-address atomics)
  if (i < a.P) {
    int rows, nbands = 0;
    if (o.valid && !small) large_bands(o.rect[2] - o.rect[0], o.rect[3] - o.rect[1], rows, nbands);
    a.lflag[ri] = (uint32_t)nbands;
  }
  uint32_t S;
  stage_small(L, lane, small, o.x, o.y, o.conic, o.power, o.rect[0], o.rect[1], o.rect[2],
              o.rect[3], S);
  __syncthreads();
  walk_small<false>(L, nullptr, S, 0u, nullptr, nullptr, false, cam.tile_w, cam.tile_h);
  __syncthreads();
  if (i < a.P) {
    a.pcount[ri] = L.cnt[lane];
    // the depth sort carries the small rects' counts; large ones are counted later
    a.dsort_v[ri] = o.valid && !small ? PAIRS_LARGE : L.cnt[lane];
  }
}

// Colour pass: SH degree <= 3 of every valid Gaussian (preprocess.cu:698-707)
// into its splat record (r, g, b, clamp bits). One wave64 per workgroup; the
// dc / SH rows of its 64 Gaussians are staged in LDS with coalesced loads (a
// thread's own 192-byte row would otherwise be 48 scattered loads per wave).
// LDS layout of k_sh_color: dc rows padded to 64 float4, then the rest rows
// padded to SHC_QMAX x 64 float4, so that the staging stores need no
// predicate (which would let the compiler sink each load into its own
// predicated branch and serialise the loads).
constexpr int SHC_DC_F4 = PRE_THREADS;                                       // 64 float4
constexpr int SHC_QMAX = (3 * 15 * PRE_THREADS / 4 + PRE_THREADS - 1) / PRE_THREADS;  // 12
constexpr size_t SHC_SHMEM = (size_t)(SHC_DC_F4 + SHC_QMAX * PRE_THREADS) * sizeof(float4);
// 6 per CU of the 256 (measured at 1M Gaussians, 1280x840: 2, 3, 4, 6, 8
// per CU and an unbounded grid give 0.996, 0.966, 0.957, 0.950, 0.968 and
// 0.971 ms per training step)
constexpr unsigned SHC_BLOCKS = 256 * 6;

// One group of 64 Gaussians of k_sh_color; JAC: also the SH Jacobians.
template <bool JAC>
__device__ __forceinline__ void sh_color_group(const PreArgs& a, float4* rows4, int grp) {
  float* rows = (float*)rows4;
  const int lane = threadIdx.x;
  const int g0 = grp * PRE_THREADS;
  const int i = g0 + lane;
  const int M = a.cam.max_coeffs;
  float* sdc = rows;                    // 64 x 3 dc
  float* ssh = rows + 4 * SHC_DC_F4;    // 64 x 3M rest
  const int nb = min(PRE_THREADS, a.P - g0);
  const float* gdc = a.G.dc + 3 * (size_t)g0;
  const float* gsh = a.G.shs + (size_t)3 * M * g0;
  // the mean is independent of the staging: start its load first
  float pm[3] = {0.0f, 0.0f, 0.0f};
  if (i < a.P) {
    pm[0] = a.G.means[3 * (size_t)i];
    pm[1] = a.G.means[3 * (size_t)i + 1];
    pm[2] = a.G.means[3 * (size_t)i + 2];
  }
  // full groups of 64 rows start 16-byte aligned when the arrays do
  // (3 * 64 and 3 * M * 64 floats are multiples of 4): copy as float4 with
  // every load of the group in flight before the first LDS store (indices
  // clamped into the group; the padding takes the surplus stores)
  const bool vec = nb == PRE_THREADS && M <= 15 &&
                   (((uintptr_t)a.G.dc | (uintptr_t)(M > 0 ? a.G.shs : a.G.dc)) & 15) == 0;
  if (vec) {
    const int nq = 3 * M * PRE_THREADS / 4;
    const float4* d4 = (const float4*)gdc;
    const float4* h4 = M > 0 ? (const float4*)gsh : d4;
    const float4 dv = d4[min(lane, 3 * PRE_THREADS / 4 - 1)];
    float4 hv[SHC_QMAX];
#pragma unroll
    for (int u = 0; u < SHC_QMAX; ++u) hv[u] = h4[min(lane + u * PRE_THREADS, max(nq - 1, 0))];
    rows4[lane] = dv;
#pragma unroll
    for (int u = 0; u < SHC_QMAX; ++u) rows4[SHC_DC_F4 + lane + u * PRE_THREADS] = hv[u];
  } else {
    for (int k = lane; k < 3 * nb; k += PRE_THREADS) sdc[k] = gdc[k];
    for (int k = lane; k < 3 * M * nb; k += PRE_THREADS) ssh[k] = gsh[k];
  }
  __syncthreads();
  if (i >= a.P) return;
  // the coefficients are staged once for all views of the batch
  for (int v = 0; v < a.V; ++v) {
    const size_t ri = (size_t)v * a.P + i;
    if (a.dkey[ri] == 0xffffffffu) continue;
    const float* cp = a.views[v].campos;
    float rgb[3];
    int clamp;
    const float ox = pm[0] - cp[0], oy = pm[1] - cp[1], oz = pm[2] - cp[2];
    sh_color_rows(ox, oy, oz, sdc + 3 * lane, ssh + 3 * M * lane, a.cam.deg, rgb, clamp);
    ((float4*)(a.splat + 12 * ri))[2] = make_float4(rgb[0], rgb[1], rgb[2], u2f((uint32_t)clamp));
    if (JAC) {  // one view: the Jacobian for the backward
      const float len = sqrt_(ox * ox + oy * oy + oz * oz);
      float J[3][3];
      sh_jacobian(ox / len, oy / len, oz / len, ssh + 3 * M * lane, a.cam.deg, J);
      float4* jo = (float4*)(a.shjac + 12 * (size_t)i);
      jo[0] = make_float4(J[0][0], J[0][1], J[0][2], u2f((uint32_t)clamp));
      jo[1] = make_float4(J[1][0], J[1][1], J[1][2], 0.0f);
      jo[2] = make_float4(J[2][0], J[2][1], J[2][2], 0.0f);
    }
  }
}

// A grid of at most SHC_BLOCKS workgroups (each takes groups of 64 Gaussians
// in a grid-stride loop): the pass runs beside the depth sort, whose
// latency-bound passes need a workgroup slot and 55 KB of LDS on every CU; a
// full grid of this pass would hold every slot until it ends.
// JAC (frames of one view that will be differentiated): the Jacobians are
// stored too (a separate instance: their registers would lower the
// occupancy of the plain colour pass).
template <bool JAC>
__global__ __launch_bounds__(PRE_THREADS) void k_sh_color(PreArgs a) {
  extern __shared__ float4 rows4[];
  const int ngroups = (a.P + PRE_THREADS - 1) / PRE_THREADS;
  for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    if (grp != (int)blockIdx.x) __syncthreads();  // the previous group's rows are read
    sh_color_group<JAC>(a, rows4, grp);
  }
}

hipError_t launch_preprocess_fast(const PreArgs& a, const ViewSetup& vs, ViewRec* table,
                                  uint32_t* zero, int zero_n, hipStream_t s) {
  if (a.P <= 0) return launch_views_setup(vs, table, zero, zero_n, s);
  const unsigned blocks = (a.P + PRE_THREADS - 1) / PRE_THREADS;
  FGS_TIMED("preprocess_fast", s,
            (k_preprocess_fast<<<dim3(blocks, (unsigned)a.V), PRE_THREADS, 0, s>>>(a, vs, table,
                                                                                zero, zero_n)));
  return hipGetLastError();
}

hipError_t launch_sh_color(const PreArgs& a, hipStream_t s) {
  if (a.P <= 0 || a.G.colors != nullptr || a.G.dc == nullptr) return hipSuccess;
  const unsigned groups = (a.P + PRE_THREADS - 1) / PRE_THREADS;
  const unsigned blocks = groups < SHC_BLOCKS ? groups : SHC_BLOCKS;
  const size_t rows = (size_t)(4 * SHC_DC_F4 + PRE_THREADS * 3 * a.cam.max_coeffs) * sizeof(float);
  const size_t shmem = rows > SHC_SHMEM ? rows : SHC_SHMEM;
  if (a.shjac)
    FGS_TIMED("sh_color", s, (k_sh_color<true><<<blocks, PRE_THREADS, shmem, s>>>(a)));
  else
    FGS_TIMED("sh_color", s, (k_sh_color<false><<<blocks, PRE_THREADS, shmem, s>>>(a)));
  return hipGetLastError();
}

// Large rects: each rect is cut into up to LARGE_BANDS bands of whole tile
// rows (about LARGE_BAND_TILES tiles per band; large_bands), and the band
// list (entry = record * LARGE_BANDS + band, the bands of a record next to
// each other) is walked one band per wave, grid-stride over the device-side
// count. A full-screen rect is then 8 waves of about 7 rows instead of one
// wave of every row, which keeps the slowest wave short. Each lane keeps one
// tile column (its tile_col terms formed once); a band of width w <= 64 is
// walked 64 / w rows per step, a wider one 64 columns at a time, row by row.
// Counting stores the band's hits in lband and adds them to the Gaussian's
// pair count; emission places the k-th hit of band b at poff + (hits of the
// record's bands < b) + k (a running ballot prefix, no atomics). The order
// of one Gaussian's pairs within its slot range does not matter: the tile
// sort that follows is stable and each Gaussian appears at most once per
// tile.
template <bool EMIT>
__device__ __forceinline__ void large_step(bool hit, int tx, int ty, uint32_t base, uint32_t& run,
                                           int gx, uint32_t toff, uint32_t g, uint32_t cap,
                                           void* tkeys, uint32_t* recs, int t16) {
  const uint64_t m = __ballot(hit);
  if (EMIT) {
    const uint32_t pos = base + run + lanes_below(m);
    if (hit && pos < cap) store_pair(tkeys, recs, t16 != 0, pos, toff + (uint32_t)(ty * gx + tx), g);
  }
  run += (uint32_t)__popcll(m);
}

template <bool EMIT>
__global__ __launch_bounds__(256) void k_large_tiles(int P, int V, const ViewRec* __restrict__ views,
                                                     const float* __restrict__ geom,
                                                     const uint32_t* __restrict__ large_ids,
                                                     const uint32_t* __restrict__ counts,
                                                     uint32_t* pcount,
                                                     const uint32_t* __restrict__ poff,
                                                     const uint32_t* __restrict__ lrank,
                                                     uint32_t* lband,
                                                     uint32_t cap, void* tkeys, uint32_t* recs,
                                                     int t16, uint32_t rank_limit,
                                                     const uint32_t* __restrict__ gate) {
  if (gate && gate[0] == 0u) return;
  const uint32_t NT = counts[3];
  const int lane = lane_id();
  const uint32_t nwaves = gridDim.x * (blockDim.x / WAVE);
  for (uint32_t q = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; q < NT; q += nwaves) {
    const uint32_t id = large_ids[q];
    const uint32_t g = id / LARGE_BANDS;  // record index v * P + i
    const int band = (int)(id % LARGE_BANDS);
    // a band past the prefix leaves before its record's geometry is read
    uint32_t rank = 0u;
    if (EMIT) {
      rank = lrank[g];
      if (rank >= rank_limit) continue;  // wave-uniform
    }
    const GeomRow r = load_geom(geom, g);
    const int wr = r.r2 - r.r0;
    int rows, nbands;
    large_bands(wr, r.r3 - r.r1, rows, nbands);  // band < nbands for a listed band
    (void)nbands;
    const int y0 = r.r1 + band * rows;
    const int y1 = min(y0 + rows, r.r3);
    const int v = V > 1 ? (int)(g / (uint32_t)P) : 0;
    const int gx = views[v].grid_x, tw = views[v].tile_w, th = views[v].tile_h;
    const uint32_t toff = views[v].tile_off;
    uint32_t base = 0u;
    if (EMIT) {
      base = poff[rank];
      for (int k = 1; k <= band; ++k) base += lband[q - k];
    }
    uint32_t run = 0;
    if (wr <= WAVE) {
      const int k = WAVE / wr;  // rows per step
      const int rs = lane / wr;
      const int tx = r.r0 + (lane - rs * wr);
      const TileCol c = tile_col(tx, tw, r.x, r.conic, r.power);
      for (int row0 = y0; row0 < y1; row0 += k) {
        const int ty = row0 + rs;
        const TileRow rw = tile_row(ty, th, r.y, r.conic, r.power);
        const bool hit = rs < k && ty < y1 && tile_hit_terms(c, rw, r.conic);
        large_step<EMIT>(hit, tx, ty, base, run, gx, toff, g, cap, tkeys, recs, t16);
      }
    } else {
      for (int c0 = r.r0; c0 < r.r2; c0 += WAVE) {
        const int tx = c0 + lane;
        const TileCol c = tile_col(tx, tw, r.x, r.conic, r.power);
        for (int ty = y0; ty < y1; ++ty) {
          const TileRow rw = tile_row(ty, th, r.y, r.conic, r.power);
          const bool hit = tx < r.r2 && tile_hit_terms(c, rw, r.conic);
          large_step<EMIT>(hit, tx, ty, base, run, gx, toff, g, cap, tkeys, recs, t16);
        }
      }
    }
    if (!EMIT && lane == 0) {
      lband[q] = run;
      if (run) atomicAdd(&pcount[g], run);
    }
  }
}

hipError_t launch_large_tiles(bool emit, int P, int V, const ViewRec* views, const float* geom,
                              const uint32_t* large_ids, const uint32_t* counts, uint32_t* pcount,
                              const uint32_t* poff, const uint32_t* lrank, uint32_t* lband,
                              uint32_t cap, void* tkeys, uint32_t* recs, bool t16, hipStream_t s,
                              uint32_t rank_limit, const uint32_t* gate) {
  if (P <= 0) return hipSuccess;
  // 16384 waves, grid-stride; each exits past the device-side count (fewer
  // for the count, which runs beside the depth sort, measured slower: the
  // pair offsets wait for it)
  const int blocks = 4096;
  if (emit)
    FGS_TIMED(gate ? "large_tiles_emit_redo" : "large_tiles_emit", s,
              (k_large_tiles<true><<<blocks, 256, 0, s>>>(P, V, views, geom, large_ids, counts,
                                                          pcount, poff, lrank, lband, cap, tkeys,
                                                          recs, t16 ? 1 : 0, rank_limit, gate)));
  else
    FGS_TIMED("large_tiles_count", s,
              (k_large_tiles<false><<<blocks, 256, 0, s>>>(P, V, views, geom, large_ids, counts,
                                                           pcount, poff, lrank, lband, cap, tkeys,
                                                           recs, t16 ? 1 : 0, rank_limit, gate)));
  return hipGetLastError();
}

// Emission of the small rects in depth-rank order: wave w walks the records
// of ranks [64 w, 64 w + 64) (order = records by depth rank). Pair slots are
// assigned in that order, so a wave's pairs fill one contiguous stretch of
// the pair array (apart from the slots of large rects among them, which
// k_large_tiles fills) and its stores combine in L2. Ranks [0, limit) only
// (a prefix of the depth order: their pairs are the slots [0, poff[limit])).
__global__ __launch_bounds__(256) void k_emit_small(int P, int NR, const ViewRec* __restrict__ views,
                                                    const float* __restrict__ geom, int split,
                                                    const uint64_t* __restrict__ order,
                                                    const uint32_t* __restrict__ poff, uint32_t cap,
                                                    void* __restrict__ tkeys,
                                                    uint32_t* __restrict__ recs, int t16,
                                                    int limit, uint32_t* __restrict__ counts,
                                                    int mode) {
  __shared__ WaveTiles lds[4];
  __shared__ WaveEmit lde[4];
  if (mode == EMIT_REDO) {
    if (counts[COUNT_REDO] == 0u) return;  // the prefix lists sufficed
  } else if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      counts[COUNT_PREFIX_N] = mode == EMIT_PREFIX ? poff[limit] : 0u;
      counts[COUNT_REDO] = 0u;
      counts[COUNT_REDO_N] = 0u;
    }
    if (mode == EMIT_PREFIX)
      for (int i = threadIdx.x; i < SORT_SYNC_WORDS; i += 256) counts[REDO_SYNC_OFF + i] = 0u;
  }
  const int rank = blockIdx.x * 256 + threadIdx.x;
  const int lane = lane_id();
  WaveTiles& L = lds[threadIdx.x / WAVE];
  WaveEmit& E = lde[threadIdx.x / WAVE];
  GeomRow r{};
  bool small = false;
  uint32_t g = 0u, pos = 0u, toff = 0u;
  int gx = 1;
  if (rank < limit) {
    g = (uint32_t)order[rank];
    r = load_geom(geom, g);
    const int area = (r.r2 - r.r0) * (r.r3 - r.r1);
    small = area > 0 && area < split;
    if (small) {
      const int v = NR > P ? (int)(g / (uint32_t)P) : 0;
      pos = poff[rank];
      gx = views[v].grid_x;
      toff = views[v].tile_off;
    }
  }
  uint32_t S;
  stage_small(L, lane, small, r.x, r.y, r.conic, r.power, r.r0, r.r1, r.r2, r.r3, S);
  E.rec[lane] = g;
  E.pos[lane] = pos;
  E.toff[lane] = toff;
  E.gx[lane] = gx;
  __syncthreads();
  // one tile shape for every view of a batch
  walk_small<true>(L, &E, S, cap, tkeys, recs, t16 != 0, views[0].tile_w, views[0].tile_h);
}

hipError_t launch_emit_small(int P, int V, const ViewRec* views, const float* geom, int split,
                             const uint64_t* order, const uint32_t* poff, uint32_t cap,
                             void* tkeys, uint32_t* recs, bool t16, hipStream_t s,
                             int64_t limit, uint32_t* counts, EmitMode mode) {
  const int64_t NR = (int64_t)P * V;
  if (limit < 0 || limit > NR || (mode == EMIT_PREFIX && limit >= NR))
    return hipErrorInvalidValue;
  if (mode != EMIT_PREFIX) limit = NR;
  if (NR > 0)
    FGS_TIMED(mode == EMIT_REDO ? "emit_small_redo" : "emit_small", s,
              (k_emit_small<<<(unsigned)((limit + 255) / 256 > 0 ? (limit + 255) / 256 : 1), 256,
                              0, s>>>(P, (int)NR, views, geom, split, order, poff, cap, tkeys,
                                      recs, t16 ? 1 : 0, (int)limit, counts, (int)mode)));
  return hipGetLastError();
}

// View table from kernel arguments (see ViewSetup).
// Thread t (< 20) of the copy of view v into the table.
__device__ __forceinline__ void write_view(const ViewSetup& src, ViewRec* out, int v, int t) {
  const ViewSrc& s = src.v[v];
  ViewRec& r = out[v];
  if (t < 16) {
    r.view[t] = s.view[t];
    r.proj[t] = s.proj[t];
  } else if (t < 19) {
    r.campos[t - 16] = s.campos[t - 16];
    r.bg[t - 16] = s.bg ? s.bg[t - 16] : 0.0f;
  } else if (t == 19) {
    r.campos[3] = 0.0f;
    r.bg[3] = 0.0f;
    r.tanfovx = s.tanfovx;
    r.tanfovy = s.tanfovy;
    r.focal_x = s.focal_x;
    r.focal_y = s.focal_y;
    r.W = s.W;
    r.H = s.H;
    r.grid_x = grid_dim(s.W, s.tile_w);
    r.grid_y = grid_dim(s.H, s.tile_h);
    r.tile_off = s.tile_off;
    r.pix_off = s.pix_off;
    r.tile_w = s.tile_w;
    r.tile_h = s.tile_h;
  }
}

__global__ void k_views_setup(ViewSetup src, ViewRec* out, uint32_t* zero, int64_t zero_n) {
  const int v = blockIdx.x;
  const int t = threadIdx.x;
  for (int64_t i = (int64_t)v * blockDim.x + t; i < zero_n; i += (int64_t)gridDim.x * blockDim.x)
    zero[i] = 0u;
  if (v < src.n) write_view(src, out, v, t);
}

hipError_t launch_views_setup(const ViewSetup& src, ViewRec* out, uint32_t* zero, int64_t zero_n,
                              hipStream_t s) {
  // enough workgroups that each clears at most ~8 words per thread
  int64_t blocks = (zero_n + 2047) / 2048;
  if (blocks > 1024) blocks = 1024;
  if (blocks < src.n) blocks = src.n;
  if (blocks > 0) k_views_setup<<<(unsigned)blocks, 256, 0, s>>>(src, out, zero, zero_n);
  return hipGetLastError();
}

}  // namespace fgs
