// Preprocess and pair emission kernels.
//
// Per Gaussian (reference: preprocessSmallCUDA, FGS/csrc/cuda_rasterizer/
// preprocess.cu:414-746): cull, project, 3D/2D covariance, conic, radius,
// opacity-aware half extents, tile rect, SH colour; then the exact
// tile/ellipse test (preprocess.cu:600-607) decides which tiles get a pair.
//
// Records written per Gaussian (structure of arrays, 16-byte aligned rows):
//   geom  8 floats: x, y, conic a, b, c, power, rect min, rect max (packed u16x2)
//   splat 12 floats: x, y, ca, cb, cc, log2 op, r, g, b, depth, clamp bits, 0
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
#include "fgs_layout.h"

namespace fgs {

__device__ __forceinline__ void store_records(int i, const Proj<float>& o, float* geom,
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
  s4[1] = make_float4(cc, w, o.rgb[0], o.rgb[1]);
  s4[2] = make_float4(o.rgb[2], o.depth, u2f((uint32_t)o.clamp), 0.0f);
}

__device__ __forceinline__ void store_invalid(int i, float* geom, float* splat) {
  float4* g4 = (float4*)(geom + 8 * (size_t)i);
  g4[0] = make_float4(0.f, 0.f, 0.f, 0.f);
  g4[1] = make_float4(0.f, 0.f, 0.f, 0.f);
  float4* s4 = (float4*)(splat + 12 * (size_t)i);
  s4[2] = make_float4(0.f, 0.f, 0.f, 0.f);
}

struct GeomRow {
  float x, y, conic[3], power;
  int r0, r1, r2, r3;
};

__device__ __forceinline__ GeomRow load_geom(const float* geom, int i) {
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
      n += tile_hit(tx, ty, o.x, o.y, o.conic, o.power) ? 1u : 0u;
  a.pcount[i] = n;
  a.dkey[i] = depth_bits(o.depth);
  store_records(i, o, a.geom, a.splat);
}

hipError_t launch_preprocess_plain(const PreArgs& a, hipStream_t s) {
  if (a.P > 0) k_preprocess_plain<<<(a.P + 255) / 256, 256, 0, s>>>(a);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_duplicate_plain(int P, const float* __restrict__ geom,
                                                         const uint32_t* __restrict__ dkey,
                                                         const uint32_t* __restrict__ pcount,
                                                         const uint32_t* __restrict__ poff,
                                                         int grid_x, uint64_t* __restrict__ keys,
                                                         uint32_t* __restrict__ vals) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P || pcount[i] == 0) return;
  const GeomRow g = load_geom(geom, i);
  uint32_t k = poff[i];
  const uint64_t d = dkey[i];
  for (int ty = g.r1; ty < g.r3; ++ty)
    for (int tx = g.r0; tx < g.r2; ++tx)
      if (tile_hit(tx, ty, g.x, g.y, g.conic, g.power)) {
        keys[k] = ((uint64_t)(uint32_t)(ty * grid_x + tx) << 32) | d;
        vals[k] = (uint32_t)i;
        ++k;
      }
}

hipError_t launch_duplicate_plain(int P, const float* geom, const uint32_t* dkey,
                                  const uint32_t* pcount, const uint32_t* poff, int grid_x,
                                  uint64_t* keys, uint32_t* vals, hipStream_t s) {
  if (P > 0)
    k_duplicate_plain<<<(P + 255) / 256, 256, 0, s>>>(P, geom, dkey, pcount, poff, grid_x, keys,
                                                      vals);
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

// Walk the concatenated small rects of one wave 64 candidate tiles per step
// (lane-packed); count (EMIT=false) or emit (EMIT=true) the pairs that pass
// the exact tile test. Within a Gaussian's block of the pair array the order
// is free: its pairs are in distinct tiles and the tile sort is stable.
template <bool EMIT>
__device__ __forceinline__ void walk_small(WaveTiles& L, uint32_t S, int gbase, int gx,
                                           const uint32_t* __restrict__ poff,
                                           uint64_t* __restrict__ packed) {
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
      if (tile_hit(tx, ty, L.x[j], L.y[j], conic, L.pw[j])) {
        const uint32_t k = atomicAdd(&L.cnt[j], 1u);
        if (EMIT) {
          const uint32_t g = (uint32_t)(gbase + j);
          packed[poff[g] + k] = ((uint64_t)(uint32_t)(ty * gx + tx) << 32) | g;
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

__global__ __launch_bounds__(256) void k_preprocess_fast(PreArgs a) {
  __shared__ WaveTiles lds[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int lane = lane_id();
  WaveTiles& L = lds[threadIdx.x / WAVE];
  const int gbase = blockIdx.x * 256 + (threadIdx.x / WAVE) * WAVE;
  Proj<float> o;
  o.valid = 0;
  if (i < a.P) {
    preprocess_gauss(i, a.G, a.cam, o);
    a.radii[i] = o.valid ? o.radius : 0;
    const uint32_t dk = o.valid ? depth_bits(o.depth) : 0xffffffffu;
    a.dkey[i] = dk;
    a.dsort_k[i] = dk;
    a.dsort_v[i] = (uint32_t)i;
    if (o.valid)
      store_records(i, o, a.geom, a.splat);
    else
      store_invalid(i, a.geom, a.splat);
  }
  const int area = o.valid ? (o.rect[2] - o.rect[0]) * (o.rect[3] - o.rect[1]) : 0;
  const bool small = o.valid && area < a.split_tiles;
  if (i < a.P) {
    a.lchunk[i] = (o.valid && !small) ? (uint32_t)((area + a.split_tiles - 1) / a.split_tiles) : 0u;
  }
  uint32_t S;
  stage_small(L, lane, small, o.x, o.y, o.conic, o.power, o.rect[0], o.rect[1], o.rect[2],
              o.rect[3], S);
  __syncthreads();
  walk_small<false>(L, S, gbase, a.cam.grid_x, nullptr, nullptr);
  __syncthreads();
  if (i < a.P) a.pcount[i] = L.cnt[lane];
}

hipError_t launch_preprocess_fast(const PreArgs& a, hipStream_t s) {
  if (a.P > 0) k_preprocess_fast<<<(a.P + 255) / 256, 256, 0, s>>>(a);
  return hipGetLastError();
}

// Large rects: tasks of `split` tiles each (task t belongs to the Gaussian
// whose [loff, loff + lchunk) holds t); waves take tasks in a grid-stride
// loop over the device-side task count and walk 64 tiles per step.
template <bool EMIT>
__global__ __launch_bounds__(256) void k_large_tiles(int P, const float* __restrict__ geom,
                                                     const uint32_t* __restrict__ loff,
                                                     const uint32_t* __restrict__ counts,
                                                     int split, int gx, uint32_t* pcount,
                                                     const uint32_t* __restrict__ poff,
                                                     uint32_t* lcur, uint64_t* packed) {
  const uint32_t L = counts[3];
  const int lane = lane_id();
  const uint32_t nwaves = gridDim.x * (blockDim.x / WAVE);
  for (uint32_t t = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; t < L; t += nwaves) {
    // last Gaussian with loff <= t
    int lo = 0, hi = P;  // loff[lo] <= t < loff[hi] (loff[P] = L)
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (loff[mid] <= t) lo = mid; else hi = mid;
    }
    const int g = lo;
    const GeomRow r = load_geom(geom, g);
    const int w = r.r2 - r.r0;
    const int area = w * (r.r3 - r.r1);
    const int c = (int)(t - loff[g]);
    const int start = c * split;
    const int end = min(area, start + split);
    int l = start + lane;
    int ty = l / w;
    int tx = l - ty * w;
    const int sy = WAVE / w, sx = WAVE - sy * w;
    uint32_t mine = 0;
    for (int k = start; k < end; k += WAVE) {
      const bool in = l < end;
      const bool hit = in && tile_hit(r.r0 + tx, r.r1 + ty, r.x, r.y, r.conic, r.power);
      const uint64_t m = __ballot(hit);
      if (!EMIT) {
        mine += (uint32_t)__popcll(m);
      } else if (m) {
        uint32_t slot0 = 0;
        if (lane == 0) slot0 = atomicAdd(&lcur[g], (uint32_t)__popcll(m));
        slot0 = __shfl(slot0, 0, WAVE);
        if (hit)
          packed[poff[g] + slot0 + lanes_below(m)] =
              ((uint64_t)(uint32_t)((r.r1 + ty) * gx + r.r0 + tx) << 32) | (uint32_t)g;
      }
      l += WAVE;
      tx += sx;
      ty += sy;
      if (tx >= w) {
        tx -= w;
        ty += 1;
      }
    }
    if (!EMIT && lane == 0 && mine) atomicAdd(&pcount[g], mine);
  }
}

hipError_t launch_large_tiles(bool emit, int P, const float* geom, const uint32_t* loff,
                              const uint32_t* counts, int split, int gx, uint32_t* pcount,
                              const uint32_t* poff, uint32_t* lcur, uint64_t* packed,
                              hipStream_t s) {
  if (P <= 0) return hipSuccess;
  const int blocks = 1024;  // 4096 persistent waves; each exits when the task count is reached
  if (emit)
    k_large_tiles<true><<<blocks, 256, 0, s>>>(P, geom, loff, counts, split, gx, pcount, poff,
                                               lcur, packed);
  else
    k_large_tiles<false><<<blocks, 256, 0, s>>>(P, geom, loff, counts, split, gx, pcount, poff,
                                                lcur, packed);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_emit_small(int P, const float* __restrict__ geom,
                                                    int split, int gx,
                                                    const uint32_t* __restrict__ poff,
                                                    uint64_t* __restrict__ packed) {
  __shared__ WaveTiles lds[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int lane = lane_id();
  WaveTiles& L = lds[threadIdx.x / WAVE];
  const int gbase = blockIdx.x * 256 + (threadIdx.x / WAVE) * WAVE;
  GeomRow r{};
  bool small = false;
  if (i < P) {
    r = load_geom(geom, i);
    const int area = (r.r2 - r.r0) * (r.r3 - r.r1);
    small = area > 0 && area < split;
  }
  uint32_t S;
  stage_small(L, lane, small, r.x, r.y, r.conic, r.power, r.r0, r.r1, r.r2, r.r3, S);
  __syncthreads();
  walk_small<true>(L, S, gbase, gx, poff, packed);
}

hipError_t launch_emit_small(int P, const float* geom, int split, int gx, const uint32_t* poff,
                             uint64_t* packed, hipStream_t s) {
  if (P > 0) k_emit_small<<<(P + 255) / 256, 256, 0, s>>>(P, geom, split, gx, poff, packed);
  return hipGetLastError();
}

// Pair counts gathered in depth order (rank r -> Gaussian order[r]).
__global__ void k_gather_u32(int P, const uint32_t* __restrict__ order,
                             const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r < P) dst[r] = src[order[r]];
}
// Offsets scattered back to Gaussian order.
__global__ void k_scatter_u32(int P, const uint32_t* __restrict__ order,
                              const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r < P) dst[order[r]] = src[r];
}
hipError_t launch_gather_scatter(bool gather, int P, const uint32_t* order, const uint32_t* src,
                                 uint32_t* dst, hipStream_t s) {
  if (P <= 0) return hipSuccess;
  if (gather)
    k_gather_u32<<<(P + 255) / 256, 256, 0, s>>>(P, order, src, dst);
  else
    k_scatter_u32<<<(P + 255) / 256, 256, 0, s>>>(P, order, src, dst);
  return hipGetLastError();
}

}  // namespace fgs
