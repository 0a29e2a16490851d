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

}  // namespace fgs
