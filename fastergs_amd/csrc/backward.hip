// Backward kernels: rasterizer adjoint, per-Gaussian preprocess adjoint and
// densification statistics.
//
// Per-pixel reverse step: FGS/csrc/cuda_rasterizer/backward.cu:549-632.
// Per-Gaussian reduction then atomics: backward.cu:634-693 (the reference
// reduces 11 values over a 32-lane warp and adds them with float atomics, in
// an order that varies run to run; here a wave64 reduction whose float part
// has a fixed tree and whose cross-quad part and atomics are fp64).
// Preprocess adjoint: backward.cu:250-539. Statistics: backward.cu:892-918.
//
// Raw per-Gaussian sums land in two arrays of rows (fgs_math.h RawField,
// raw_lin_index): the three conic sums in double, the eight others in
// float; the per-Gaussian passes round them to float, apply the constant
// factors and produce the reference's gradient tensors.
//
// Summation: per splat, every lane accumulates its pixels' terms in float in
// a fixed order, and the wave adds its 64 lane values in float along a fixed
// tree (wave_flush); the wave totals are added across waves with atomics,
// in fp64 for the conic sums and in float for the others. The conic -> 2D
// covariance adjoint (inverse of a possibly elongated 2x2 matrix) amplifies
// the rounding of the conic sums by up to its squared condition number, so
// those are added order-independently: fp64 reassociation (~1e-16 relative)
// is far below their float rounding, so they -- and the geometric gradients
// they feed -- come out the same from run to run and equal to the CPU
// oracle's, which forms the same float wave totals and adds them in double.
// The other sums enter their gradients linearly; float atomics move those
// by ~1e-7 relative.
#include <hip/hip_runtime.h>

#include "fgs_device.h"
#include "fgs_kernels.h"
#include "fgs_prof.h"
#include "fgs_layout.h"

namespace fgs {

// 1 / (1 - alpha) of the reverse walk, correctly rounded: v_rcp_f32 (1 ulp)
// refined by one Newton step with fused multiply-adds. For alpha <= 0.99 the
// argument lies in [0.01, 1]; over every float in [2^-7, 1] this equals the
// IEEE quotient 1.0f / x (checked exhaustively on the device by
// fgs_rcp_selftest), so the walk recovers the same T as the oracle's
// division at 3 instructions instead of the ~10 of a general IEEE division.
struct ExactRcp {
  __device__ float operator()(float x) const {
    const float y = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, y, 1.0f);
    return __builtin_fmaf(e, y, y);
  }
};

// Reduce NF (<= 16) raw fields over the wave and add field f to the raw sums
// of record id (fp64 atomics for the conic fields, float for the others). Transposing butterfly: at
// each step a lane keeps one field of every pair and hands the other to a
// partner lane that differs in the step's lane bit (xor 1, xor 2 by
// quad_perm; bit 2 and bit 3 by row rotations of 4 and 8, whose partners hold
// the same fields), so the number of fields per lane halves while the lanes
// they cover double; after four steps lane l holds field l % 16 summed over
// its row, and two shuffles add the four rows. About 40 instructions for 9
// fields instead of six DPP steps per field. The float tree it evaluates is
// fixed (float addition commutes, so which lane of a pair adds is
// immaterial): with a_l the lane values, q_j = (a_4j + a_4j+1) + (a_4j+2 +
// a_4j+3) the lane-quad sums, row r's sum of field f is (q_g + q_g-1) +
// (q_g-2 + q_g-3) over the quads 4r + (. mod 4) with g = (f / 4) mod 4
// (row_ror:n reads lane (i - n) mod 16), and the total is (row0 + row1) +
// (row2 + row3).
template <int NF>
__device__ __forceinline__ void wave_flush(const float* acc, float* raw_lin, double* raw_con,
                                           uint32_t id) {
  static_assert(NF <= 16, "at most 16 fields");
  const int l = lane_id();
  const bool b0 = l & 1, b1 = (l >> 1) & 1, b2 = (l >> 2) & 1, b3 = (l >> 3) & 1;
  float s1[8], s2[4], s3[2];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const float a = 2 * m < NF ? acc[2 * m] : 0.0f;
    const float b = 2 * m + 1 < NF ? acc[2 * m + 1] : 0.0f;
    s1[m] = 2 * m < NF ? (b0 ? b : a) + dpp_f<0xb1>(b0 ? a : b) : 0.0f;  // quad_perm [1,0,3,2]
  }
#pragma unroll
  for (int n = 0; n < 4; ++n)
    s2[n] = 4 * n < NF ? (b1 ? s1[2 * n + 1] : s1[2 * n]) +
                             dpp_f<0x4e>(b1 ? s1[2 * n] : s1[2 * n + 1])  // quad_perm [2,3,0,1]
                       : 0.0f;
#pragma unroll
  for (int p = 0; p < 2; ++p)
    s3[p] = 8 * p < NF ? (b2 ? s2[2 * p + 1] : s2[2 * p]) +
                             dpp_f<0x124>(b2 ? s2[2 * p] : s2[2 * p + 1])  // row_ror:4
                       : 0.0f;
  float v = (b3 ? s3[1] : s3[0]) + dpp_f<0x128>(b3 ? s3[0] : s3[1]);  // row_ror:8
  v += __shfl_xor(v, 16, WAVE);
  v += __shfl_xor(v, 32, WAVE);
  if (l >= RAW_KXX && l <= RAW_KYY)
    unsafeAtomicAdd(raw_con + (size_t)RAW_CON_STRIDE * id + (l - RAW_KXX), (double)v);
  else if (l < NF)
    atomic_add_f(raw_lin + (size_t)RAW_LIN_STRIDE * id + raw_lin_index(l), v);
}

// Exhaustive check of ExactRcp against the IEEE quotient over [2^-7, 1].
__global__ __launch_bounds__(256) void k_rcp_selftest(uint32_t lo, uint32_t n,
                                                      unsigned long long* bad) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float x = __uint_as_float(lo + i);
  const bool diff = __float_as_uint(ExactRcp()(x)) != __float_as_uint(1.0f / x);
  const uint64_t m = __ballot(diff);
  if (m && lane_id() == __builtin_ctzll(m)) atomicAdd(bad, (unsigned long long)__popcll(m));
}

hipError_t launch_rcp_selftest(unsigned long long* bad, hipStream_t s) {
  const uint32_t lo = 0x3C000000u, hi = 0x3F800000u;  // 2^-7 .. 1.0
  const uint32_t n = hi - lo + 1;
  hipError_t e = fill_u32((uint32_t*)bad, 0u, 2, s);
  if (e != hipSuccess) return e;
  k_rcp_selftest<<<(n + 255) / 256, 256, 0, s>>>(lo, n, bad);
  return hipGetLastError();
}

// Raw sums of one record in RawField order (float; the conic ones rounded
// from double).
__device__ __forceinline__ void load_raw(const float* raw_lin, const double* raw_con, size_t i,
                                         float* raw) {
  const float4* l4 = (const float4*)(raw_lin + RAW_LIN_STRIDE * i);
  const double2* c2 = (const double2*)(raw_con + RAW_CON_STRIDE * i);
  const float4 a = l4[0], b = l4[1];
  const double2 k01 = c2[0];
  const double k2 = raw_con[RAW_CON_STRIDE * i + 2];
  raw[RAW_COL0] = a.x;
  raw[RAW_COL1] = a.y;
  raw[RAW_COL2] = a.z;
  raw[RAW_MX] = a.w;
  raw[RAW_MY] = b.x;
  raw[RAW_KXX] = (float)k01.x;
  raw[RAW_KXY] = (float)k01.y;
  raw[RAW_KYY] = (float)k2;
  raw[RAW_POW] = b.y;
  raw[RAW_AX] = b.z;
  raw[RAW_AY] = b.w;
  raw[RAW_STRIDE - 1] = 0.0f;
}

// ---------------------------------------------------------- PLAIN backward
// One workgroup per tile, one pixel per thread, reverse over the tile list
// from the tile's furthest contributor; each pixel starts at T_final with an
// empty residual and recovers T by division (no checkpoints).
template <bool ABS>
__global__ __launch_bounds__(256) void k_backward_plain(BwdArgs a) {
  constexpr int NF = ABS ? RAW_N : RAW_POW + 1;
  __shared__ float4 sA[256];
  __shared__ float4 sB[256];
  __shared__ float sC[256];
  __shared__ uint32_t sId[256];
  const int tile = blockIdx.x;
  const int tx = tile % a.grid_x, ty = tile / a.grid_x;
  const int t = threadIdx.x;
  const int tw = a.tile_w;
  const int px = tx * tw + (t & (tw - 1)), py = ty * tile_h_of(tw) + (t >> tile_w_log2(tw));
  const bool inside = px < a.W && py < a.H;
  const uint32_t start = a.ranges[2 * tile];
  const int mc = (int)a.maxc[tile];
  const size_t HW = (size_t)a.W * a.H;
  const size_t g = inside ? (size_t)py * a.W + px : 0;
  const int cnt = inside ? a.contrib[g] : 0;
  float T = inside ? a.tfinal[g] : 0.0f;
  const float Tf = T;
  const float dp0 = inside ? a.dimg[g] : 0.0f;
  const float dp1 = inside ? a.dimg[HW + g] : 0.0f;
  const float dp2 = inside ? a.dimg[2 * HW + g] : 0.0f;
  const float tb = Tf * (a.bg[0] * dp0 + a.bg[1] * dp1 + a.bg[2] * dp2);
  const float fx = (float)px, fy = (float)py;
  float Q = tb;  // empty residual: background only
  for (int top = mc; top > 0; top -= 256) {
    const int lo = top > 256 ? top - 256 : 0;
    const int n = top - lo;
    __syncthreads();
    if (t < n) {
      const uint32_t id = a.vals[start + lo + t];
      const float4* s = (const float4*)(a.splat + 12 * (size_t)id);
      const float4 s1 = s[1], s2 = s[2];
      sA[t] = s[0];
      sB[t] = make_float4(s1.x, s1.y, s2.x, s2.y);  // cc, log2 op, r, g
      sC[t] = s2.z;
      sId[t] = id;
    }
    __syncthreads();
    for (int j = n - 1; j >= 0; --j) {
      const bool act = lo + j < cnt;
      if (__ballot(act) == 0) continue;
      float acc[RAW_N];
#pragma unroll
      for (int f = 0; f < RAW_N; ++f) acc[f] = 0.0f;
      if (act) {
        const float4 A = sA[j], Bv = sB[j];
        pix_bwd<float, ExactRcp, ABS>(A.x, A.y, A.z, A.w, Bv.x, Bv.y, Bv.z, Bv.w, sC[j], fx, fy,
                                      dp0, dp1, dp2, T, Q, acc);
      }
      wave_flush<NF>(acc, a.raw_lin, a.raw_con, sId[j]);
    }
  }
}

hipError_t launch_backward_plain(const BwdArgs& a, hipStream_t s) {
  if (a.T <= 0) return hipSuccess;
  if (a.abs_grad)
    FGS_TIMED("backward_plain", s, (k_backward_plain<true><<<a.T, 256, 0, s>>>(a)));
  else
    FGS_TIMED("backward_plain", s, (k_backward_plain<false><<<a.T, 256, 0, s>>>(a)));
  return hipGetLastError();
}

// ----------------------------------------------------------- FAST backward
// One wave64 per needed bucket (backward.cu:695-878): the buckets of each
// tile below its furthest contributor form a compact task list (prefix over
// tiles, `actoff`), taken by persistent waves in a grid-stride loop. A wave
// covers all 256 pixels of the tile (a 2x2 quad per lane, the quads in
// row-major order over the tile's tw/2 x th/2 grid), restarts each pixel from
// the checkpoint stored after this bucket (T, C) and walks the bucket's
// splats in reverse; the splat records are wave-uniform scalar loads, the
// next one prefetched while the current one is processed. Per splat the 4
// pixel contributions are summed in registers, then over the wave with DPP,
// and one atomic instruction adds the 9 (or 11) raw sums.
template <bool ABS>
__global__ __launch_bounds__(256, 6) void k_backward_fast(BwdArgs a) {
  constexpr int NF = ABS ? RAW_N : RAW_POW + 1;
  const uint32_t A = a.actoff[a.T];
  const int lane = lane_id();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + threadIdx.x / WAVE);
  const uint32_t nwaves = gridDim.x * 4;
  for (uint32_t u = wave; u < A; u += nwaves) {
    int lo = 0, hi = a.T;  // actoff[lo] <= u < actoff[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (a.actoff[mid] <= u) lo = mid; else hi = mid;
    }
    const int tile = lo;
    const uint32_t lb = u - a.actoff[tile];
    const uint32_t start = a.ranges[2 * tile], end = a.ranges[2 * tile + 1];
    const uint32_t s = start + lb * (uint32_t)a.bucket;
    const int n = (int)min((uint32_t)a.bucket, end - s);
    const size_t bq = (size_t)(a.boff[tile] + lb) * TILE_PIX;
    // the tile's view: size, tile grid, background and image offsets
    const ViewRec& vr = a.views[a.V > 1 ? view_of_tile(a.views, a.V, (uint32_t)tile) : 0];
    const int W = vr.W, H = vr.H, gx = vr.grid_x, tw = vr.tile_w, th = vr.tile_h;
    const int qx = 2 * (lane & (tw / 2 - 1)), qy = 2 * (lane >> (tile_w_log2(tw) - 1));
    const size_t HW = (size_t)W * H;
    const size_t poff = vr.pix_off;
    const float* dimg = a.dimg + 3 * poff;
    const int lt = tile - (int)vr.tile_off;
    const int tx = lt % gx, ty = lt / gx;
    // Every pixel runs the same branch-free step: an inactive pixel is
    // placed 1e7 pixels away (power far below the cut-off, alpha 0, every
    // contribution 0) with zero dL/dpixel.
    float T[4], Q[4], d0[4], d1[4], d2[4], fx[4], fy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int lx = qx + (k & 1), ly = qy + (k >> 1);
      const int px = tx * tw + lx, py = ty * th + ly;
      fy[k] = (float)py;
      const bool inside = px < W && py < H;
      const size_t gl = inside ? (size_t)py * W + px : 0;  // pixel in the view
      const size_t g = poff + gl;                            // pixel in the batch
      const bool act = inside && lb * (uint32_t)a.bucket < (uint32_t)a.contrib[g];
      fx[k] = act ? (float)px : 1.0e7f;
      T[k] = 1.0f; Q[k] = 0.0f; d0[k] = d1[k] = d2[k] = 0.0f;
      if (act) {
        const float4 ck = ((const float4*)a.acc)[bq + ly * tw + lx];
        const float4 pc = ((const float4*)a.pixc)[g];
        T[k] = ck.w;
        d0[k] = dimg[gl];
        d1[k] = dimg[HW + gl];
        d2[k] = dimg[2 * HW + gl];
        const float tb = a.tfinal[g] * (vr.bg[0] * d0[k] + vr.bg[1] * d1[k] + vr.bg[2] * d2[k]);
        Q[k] = residual_dot(pc.x - ck.x, pc.y - ck.y, pc.z - ck.z, d0[k], d1[k], d2[k], tb);
      }
    }
    uint32_t id = a.vals[s + n - 1];
    // record: (x, y, ca, cb) (cc, log2 op, depth, -) (r, g, b, clamp); the
    // walk uses A = first, B = (cc, log2 op, r, g), C = b
    const float4* rp = (const float4*)(a.splat + 12 * (size_t)id);
    float4 A0 = rp[0], B0;
    {
      const float4 b1 = rp[1], c1 = rp[2];
      B0 = make_float4(b1.x, b1.y, c1.x, c1.y);
    }
    float C0 = rp[2].z;
    for (int k = n - 1; k >= 0; --k) {
      // prefetch the previous splat of the bucket
      const uint32_t idn = k > 0 ? a.vals[s + k - 1] : id;
      const float4* rn = (const float4*)(a.splat + 12 * (size_t)idn);
      const float4 A1 = rn[0], n1 = rn[1], n2 = rn[2];
      const float4 B1 = make_float4(n1.x, n1.y, n2.x, n2.y);
      const float C1 = rn[2].z;
      // -0 start values: the first additions fold away (x + -0 = x)
      float acc[RAW_N];
#pragma unroll
      for (int f = 0; f < RAW_N; ++f) acc[f] = -0.0f;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        pix_bwd<float, ExactRcp, ABS>(A0.x, A0.y, A0.z, A0.w, B0.x, B0.y, B0.z, B0.w, C0, fx[q],
                                     fy[q], d0[q], d1[q], d2[q], T[q], Q[q], acc);
      wave_flush<NF>(acc, a.raw_lin, a.raw_con, id);
      id = idn;
      A0 = A1;
      B0 = B1;
      C0 = C1;
    }
  }
}

hipError_t launch_backward_fast(const BwdArgs& a, hipStream_t s) {
  if (a.T <= 0 || a.NB <= 0) return hipSuccess;
  const int64_t want = (a.NB + 3) / 4;
  const int blocks = (int)(want < 4096 ? want : 4096);
  if (a.abs_grad)
    FGS_TIMED("backward_fast", s, (k_backward_fast<true><<<blocks, 256, 0, s>>>(a)));
  else
    FGS_TIMED("backward_fast", s, (k_backward_fast<false><<<blocks, 256, 0, s>>>(a)));
  return hipGetLastError();
}

// --------------------------------------------------- per-Gaussian adjoint
// 128 Gaussians per workgroup; their SH rest rows are staged in LDS with
// coalesced loads, the adjoint overwrites each row in place with its
// gradient, and the rows are stored back coalesced.
constexpr int GBWD_THREADS = 128;
__global__ __launch_bounds__(GBWD_THREADS) void k_gauss_bwd(GaussBwdArgs a) {
  extern __shared__ float rows[];  // GBWD_THREADS x 3M
  const int g0 = blockIdx.x * GBWD_THREADS;
  const int i = g0 + threadIdx.x;
  const int M = a.cam.max_coeffs;
  const bool sh = a.dsh != nullptr && M > 0;
  const int nb = min(GBWD_THREADS, a.P - g0);
  if (sh)
    for (int k = threadIdx.x; k < 3 * M * nb; k += GBWD_THREADS)
      rows[k] = a.G.shs[(size_t)3 * M * g0 + k];
  __syncthreads();
  if (i < a.P) {
    float raw[RAW_STRIDE];
    load_raw(a.raw_lin, a.raw_con, (size_t)i, raw);
    Grad2D<float> g2;
    finalize_raw(raw, a.G.opac[i], a.cam.W, a.cam.H, g2);
    if (raw[RAW_POW] == 0.0f) g2.opacity = 0.0f;
    ((float4*)a.dmean2D)[i] = make_float4(g2.mean2d[0], g2.mean2d[1], g2.mean2d[2], g2.mean2d[3]);
    if (a.dconic)
      ((float4*)a.dconic)[i] = make_float4(g2.conic[0], g2.conic[1], g2.conic[2], g2.conic[3]);
    a.dopacity[i] = g2.opacity;
    if (a.dcolor) {
      a.dcolor[3 * (size_t)i] = g2.color[0];
      a.dcolor[3 * (size_t)i + 1] = g2.color[1];
      a.dcolor[3 * (size_t)i + 2] = g2.color[2];
    }
    float* row = rows + 3 * M * threadIdx.x;
    float ddc[3] = {0.0f, 0.0f, 0.0f};
    float tcov[6];  // dcov3D not requested
    GaussGrad<float> out;
    out.dmean3d = a.dmean3D + 3 * (size_t)i;
    out.dcov3d = a.dcov3D ? a.dcov3D + 6 * (size_t)i : tcov;
    out.dscale = a.dscale + 3 * (size_t)i;
    out.drot = a.drot + 4 * (size_t)i;
    out.ddc = ddc;
    out.dsh = sh ? row : nullptr;
    int written = 0;  // leading SH entries holding gradients
    if (a.radii[i] > 0) {
      const int clamp = (int)f2u(a.splat[12 * (size_t)i + 11]);
      gauss_bwd(i, a.G, a.cam, clamp, g2, out, sh ? row : nullptr);
      if (a.G.shs != nullptr && a.G.colors == nullptr) written = 3 * (sh_count(a.cam.deg) - 1);
    } else {
      for (int k = 0; k < 3; ++k) out.dmean3d[k] = 0.0f;
      for (int k = 0; k < 6; ++k) out.dcov3d[k] = 0.0f;
      for (int k = 0; k < 3; ++k) out.dscale[k] = 0.0f;
      for (int k = 0; k < 4; ++k) out.drot[k] = 0.0f;
    }
    if (sh)
      for (int k = written; k < 3 * M; ++k) row[k] = 0.0f;
    a.ddc[3 * (size_t)i] = ddc[0];
    a.ddc[3 * (size_t)i + 1] = ddc[1];
    a.ddc[3 * (size_t)i + 2] = ddc[2];
  }
  __syncthreads();
  if (sh)
    for (int k = threadIdx.x; k < 3 * M * nb; k += GBWD_THREADS)
      a.dsh[(size_t)3 * M * g0 + k] = rows[k];
}

// Batch of V > 1 views: one thread per Gaussian runs the adjoint of every
// view (records v * P + i) and sums the parameter gradients in view order;
// the screen-space gradients (dmean2D, dconic) are written per view. The SH
// rest rows are staged once; per view the adjoint writes its SH gradient to
// a scratch row that is added to an accumulator row (both in LDS).
constexpr int GBWDV_THREADS = 64;
__global__ __launch_bounds__(GBWDV_THREADS) void k_gauss_bwd_views(GaussBwdArgs a) {
  extern __shared__ float rows[];  // 3 x (GBWDV_THREADS x 3M): input, scratch, sum
  const int g0 = blockIdx.x * GBWDV_THREADS;
  const int i = g0 + threadIdx.x;
  const int M = a.cam.max_coeffs;
  const int RM = 3 * M;
  const bool sh = a.dsh != nullptr && M > 0;
  const int nb = min(GBWDV_THREADS, a.P - g0);
  float* shin = rows;
  float* shtmp = rows + GBWDV_THREADS * RM;
  float* shacc = rows + 2 * GBWDV_THREADS * RM;
  if (sh)
    for (int k = threadIdx.x; k < RM * nb; k += GBWDV_THREADS) {
      shin[k] = a.G.shs[(size_t)RM * g0 + k];
      shacc[k] = 0.0f;
    }
  __syncthreads();
  if (i < a.P) {
    float sm[3] = {0, 0, 0}, sc[6] = {0, 0, 0, 0, 0, 0}, ss[3] = {0, 0, 0}, sr[4] = {0, 0, 0, 0};
    float sdc[3] = {0, 0, 0}, scol[3] = {0, 0, 0}, sop = 0.0f;
    float* row_in = shin + RM * threadIdx.x;
    float* row_tmp = shtmp + RM * threadIdx.x;
    float* row_acc = shacc + RM * threadIdx.x;
    for (int v = 0; v < a.V; ++v) {
      const ViewRec& vr = a.views[v];
      const size_t ri = (size_t)v * a.P + i;
      float raw[RAW_STRIDE];
      load_raw(a.raw_lin, a.raw_con, ri, raw);
      Grad2D<float> g2;
      finalize_raw(raw, a.G.opac[i], vr.W, vr.H, g2);
      if (raw[RAW_POW] == 0.0f) g2.opacity = 0.0f;
      ((float4*)a.dmean2D)[ri] = make_float4(g2.mean2d[0], g2.mean2d[1], g2.mean2d[2], g2.mean2d[3]);
      if (a.dconic)
        ((float4*)a.dconic)[ri] = make_float4(g2.conic[0], g2.conic[1], g2.conic[2], g2.conic[3]);
      sop += g2.opacity;
      for (int k = 0; k < 3; ++k) scol[k] += g2.color[k];
      if (a.radii[ri] > 0) {
        float tm[3], tc[6], ts[3], tr[4], tdc[3] = {0.0f, 0.0f, 0.0f};
        GaussGrad<float> out;
        out.dmean3d = tm;
        out.dcov3d = tc;
        out.dscale = ts;
        out.drot = tr;
        out.ddc = tdc;
        out.dsh = sh ? row_tmp : nullptr;
        const int clamp = (int)f2u(a.splat[12 * ri + 11]);
        gauss_bwd(i, a.G, view_cam(a.cam, vr), clamp, g2, out, sh ? row_in : nullptr);
        for (int k = 0; k < 3; ++k) sm[k] += tm[k];
        for (int k = 0; k < 6; ++k) sc[k] += tc[k];
        for (int k = 0; k < 3; ++k) ss[k] += ts[k];
        for (int k = 0; k < 4; ++k) sr[k] += tr[k];
        for (int k = 0; k < 3; ++k) sdc[k] += tdc[k];
        if (sh && a.G.shs != nullptr && a.G.colors == nullptr) {
          const int written = 3 * (sh_count(a.cam.deg) - 1);
          for (int k = 0; k < written; ++k) row_acc[k] += row_tmp[k];
        }
      }
    }
    a.dopacity[i] = sop;
    for (int k = 0; k < 3; ++k) {
      if (a.dcolor) a.dcolor[3 * (size_t)i + k] = scol[k];
      a.dmean3D[3 * (size_t)i + k] = sm[k];
      a.dscale[3 * (size_t)i + k] = ss[k];
      a.ddc[3 * (size_t)i + k] = sdc[k];
    }
    if (a.dcov3D)
      for (int k = 0; k < 6; ++k) a.dcov3D[6 * (size_t)i + k] = sc[k];
    for (int k = 0; k < 4; ++k) a.drot[4 * (size_t)i + k] = sr[k];
  }
  __syncthreads();
  if (sh)
    for (int k = threadIdx.x; k < RM * nb; k += GBWDV_THREADS)
      a.dsh[(size_t)RM * g0 + k] = shacc[k];
}

// Split adjoint of one view, pass 1: the SH part. One wave64 per workgroup,
// one Gaussian per lane; the 64 coefficient rows (64 x 3M floats, contiguous)
// are staged in LDS with float4 loads, each lane reads its row (stride 3M
// words: odd for M = 15, so conflict-free), overwrites it with the
// coefficient gradients, and the rows are stored back with float4 stores.
// Lane i also writes ddc and the mean-gradient term sh_dm of its Gaussian.
constexpr int SHB_THREADS = WAVE;
// Rows staged as float4 with every load in flight before the first LDS
// store: the group's loads are unconditional (indices clamped into the
// group) and the LDS buffer is padded to SHB_QMAX x 64 float4 so that the
// stores need no predicate either.
constexpr int SHB_QMAX = (3 * 15 * SHB_THREADS / 4 + SHB_THREADS - 1) / SHB_THREADS;  // 12
constexpr size_t SHB_SHMEM = (size_t)SHB_QMAX * SHB_THREADS * sizeof(float4);

__global__ __launch_bounds__(SHB_THREADS) void k_sh_bwd(GaussBwdArgs a) {
  extern __shared__ float4 rows4[];  // SHB_THREADS x 3M floats (padded)
  float* rows = (float*)rows4;
  const int lane = threadIdx.x;
  const int g0 = blockIdx.x * SHB_THREADS;
  const int i = g0 + lane;
  const int M = a.cam.max_coeffs;
  const int RM = 3 * M;
  const int nb = min(SHB_THREADS, a.P - g0);
  const float* src = a.G.shs + (size_t)RM * g0;
  float* dst = a.dsh + (size_t)RM * g0;
  // per-Gaussian inputs first, so their loads overlap the staging
  const bool in = i < a.P;
  const size_t ic = in ? (size_t)i : 0;
  const bool vis = in && a.radii[ic] > 0;
  const float* pm = a.G.means + 3 * ic;
  const float px = pm[0], py = pm[1], pz = pm[2];
  const float4 c4 = ((const float4*)(a.raw_lin + RAW_LIN_STRIDE * ic))[0];
  const float col[3] = {c4.x, c4.y, c4.z};
  const int clampbits = (int)f2u(a.splat[12 * ic + 11]);
  // full groups of 64 rows start 16-byte aligned when the arrays do
  const bool vec = M > 0 && M <= 15 && nb == SHB_THREADS && RM * SHB_THREADS % 4 == 0 &&
                   (((uintptr_t)a.G.shs | (uintptr_t)a.dsh) & 15) == 0;
  const int nq = RM * SHB_THREADS / 4;
  if (vec) {
    float4 hv[SHB_QMAX];
#pragma unroll
    for (int u = 0; u < SHB_QMAX; ++u)
      hv[u] = ((const float4*)src)[min(lane + u * SHB_THREADS, nq - 1)];
#pragma unroll
    for (int u = 0; u < SHB_QMAX; ++u) rows4[lane + u * SHB_THREADS] = hv[u];
  } else {
    for (int j = lane; j < RM * nb; j += SHB_THREADS) rows[j] = src[j];
  }
  __syncthreads();
  if (in) {
    float* row = rows + RM * lane;
    float dm[3] = {0.0f, 0.0f, 0.0f}, ddc[3] = {0.0f, 0.0f, 0.0f};
    if (vis) {
      const float ox = px - a.cam.campos[0], oy = py - a.cam.campos[1], oz = pz - a.cam.campos[2];
      float dRGB[3];
      sh_drgb(clampbits, col, dRGB);
      const float len = sqrt_(ox * ox + oy * oy + oz * oz);
      const float x = ox / len, y = oy / len, z = oz / len;
      float bas[16], J[3][3], ddir[3];
      sh_basis(x, y, z, bas);
      const int n = sh_count(a.cam.deg);
      sh_jacobian(x, y, z, row, a.cam.deg, J);
      sh_ddir(J, dRGB, ddir);
#pragma unroll
      for (int k = 1; k < 16; ++k)
        if (k < n)
          for (int ch = 0; ch < 3; ++ch) row[3 * (k - 1) + ch] = bas[k] * dRGB[ch];
      for (int ch = 0; ch < 3; ++ch) ddc[ch] = bas[0] * dRGB[ch];
      for (int j = 3 * (n - 1); j < RM; ++j) row[j] = 0.0f;
      sh_dmean(ox, oy, oz, ddir, dm);
    } else {
      for (int j = 0; j < RM; ++j) row[j] = 0.0f;
    }
    ((float4*)a.sh_dm)[i] = make_float4(dm[0], dm[1], dm[2], 0.0f);
    for (int ch = 0; ch < 3; ++ch) a.ddc[3 * (size_t)i + ch] = ddc[ch];
  }
  __syncthreads();
  if (vec) {
    float4 ov[SHB_QMAX];
#pragma unroll
    for (int u = 0; u < SHB_QMAX; ++u) ov[u] = rows4[lane + u * SHB_THREADS];
#pragma unroll
    for (int u = 0; u < SHB_QMAX; ++u)
      if (lane + u * SHB_THREADS < nq) ((float4*)dst)[lane + u * SHB_THREADS] = ov[u];
  } else {
    for (int j = lane; j < RM * nb; j += SHB_THREADS) dst[j] = rows[j];
  }
}

// Split adjoint of one view, pass 2: everything but the SH coefficients,
// one thread per Gaussian without LDS (full occupancy), adding pass 1's
// mean-gradient term.
__global__ __launch_bounds__(256) void k_gauss_bwd_geom(GaussBwdArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.P) return;
  float raw[RAW_STRIDE];
  load_raw(a.raw_lin, a.raw_con, (size_t)i, raw);
  Grad2D<float> g2;
  finalize_raw(raw, a.G.opac[i], a.cam.W, a.cam.H, g2);
  if (raw[RAW_POW] == 0.0f) g2.opacity = 0.0f;
  ((float4*)a.dmean2D)[i] = make_float4(g2.mean2d[0], g2.mean2d[1], g2.mean2d[2], g2.mean2d[3]);
  if (a.dconic)
    ((float4*)a.dconic)[i] = make_float4(g2.conic[0], g2.conic[1], g2.conic[2], g2.conic[3]);
  a.dopacity[i] = g2.opacity;
  if (a.dcolor) {
    a.dcolor[3 * (size_t)i] = g2.color[0];
    a.dcolor[3 * (size_t)i + 1] = g2.color[1];
    a.dcolor[3 * (size_t)i + 2] = g2.color[2];
  }
  float tm[3], tc[6], ts[3], tr[4], tdc[3];
  GaussGrad<float> out;
  out.dmean3d = tm;
  out.dcov3d = tc;
  out.dscale = ts;
  out.drot = tr;
  out.ddc = tdc;
  out.dsh = nullptr;
  if (a.radii[i] > 0) {
    const float4 sd = ((const float4*)a.sh_dm)[i];
    const float dms[3] = {sd.x, sd.y, sd.z};
    const int clamp = (int)f2u(a.splat[12 * (size_t)i + 11]);
    gauss_bwd(i, a.G, a.cam, clamp, g2, out, (const float*)nullptr, dms);
  } else {
    for (int k = 0; k < 3; ++k) tm[k] = ts[k] = 0.0f;
    for (int k = 0; k < 6; ++k) tc[k] = 0.0f;
    for (int k = 0; k < 4; ++k) tr[k] = 0.0f;
  }
  for (int k = 0; k < 3; ++k) {
    a.dmean3D[3 * (size_t)i + k] = tm[k];
    a.dscale[3 * (size_t)i + k] = ts[k];
  }
  if (a.dcov3D)
    for (int k = 0; k < 6; ++k) a.dcov3D[6 * (size_t)i + k] = tc[k];
  ((float4*)a.drot)[i] = make_float4(tr[0], tr[1], tr[2], tr[3]);
}

// Adjoint of one view when the forward's colour pass stored the SH
// Jacobians: the geometry adjoint of k_gauss_bwd_geom and the SH part in one
// pass without reading the coefficients (their gradients bas_k dRGB need
// only the direction; the mean term takes dL/d(direction) from the
// Jacobian). One wave64 per workgroup, one Gaussian per lane; each lane
// writes its coefficient-gradient row into LDS and the rows leave with
// coalesced float4 stores.
// The coefficient-gradient rows go out through LDS 32 rows at a time (6 KB
// per wave instead of 12 KB for all 64: the LDS a wave holds bounds the
// waves per CU, and this pass is HBM-bound).
constexpr int SHJ_ROWS = 32;
constexpr int SHJ_QMAX = (3 * 15 * SHJ_ROWS / 4 + SHB_THREADS - 1) / SHB_THREADS;  // 6
constexpr size_t SHJ_SHMEM = (size_t)SHJ_QMAX * SHB_THREADS * sizeof(float4);
__global__ __launch_bounds__(SHB_THREADS) void k_gauss_bwd_jac(GaussBwdArgs a) {
  extern __shared__ float4 rows4[];  // SHJ_ROWS x 3M floats (padded)
  float* rows = (float*)rows4;
  const int lane = threadIdx.x;
  const int g0 = blockIdx.x * SHB_THREADS;
  const int i = g0 + lane;
  const int M = a.cam.max_coeffs;
  const int RM = 3 * M;
  const bool sh = a.dsh != nullptr && M > 0;
  const bool colour_sh = a.G.colors == nullptr && a.G.dc != nullptr;
  float dRGB[3] = {0.0f, 0.0f, 0.0f};  // the colour adjoint, kept for the SH rows
  bool rowv = false;                   // this lane's SH row is non-zero
  if (i < a.P) {
    float raw[RAW_STRIDE];
    load_raw(a.raw_lin, a.raw_con, (size_t)i, raw);
    Grad2D<float> g2;
    finalize_raw(raw, a.G.opac[i], a.cam.W, a.cam.H, g2);
    if (raw[RAW_POW] == 0.0f) g2.opacity = 0.0f;
    ((float4*)a.dmean2D)[i] = make_float4(g2.mean2d[0], g2.mean2d[1], g2.mean2d[2], g2.mean2d[3]);
    if (a.dconic)
      ((float4*)a.dconic)[i] = make_float4(g2.conic[0], g2.conic[1], g2.conic[2], g2.conic[3]);
    a.dopacity[i] = g2.opacity;
    if (a.dcolor) {
      a.dcolor[3 * (size_t)i] = g2.color[0];
      a.dcolor[3 * (size_t)i + 1] = g2.color[1];
      a.dcolor[3 * (size_t)i + 2] = g2.color[2];
    }
    float tm[3], tc[6], ts[3], tr[4], ddc[3] = {0.0f, 0.0f, 0.0f};
    const int radius = a.radii[i];
    if (radius > 0) {
      // densification statistics (backward.cu:892-918, as k_densify_stats)
      if (a.denom) a.denom[i] += 1.0f;
      if (a.max_radii) a.max_radii[i] = fmaxf(a.max_radii[i], (float)radius);
      if (a.accum)
        a.accum[i] += sqrtf(g2.mean2d[0] * g2.mean2d[0] + g2.mean2d[1] * g2.mean2d[1]);
      if (a.accum_abs)
        a.accum_abs[i] += sqrtf(g2.mean2d[2] * g2.mean2d[2] + g2.mean2d[3] * g2.mean2d[3]);
      int clamp = 0;
      float dms[3] = {0.0f, 0.0f, 0.0f};
      if (colour_sh) {
        const float* pm = a.G.means + 3 * (size_t)i;
        const float ox = pm[0] - a.cam.campos[0], oy = pm[1] - a.cam.campos[1],
                    oz = pm[2] - a.cam.campos[2];
        const float4* jr = (const float4*)(a.shjac + 12 * (size_t)i);
        const float4 j0 = jr[0], j1 = jr[1], j2 = jr[2];
        clamp = (int)f2u(j0.w);
        float J[3][3], ddir[3];
        sh_drgb(clamp, g2.color, dRGB);
        J[0][0] = j0.x; J[0][1] = j0.y; J[0][2] = j0.z;
        J[1][0] = j1.x; J[1][1] = j1.y; J[1][2] = j1.z;
        J[2][0] = j2.x; J[2][1] = j2.y; J[2][2] = j2.z;
        sh_ddir(J, dRGB, ddir);
        // degree-0 basis: a constant
        float bas[16];
        const float len = sqrt_(ox * ox + oy * oy + oz * oz);
        sh_basis(ox / len, oy / len, oz / len, bas);
        for (int ch = 0; ch < 3; ++ch) ddc[ch] = bas[0] * dRGB[ch];
        rowv = sh;
        sh_dmean(ox, oy, oz, ddir, dms);
      }
      GaussGrad<float> out;
      out.dmean3d = tm;
      out.dcov3d = tc;
      out.dscale = ts;
      out.drot = tr;
      out.ddc = nullptr;
      out.dsh = nullptr;
      gauss_bwd(i, a.G, a.cam, clamp, g2, out, (const float*)nullptr, dms);
    } else {
      for (int k = 0; k < 3; ++k) tm[k] = ts[k] = 0.0f;
      for (int k = 0; k < 6; ++k) tc[k] = 0.0f;
      for (int k = 0; k < 4; ++k) tr[k] = 0.0f;
    }
    for (int k = 0; k < 3; ++k) {
      a.dmean3D[3 * (size_t)i + k] = tm[k];
      a.dscale[3 * (size_t)i + k] = ts[k];
      if (a.ddc) a.ddc[3 * (size_t)i + k] = ddc[k];
    }
    if (a.dcov3D)
      for (int k = 0; k < 6; ++k) a.dcov3D[6 * (size_t)i + k] = tc[k];
    ((float4*)a.drot)[i] = make_float4(tr[0], tr[1], tr[2], tr[3]);
  }
  if (!sh) return;
  // coefficient gradients bas_k * dRGB (k >= 1), the basis recomputed from
  // the direction (the geometric adjoint's registers are free by now)
  float bas[16];
  int n = 0;
  if (rowv) {
    const float* pm = a.G.means + 3 * (size_t)i;
    const float ox = pm[0] - a.cam.campos[0], oy = pm[1] - a.cam.campos[1],
                oz = pm[2] - a.cam.campos[2];
    const float len = sqrt_(ox * ox + oy * oy + oz * oz);
    sh_basis(ox / len, oy / len, oz / len, bas);
    n = sh_count(a.cam.deg);
  }
  const bool al16 = ((uintptr_t)a.dsh & 15) == 0 && M <= 15;
  for (int h = 0; h < SHB_THREADS / SHJ_ROWS; ++h) {
    const int gb = g0 + h * SHJ_ROWS;
    if (gb >= a.P) break;  // uniform
    if (h) __syncthreads();  // the previous rows are copied out
    if (lane / SHJ_ROWS == h) {
      float* row = rows + RM * (lane % SHJ_ROWS);
#pragma unroll
      for (int k = 1; k < 16; ++k)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch)
          if (3 * (k - 1) + ch < RM) row[3 * (k - 1) + ch] = k < n ? bas[k] * dRGB[ch] : 0.0f;
    }
    __syncthreads();
    const int nr = min(SHJ_ROWS, a.P - gb);
    float* dst = a.dsh + (size_t)RM * gb;
    if (al16 && nr == SHJ_ROWS) {
      const int nq = RM * SHJ_ROWS / 4;
      float4 ov[SHJ_QMAX];
#pragma unroll
      for (int u = 0; u < SHJ_QMAX; ++u) ov[u] = rows4[lane + u * SHB_THREADS];
#pragma unroll
      for (int u = 0; u < SHJ_QMAX; ++u)
        if (lane + u * SHB_THREADS < nq) ((float4*)dst)[lane + u * SHB_THREADS] = ov[u];
    } else {
      for (int j = lane; j < RM * nr; j += SHB_THREADS) dst[j] = rows[j];
    }
  }
}

// The adjoint pass for a's shape; stats_done: it also added the
// densification statistics.
static hipError_t launch_gauss_adjoint(const GaussBwdArgs& a, hipStream_t s, bool& stats_done) {
  stats_done = false;
  const int M = a.cam.max_coeffs;
  const bool sh = a.dsh != nullptr && M > 0;
  const bool colour_sh = a.G.colors == nullptr && a.G.dc != nullptr;
  if (a.shjac != nullptr && (a.views == nullptr || a.V <= 1) && M <= 15) {
    const size_t shmem = sh ? SHJ_SHMEM : 0;
    FGS_TIMED("gauss_bwd_jac", s,
              (k_gauss_bwd_jac<<<(unsigned)((a.P + SHB_THREADS - 1) / SHB_THREADS), SHB_THREADS,
                                 shmem, s>>>(a)));
    stats_done = true;
    return hipGetLastError();
  }
  if (a.sh_dm != nullptr && (a.views == nullptr || a.V <= 1) && M <= 15) {
    if (colour_sh) {
      const size_t need = (size_t)SHB_THREADS * 3 * M * sizeof(float);
      const size_t shmem = need > SHB_SHMEM ? need : SHB_SHMEM;
      FGS_TIMED("sh_bwd", s,
                (k_sh_bwd<<<(unsigned)((a.P + SHB_THREADS - 1) / SHB_THREADS), SHB_THREADS,
                             shmem, s>>>(a)));
    } else {  // precomputed colours: no SH term, zero dc gradient
      hipError_t e = fill_f32(a.sh_dm, 0.0f, (int64_t)a.P * 4, s);
      if (e == hipSuccess && a.ddc) e = fill_f32(a.ddc, 0.0f, (int64_t)a.P * 3, s);
      if (e != hipSuccess) return e;
    }
    FGS_TIMED("gauss_bwd_geom", s, (k_gauss_bwd_geom<<<(a.P + 255) / 256, 256, 0, s>>>(a)));
    return hipGetLastError();
  }
  if (a.views != nullptr && a.V > 1) {
    const size_t shmem = sh ? (size_t)3 * GBWDV_THREADS * 3 * M * sizeof(float) : 0;
    FGS_TIMED("gauss_bwd_views", s,
              (k_gauss_bwd_views<<<(a.P + GBWDV_THREADS - 1) / GBWDV_THREADS, GBWDV_THREADS, shmem,
                                   s>>>(a)));
    return hipGetLastError();
  }
  const size_t shmem = sh ? (size_t)GBWD_THREADS * 3 * M * sizeof(float) : 0;
  FGS_TIMED("gauss_bwd", s,
            (k_gauss_bwd<<<(a.P + GBWD_THREADS - 1) / GBWD_THREADS, GBWD_THREADS, shmem, s>>>(a)));
  return hipGetLastError();
}

hipError_t launch_gauss_bwd(const GaussBwdArgs& a, hipStream_t s) {
  if (a.P <= 0) return hipSuccess;
  bool stats_done = false;
  hipError_t e = launch_gauss_adjoint(a, s, stats_done);
  if (e == hipSuccess && !stats_done && (a.accum || a.accum_abs || a.max_radii || a.denom))
    e = launch_densify_stats(a.P, a.views != nullptr && a.V > 1 ? a.V : 1, a.radii, a.dmean2D,
                             a.accum, a.accum_abs, a.max_radii, a.denom, s);
  return e;
}

__global__ __launch_bounds__(256) void k_densify_stats(int P, int V,
                                                       const int32_t* __restrict__ radii,
                                                       const float* __restrict__ dmean2D,
                                                       float* accum, float* accum_abs,
                                                       float* max_radii, float* denom) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  for (int v = 0; v < V; ++v) {  // one view after the other, as separate calls would
    const size_t ri = (size_t)v * P + i;
    if (!(radii[ri] > 0)) continue;
    const float4 g = ((const float4*)dmean2D)[ri];
    if (denom) denom[i] += 1.0f;
    if (max_radii) max_radii[i] = fmaxf(max_radii[i], (float)radii[ri]);
    if (accum) accum[i] += sqrtf(g.x * g.x + g.y * g.y);
    if (accum_abs) accum_abs[i] += sqrtf(g.z * g.z + g.w * g.w);
  }
}

hipError_t launch_densify_stats(int P, int V, const int32_t* radii, const float* dmean2D,
                                float* accum, float* accum_abs, float* max_radii, float* denom,
                                hipStream_t s) {
  if (P > 0 && V > 0)
    k_densify_stats<<<(P + 255) / 256, 256, 0, s>>>(P, V, radii, dmean2D, accum, accum_abs,
                                                    max_radii, denom);
  return hipGetLastError();
}

}  // namespace fgs
