// CPU oracle of the FasterGS rasterizer hot path. TEST INFRASTRUCTURE ONLY:
// it is the checker that the gfx950 kernels are compared against and the
// timed CPU baseline; no product code path may call it.
//
// It restates the reference pipeline (FGS = gs_submodules/diff-gaussian-
// rasterization_faster in the FasterGS repository) on the CPU:
//   preprocess + exact tile test + key emission ... FGS/csrc/cuda_rasterizer/preprocess.cu:414-746
//   stable sort of (tile<<32 | depth bits)       ... FGS/csrc/cuda_rasterizer/sort.cu:28-44
//   identifyTileRanges                           ... FGS/csrc/cuda_rasterizer/render.cu:54-79
//   per-tile bucket count / offsets / ranges     ... render.cu:13-49, :418-434
//   renderBucket (bucket from T=1)               ... render.cu:130-275
//   mergeResults (bucket merge, T<1e-4 cutoff)   ... render.cu:286-391
//   bucket backward                              ... FGS/csrc/cuda_rasterizer/backward.cu:695-878
//   preprocess backward                          ... backward.cu:250-539, :922-954
//   densification statistics                     ... backward.cu:892-918
// The per-Gaussian and per-pixel arithmetic is shared with the kernels via
// fastergs_amd/csrc/fgs_math.h (see the header comment there). Differences
// from the reference that are decisions of this project:
//   * pair order is canonical: (tile, depth bits, Gaussian index); the
//     reference breaks depth ties by atomic emission order;
//   * exp2/log2 are the project's own polynomials instead of ex2.approx /
//     lg2.approx, and sqrt/div are IEEE-correct instead of sqrt.approx;
//   * the radius formula is evaluated in float instead of double.
// Mode 0 ("bucketed") restates the reference compositing exactly; mode 1
// ("sequential") is the plain path: one running (T,C) per pixel over the
// whole list, with the same T<1e-4 test at bucket boundaries, and a
// per-pixel reverse backward starting from T_final.
//
// Per-Gaussian gradient sums: float partial sums over fixed groups of
// pixels (the lane quads of the GPU walk, see backward()), added in double
// per sorted pair and then in sorted-pair order, so results do not depend on
// the thread count.
#include <parallel/algorithm>
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../fastergs_amd/csrc/fgs_math.h"

namespace {

using fgs::CamRef;
using fgs::GaussRef;

template <typename R>
struct State {
  int P = 0, W = 0, H = 0, gx = 0, gy = 0, T = 0, B = 32, mode = 0, deg = 0, M = 0;
  int tw = 16, th = 16;  // tile shape (tw x th = 256 pixels)
  R tanfovx, tanfovy, focal_x, focal_y, scale_mod, mult;
  std::vector<R> view, proj, campos, bg;
  std::vector<R> means, opac, scales, rots, dc, shs, colors;
  std::vector<uint8_t> cull;
  bool has_sh = false, has_colors = false, has_cull = false;

  // forward products
  std::vector<fgs::Proj<R>> pj;
  std::vector<int32_t> radii;
  std::vector<uint32_t> pair_count;   // per Gaussian
  std::vector<uint64_t> keys;         // sorted
  std::vector<uint32_t> vals;         // sorted
  std::vector<uint32_t> ranges;       // T x 2
  std::vector<uint32_t> bucket_off;   // T, inclusive scan of bucket counts
  std::vector<uint32_t> bucket_ranges;  // NB x 2
  std::vector<uint32_t> bucket_tile;    // NB
  std::vector<R> out, tfinal, pixc, acc_t, acc_c;
  std::vector<int32_t> contrib;
  std::vector<uint32_t> max_contrib;
  // per splat render data: x, y, ca, cb, cc, w, r, g, b
  std::vector<R> splat;

  // backward products
  std::vector<double> raw;  // P x RAW_N
  std::vector<R> dmean2d, dconic, dopac, dcolor, dmean3d, dcov3d, ddc, dsh, dscale, drot;

  long long R_pairs() const { return (long long)vals.size(); }
  long long NB() const { return (long long)bucket_tile.size(); }

  CamRef<R> cam() const {
    CamRef<R> c;
    c.view = view.data();
    c.proj = proj.data();
    c.campos = campos.data();
    c.tanfovx = tanfovx;
    c.tanfovy = tanfovy;
    c.focal_x = focal_x;
    c.focal_y = focal_y;
    c.scale_mod = scale_mod;
    c.mult = mult;
    c.W = W;
    c.H = H;
    c.grid_x = gx;
    c.grid_y = gy;
    c.tile_w = tw;
    c.tile_h = th;
    c.deg = deg;
    c.max_coeffs = M;
    return c;
  }
  GaussRef<R> gauss() const {
    GaussRef<R> g;
    g.means = means.data();
    g.opac = opac.data();
    g.scales = scales.data();
    g.rots = rots.data();
    g.dc = has_sh ? dc.data() : nullptr;
    g.shs = has_sh ? shs.data() : nullptr;
    g.colors = has_colors ? colors.data() : nullptr;
    g.culling = has_cull ? cull.data() : nullptr;
    return g;
  }
};

template <typename R, typename S>
void copy_in(std::vector<R>& dst, const S* src, size_t n) {
  dst.resize(n);
  for (size_t k = 0; k < n; ++k) dst[k] = (R)src[k];
}

template <typename R>
void preprocess(State<R>& st) {
  const auto cam = st.cam();
  const auto G = st.gauss();
  st.pj.assign(st.P, fgs::Proj<R>());
  st.radii.assign(st.P, 0);
  st.pair_count.assign(st.P, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < st.P; ++i) {
    fgs::Proj<R>& o = st.pj[i];
    fgs::preprocess_gauss(i, G, cam, o);
    if (!o.valid) continue;
    st.radii[i] = o.radius;
    uint32_t n = 0;
    for (int ty = o.rect[1]; ty < o.rect[3]; ++ty)
      for (int tx = o.rect[0]; tx < o.rect[2]; ++tx)
        n += fgs::tile_hit(tx, ty, st.tw, st.th, o.x, o.y, o.conic, o.power) ? 1u : 0u;
    st.pair_count[i] = n;
  }
}

struct Pair {
  uint64_t key;
  uint32_t val;
};

template <typename R>
void emit_and_sort(State<R>& st) {
  std::vector<uint64_t> off(st.P + 1, 0);
  for (int i = 0; i < st.P; ++i) off[i + 1] = off[i] + st.pair_count[i];
  const size_t Rn = off[st.P];
  std::vector<Pair> pairs(Rn);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < st.P; ++i) {
    const fgs::Proj<R>& o = st.pj[i];
    if (!o.valid || st.pair_count[i] == 0) continue;
    size_t k = off[i];
    const float dz = (float)o.depth;
    for (int ty = o.rect[1]; ty < o.rect[3]; ++ty)
      for (int tx = o.rect[0]; tx < o.rect[2]; ++tx)
        if (fgs::tile_hit(tx, ty, st.tw, st.th, o.x, o.y, o.conic, o.power)) {
          const uint64_t tile = (uint64_t)(ty * st.gx + tx);
          pairs[k].key = (tile << 32) | fgs::depth_bits(dz);
          pairs[k].val = (uint32_t)i;
          ++k;
        }
  }
  // Canonical order (key, index): equals a stable sort by key of the
  // index-ordered emission.
  __gnu_parallel::sort(pairs.begin(), pairs.end(), [](const Pair& a, const Pair& b) {
    return a.key < b.key || (a.key == b.key && a.val < b.val);
  });
  st.keys.resize(Rn);
  st.vals.resize(Rn);
  for (size_t k = 0; k < Rn; ++k) {
    st.keys[k] = pairs[k].key;
    st.vals[k] = pairs[k].val;
  }
}

template <typename R>
void ranges_and_buckets(State<R>& st) {
  st.ranges.assign(2 * (size_t)st.T, 0);
  const size_t Rn = st.vals.size();
  for (size_t k = 0; k < Rn; ++k) {  // identifyTileRanges (render.cu:54-79)
    const uint32_t t = (uint32_t)(st.keys[k] >> 32);
    if (k == 0 || (uint32_t)(st.keys[k - 1] >> 32) != t) st.ranges[2 * t] = (uint32_t)k;
    if (k + 1 == Rn || (uint32_t)(st.keys[k + 1] >> 32) != t) st.ranges[2 * t + 1] = (uint32_t)(k + 1);
  }
  st.bucket_off.assign(st.T, 0);
  uint32_t acc = 0;
  for (int t = 0; t < st.T; ++t) {
    const uint32_t len = st.ranges[2 * t + 1] - st.ranges[2 * t];
    acc += (len + st.B - 1) / st.B;
    st.bucket_off[t] = acc;
  }
  st.bucket_ranges.assign(2 * (size_t)acc, 0);
  st.bucket_tile.assign(acc, 0);
  for (int t = 0; t < st.T; ++t) {
    uint32_t b = t == 0 ? 0 : st.bucket_off[t - 1];
    uint32_t s = st.ranges[2 * t];
    const uint32_t e = st.ranges[2 * t + 1];
    while (s < e) {
      const uint32_t n = std::min<uint32_t>(st.B, e - s);
      st.bucket_ranges[2 * b] = s;
      st.bucket_ranges[2 * b + 1] = s + n;
      st.bucket_tile[b] = (uint32_t)t;
      s += n;
      ++b;
    }
  }
}

template <typename R>
void build_splats(State<R>& st) {
  st.splat.assign((size_t)st.P * 9, R(0));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < st.P; ++i) {
    const fgs::Proj<R>& o = st.pj[i];
    if (!o.valid) continue;
    R* s = &st.splat[9 * (size_t)i];
    s[0] = o.x;
    s[1] = o.y;
    fgs::log2_conic(o.conic, o.log2op, s[2], s[3], s[4], s[5]);
    s[6] = o.rgb[0];
    s[7] = o.rgb[1];
    s[8] = o.rgb[2];
  }
}

template <typename R>
void composite(State<R>& st) {
  const size_t HW = (size_t)st.W * st.H;
  st.out.assign(3 * HW, R(0));
  st.tfinal.assign(HW, R(0));
  st.pixc.assign(3 * HW, R(0));
  st.contrib.assign(HW, 0);
  st.max_contrib.assign(st.T, 0);
  const size_t NB = st.bucket_tile.size();
  if (st.mode == 0) {
    st.acc_t.assign(NB * fgs::TILE_PIX, R(0));
    st.acc_c.assign(NB * fgs::TILE_PIX * 3, R(0));
  } else {
    st.acc_t.clear();
    st.acc_c.clear();
  }
  const R bg0 = st.bg[0], bg1 = st.bg[1], bg2 = st.bg[2];
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < st.T; ++t) {
    const int tx = t % st.gx, ty = t / st.gx;
    const uint32_t b0 = t == 0 ? 0 : st.bucket_off[t - 1];
    const uint32_t b1 = st.bucket_off[t];
    uint32_t mc = 0;
    for (int ly = 0; ly < st.th; ++ly)
      for (int lx = 0; lx < st.tw; ++lx) {
        const int px = tx * st.tw + lx, py = ty * st.th + ly;
        if (px >= st.W || py >= st.H) continue;
        const int lp = ly * st.tw + lx;
        const R fx = (R)px, fy = (R)py;
        R T = R(1), C0 = R(0), C1 = R(0), C2 = R(0);
        int32_t cnt = 0;
        for (uint32_t b = b0; b < b1; ++b) {
          const uint32_t s = st.bucket_ranges[2 * b], e = st.bucket_ranges[2 * b + 1];
          if (st.mode == 0) {
            R Tb = R(1), D0 = R(0), D1 = R(0), D2 = R(0);
            for (uint32_t k = s; k < e; ++k) {
              const R* sp = &st.splat[9 * (size_t)st.vals[k]];
              const R a = fgs::splat_alpha(sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], fx, fy);
              fgs::blend(a, sp[6], sp[7], sp[8], Tb, D0, D1, D2);
            }
            fgs::merge(Tb, D0, D1, D2, T, C0, C1, C2);
            const size_t q = (size_t)b * fgs::TILE_PIX + lp;
            st.acc_t[q] = T;
            st.acc_c[3 * q] = C0;
            st.acc_c[3 * q + 1] = C1;
            st.acc_c[3 * q + 2] = C2;
          } else {
            for (uint32_t k = s; k < e; ++k) {
              const R* sp = &st.splat[9 * (size_t)st.vals[k]];
              const R a = fgs::splat_alpha(sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], fx, fy);
              fgs::blend(a, sp[6], sp[7], sp[8], T, C0, C1, C2);
            }
          }
          cnt += (int32_t)(e - s);
          if (T < R(fgs::T_CUT_F)) break;
        }
        const size_t g = (size_t)py * st.W + px;
        st.contrib[g] = cnt;
        st.tfinal[g] = T;
        st.out[g] = fgs::fma_(T, bg0, C0);
        st.out[HW + g] = fgs::fma_(T, bg1, C1);
        st.out[2 * HW + g] = fgs::fma_(T, bg2, C2);
        st.pixc[3 * g] = C0;
        st.pixc[3 * g + 1] = C1;
        st.pixc[3 * g + 2] = C2;
        mc = std::max(mc, (uint32_t)cnt);
      }
    st.max_contrib[t] = mc;
  }
}

template <typename R>
void forward(State<R>& st) {
  preprocess(st);
  emit_and_sort(st);
  ranges_and_buckets(st);
  build_splats(st);
  composite(st);
}

template <typename R>
void gauss_grads(State<R>& st, const float* raw_f);

// Raw per-Gaussian sums of the rasterizer adjoint, formed exactly as the
// gfx950 reverse walks form them, so that GPU and oracle agree to the last
// bit of the float-rounded sums. A wave of the walk covers 64 lanes; per
// splat each lane accumulates its pixels' terms in float in a fixed order
// (fused accumulation from -0 / 0), the wave adds its 64 lane values in float
// along a fixed tree (wave_total), and the wave totals are added in double.
//   FAST  (mode 0): one wave per bucket, lane l holds the 2x2 pixels
//                   (2 (l % 8) + i, 2 (l / 8) + j), in the order (0,0) (1,0)
//                   (0,1) (1,1);
//   PLAIN (mode 1): four waves per tile, lane l of wave w holds pixel
//                   (l % 16, 4 w + l / 16).
// The reference adds float warp sums with float atomics (backward.cu:634-693),
// whose order varies between runs; the double addition of fixed float wave
// totals is that computation with a deterministic result.

// Float tree of the GPU's wave reduction for field f over lane values a[64]:
// lane-quad sums q_j = (a_4j + a_4j+1) + (a_4j+2 + a_4j+3); per row of 16
// lanes (q_4r .. q_4r+3, indices mod 4 below) (q_g + q_g-1) + (q_g-2 + q_g-3)
// with g = (f / 4) mod 4, the order the transposing butterfly's row
// rotations produce; then (row0 + row1) + (row2 + row3).
template <typename R>
R wave_total(const R (*a)[fgs::RAW_N], int f) {
  R q[16];
  for (int j = 0; j < 16; ++j)
    q[j] = (a[4 * j][f] + a[4 * j + 1][f]) + (a[4 * j + 2][f] + a[4 * j + 3][f]);
  const int g = (f >> 2) & 3;
  R row[4];
  for (int r = 0; r < 4; ++r) {
    const R* qq = &q[4 * r];
    row[r] = (qq[g] + qq[(g + 3) & 3]) + (qq[(g + 2) & 3] + qq[(g + 1) & 3]);
  }
  return (row[0] + row[1]) + (row[2] + row[3]);
}

template <typename R>
void backward(State<R>& st, const R* dimg) {
  const size_t HW = (size_t)st.W * st.H;
  const size_t Rn = st.vals.size();
  std::vector<double> rows(Rn * fgs::RAW_N, 0.0);
  const R bg0 = st.bg[0], bg1 = st.bg[1], bg2 = st.bg[2];
  constexpr int NF = fgs::RAW_N;
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < st.T; ++t) {
    const int tx = t % st.gx, ty = t / st.gx;
    const uint32_t b0 = t == 0 ? 0 : st.bucket_off[t - 1];
    const uint32_t b1 = st.bucket_off[t];
    const uint32_t tile_start = st.ranges[2 * t];
    // per pixel of the tile (index ly * tw + lx)
    struct Pix {
      bool in = false;
      R fx, fy, dp0, dp1, dp2, tb;
      uint32_t cnt = 0;
      size_t g = 0;
    };
    Pix pix[fgs::TILE_PIX];
    for (int lp = 0; lp < fgs::TILE_PIX; ++lp) {
      const int px = tx * st.tw + lp % st.tw, py = ty * st.th + lp / st.tw;
      Pix& q = pix[lp];
      if (px >= st.W || py >= st.H) continue;
      q.in = true;
      q.g = (size_t)py * st.W + px;
      q.fx = (R)px;
      q.fy = (R)py;
      q.dp0 = dimg[q.g];
      q.dp1 = dimg[HW + q.g];
      q.dp2 = dimg[2 * HW + q.g];
      q.tb = st.tfinal[q.g] * (bg0 * q.dp0 + bg1 * q.dp1 + bg2 * q.dp2);
      q.cnt = (uint32_t)st.contrib[q.g];
    }
    auto add_wave = [&](uint32_t k, R (*lane)[NF]) {
      double* row = &rows[k * NF];
      for (int f = 0; f < NF; ++f) row[f] += (double)wave_total(lane, f);
    };
    auto step = [&](uint32_t k, const Pix& q, R& T, R& Q, R* acc) {
      const R* sp = &st.splat[9 * (size_t)st.vals[k]];
      fgs::pix_bwd(sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], sp[6], sp[7], sp[8], q.fx, q.fy,
                   q.dp0, q.dp1, q.dp2, T, Q, acc);
    };
    R lane[64][NF];
    if (st.mode == 0) {
      for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t s = st.bucket_ranges[2 * b], e = st.bucket_ranges[2 * b + 1];
        int lp[64][4];
        bool act[64][4];
        R T[64][4], Q[64][4];
        bool any = false;
        for (int l = 0; l < 64; ++l)
          for (int m = 0; m < 4; ++m) {
            // lane l's 2x2 quad: quads in row-major order over the tile's
            // tw/2 x th/2 grid
            const int lx = 2 * (l % (st.tw / 2)) + (m & 1), ly = 2 * (l / (st.tw / 2)) + (m >> 1);
            const int p = ly * st.tw + lx;
            lp[l][m] = p;
            act[l][m] = pix[p].in && s - tile_start < pix[p].cnt;  // backward.cu:663-664
            any = any || act[l][m];
            if (!act[l][m]) continue;
            const size_t q = (size_t)b * fgs::TILE_PIX + p;
            const size_t g = pix[p].g;
            T[l][m] = st.acc_t[q];
            Q[l][m] = fgs::residual_dot(st.pixc[3 * g] - st.acc_c[3 * q],
                                        st.pixc[3 * g + 1] - st.acc_c[3 * q + 1],
                                        st.pixc[3 * g + 2] - st.acc_c[3 * q + 2], pix[p].dp0,
                                        pix[p].dp1, pix[p].dp2, pix[p].tb);
          }
        if (!any) break;  // later buckets are needed by no pixel either
        for (uint32_t k = e; k-- > s;) {
          for (int l = 0; l < 64; ++l) {
            for (int f = 0; f < NF; ++f) lane[l][f] = R(-0.0);
            for (int m = 0; m < 4; ++m)
              if (act[l][m]) step(k, pix[lp[l][m]], T[l][m], Q[l][m], lane[l]);
          }
          add_wave(k, lane);
        }
      }
    } else {
      for (int w = 0; w < 4; ++w) {  // pixels 64 w .. 64 w + 63 (row-major)
        R T[64], Q[64];
        uint32_t top = 0;
        for (int l = 0; l < 64; ++l) {
          const Pix& q = pix[64 * w + l];
          T[l] = q.in ? st.tfinal[q.g] : R(0);
          Q[l] = q.in ? q.tb : R(0);
          if (q.in) top = std::max(top, q.cnt);
        }
        for (uint32_t k = tile_start + top; k-- > tile_start;) {
          for (int l = 0; l < 64; ++l) {
            for (int f = 0; f < NF; ++f) lane[l][f] = R(0);
            const Pix& q = pix[64 * w + l];
            if (q.in && k - tile_start < q.cnt) step(k, q, T[l], Q[l], lane[l]);
          }
          add_wave(k, lane);
        }
      }
    }
  }
  st.raw.assign((size_t)st.P * fgs::RAW_N, 0.0);
  for (size_t k = 0; k < Rn; ++k) {
    double* dst = &st.raw[(size_t)st.vals[k] * fgs::RAW_N];
    const double* src = &rows[k * fgs::RAW_N];
    for (int f = 0; f < fgs::RAW_N; ++f) dst[f] += src[f];
  }

  gauss_grads(st, nullptr);
}

// Per-Gaussian part of the backward (backward.cu:250-539) from raw sums
// given as float rows of RAW_STRIDE (raw_f) or as the double sums in st.raw.
template <typename R>
void gauss_grads(State<R>& st, const float* raw_f) {
  const int P = st.P, M = st.M;
  st.dmean2d.assign((size_t)P * 4, R(0));
  st.dconic.assign((size_t)P * 4, R(0));
  st.dopac.assign(P, R(0));
  st.dcolor.assign((size_t)P * 3, R(0));
  st.dmean3d.assign((size_t)P * 3, R(0));
  st.dcov3d.assign((size_t)P * 6, R(0));
  st.ddc.assign((size_t)P * 3, R(0));
  st.dsh.assign((size_t)P * M * 3, R(0));
  st.dscale.assign((size_t)P * 3, R(0));
  st.drot.assign((size_t)P * 4, R(0));
  const auto cam = st.cam();
  const auto G = st.gauss();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < P; ++i) {
    R rawr[fgs::RAW_STRIDE];
    bool any = false;
    for (int f = 0; f < fgs::RAW_N; ++f) {
      rawr[f] = raw_f ? (R)raw_f[(size_t)i * fgs::RAW_STRIDE + f]
                      : (R)st.raw[(size_t)i * fgs::RAW_N + f];
      any = any || rawr[f] != R(0);
    }
    if (!any) continue;
    fgs::Grad2D<R> g2;
    fgs::finalize_raw(rawr, st.opac[i], st.W, st.H, g2);
    if (rawr[fgs::RAW_POW] == R(0)) g2.opacity = R(0);
    for (int k = 0; k < 4; ++k) st.dmean2d[4 * (size_t)i + k] = g2.mean2d[k];
    for (int k = 0; k < 4; ++k) st.dconic[4 * (size_t)i + k] = g2.conic[k];
    st.dopac[i] = g2.opacity;
    for (int k = 0; k < 3; ++k) st.dcolor[3 * (size_t)i + k] = g2.color[k];
    if (!(st.radii[i] > 0)) continue;  // backward.cu:257, :501
    fgs::GaussGrad<R> out;
    out.dmean3d = &st.dmean3d[3 * (size_t)i];
    out.dcov3d = &st.dcov3d[6 * (size_t)i];
    out.dscale = &st.dscale[3 * (size_t)i];
    out.drot = &st.drot[4 * (size_t)i];
    out.ddc = &st.ddc[3 * (size_t)i];
    out.dsh = M > 0 ? &st.dsh[(size_t)i * M * 3] : nullptr;
    fgs::gauss_bwd(i, G, cam, st.pj[i].clamp, g2, out);
  }
}

template <typename R>
State<R>* create(int P, const R* means, const R* opac, const R* scales, const R* rots, const R* dc,
                 const R* shs, int M, const R* colors, const uint8_t* cull, const R* view,
                 const R* proj, const R* campos, const R* bg, double tanfovx, double tanfovy,
                 double scale_mod, double mult, int W, int H, int deg, int bucket, int mode,
                 int tile_w) {
  auto* st = new State<R>();
  st->P = P;
  st->W = W;
  st->H = H;
  st->tw = tile_w ? tile_w : fgs::TILE_W_REF;
  st->th = fgs::tile_h_of(st->tw);
  st->gx = fgs::grid_dim(W, st->tw);
  st->gy = fgs::grid_dim(H, st->th);
  st->T = st->gx * st->gy;
  st->B = bucket;
  st->mode = mode;
  st->deg = deg;
  st->M = M;
  // The camera scalars are float in the API (pybind.cpp:26-35); focal is
  // computed in float like preprocess.cu:941-942.
  const float tfx = (float)tanfovx, tfy = (float)tanfovy;
  st->tanfovx = (R)tfx;
  st->tanfovy = (R)tfy;
  st->focal_x = (R)((float)W / (2.0f * tfx));
  st->focal_y = (R)((float)H / (2.0f * tfy));
  st->scale_mod = (R)(float)scale_mod;
  st->mult = (R)(float)mult;
  copy_in(st->view, view, 16);
  copy_in(st->proj, proj, 16);
  copy_in(st->campos, campos, 3);
  copy_in(st->bg, bg, 3);
  copy_in(st->means, means, 3 * (size_t)P);
  copy_in(st->opac, opac, (size_t)P);
  copy_in(st->scales, scales, 3 * (size_t)P);
  copy_in(st->rots, rots, 4 * (size_t)P);
  st->has_colors = colors != nullptr;
  st->has_sh = !st->has_colors && shs != nullptr && dc != nullptr;
  if (st->has_sh) {
    copy_in(st->dc, dc, 3 * (size_t)P);
    copy_in(st->shs, shs, (size_t)P * M * 3);
  } else {
    st->M = 0;
  }
  if (st->has_colors) copy_in(st->colors, colors, 3 * (size_t)P);
  st->has_cull = cull != nullptr;
  if (st->has_cull) st->cull.assign(cull, cull + P);
  return st;
}

template <typename T>
long long vec_get(const std::vector<T>& v, void* dst) {
  if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(T));
  return (long long)v.size();
}

template <typename R>
long long get_named(State<R>* st, const char* name, void* dst) {
  const std::string n(name);
  if (n == "radii") return vec_get(st->radii, dst);
  if (n == "pair_count") return vec_get(st->pair_count, dst);
  if (n == "keys") return vec_get(st->keys, dst);
  if (n == "vals") return vec_get(st->vals, dst);
  if (n == "ranges") return vec_get(st->ranges, dst);
  if (n == "bucket_off") return vec_get(st->bucket_off, dst);
  if (n == "bucket_ranges") return vec_get(st->bucket_ranges, dst);
  if (n == "bucket_tile") return vec_get(st->bucket_tile, dst);
  if (n == "out") return vec_get(st->out, dst);
  if (n == "tfinal") return vec_get(st->tfinal, dst);
  if (n == "pixc") return vec_get(st->pixc, dst);
  if (n == "contrib") return vec_get(st->contrib, dst);
  if (n == "max_contrib") return vec_get(st->max_contrib, dst);
  if (n == "acc_t") return vec_get(st->acc_t, dst);
  if (n == "acc_c") return vec_get(st->acc_c, dst);
  if (n == "splat") return vec_get(st->splat, dst);
  if (n == "raw") return vec_get(st->raw, dst);
  if (n == "dmean2d") return vec_get(st->dmean2d, dst);
  if (n == "dconic") return vec_get(st->dconic, dst);
  if (n == "dopac") return vec_get(st->dopac, dst);
  if (n == "dcolor") return vec_get(st->dcolor, dst);
  if (n == "dmean3d") return vec_get(st->dmean3d, dst);
  if (n == "dcov3d") return vec_get(st->dcov3d, dst);
  if (n == "ddc") return vec_get(st->ddc, dst);
  if (n == "dsh") return vec_get(st->dsh, dst);
  if (n == "dscale") return vec_get(st->dscale, dst);
  if (n == "drot") return vec_get(st->drot, dst);
  if (n == "proj") {  // per Gaussian: valid, radius, rect[4], clamp as float bits + x,y,depth,conic3,power,log2op,rgb3
    const size_t P = st->pj.size();
    if (dst) {
      R* d = (R*)dst;
      for (size_t i = 0; i < P; ++i) {
        const auto& o = st->pj[i];
        R* r = d + 16 * i;
        r[0] = (R)o.valid;
        r[1] = (R)o.rect[0];
        r[2] = (R)o.rect[1];
        r[3] = (R)o.rect[2];
        r[4] = (R)o.rect[3];
        r[5] = (R)o.clamp;
        r[6] = o.valid ? o.x : R(0);
        r[7] = o.valid ? o.y : R(0);
        r[8] = o.valid ? o.depth : R(0);
        r[9] = o.valid ? o.conic[0] : R(0);
        r[10] = o.valid ? o.conic[1] : R(0);
        r[11] = o.valid ? o.conic[2] : R(0);
        r[12] = o.valid ? o.power : R(0);
        r[13] = o.valid ? o.rgb[0] : R(0);
        r[14] = o.valid ? o.rgb[1] : R(0);
        r[15] = o.valid ? o.rgb[2] : R(0);
      }
    }
    return (long long)(16 * P);
  }
  return -1;
}

// Importance accumulation of render_simp (render_simp.cu:36-301) over the
// sorted lists of a forward: per tile every pixel blends front to back
// without stopping on its own; after each pair of splats (a leading single
// splat when the list length is odd) the tile stops once all 256 of its
// pixels -- those outside the image keep T = 1 -- have T < 1e-4. Per splat:
// the sum of the weights alpha*T and the number of pixels with alpha >=
// 1/255; per pixel: +1 for the splat of largest weight (the first of equals).
template <typename R>
void render_simp(const State<R>* st, double* wsum, int32_t* wcount, double* maxcount) {
  const int P = st->P;
  std::fill(wsum, wsum + P, 0.0);
  std::fill(wcount, wcount + P, 0);
  std::fill(maxcount, maxcount + P, 0.0);
  for (int t = 0; t < st->T; ++t) {
    const int tx = t % st->gx, ty = t / st->gx;
    const uint32_t s = st->ranges[2 * t], e = st->ranges[2 * t + 1];
    R T[fgs::TILE_PIX], wmax[fgs::TILE_PIX];
    uint32_t imax[fgs::TILE_PIX];
    bool flag[fgs::TILE_PIX];
    for (int p = 0; p < fgs::TILE_PIX; ++p) {
      T[p] = R(1);
      wmax[p] = R(0);
      imax[p] = 0;
      flag[p] = false;
    }
    auto shade = [&](uint32_t k) {
      const uint32_t id = st->vals[k];
      const R* sp = &st->splat[9 * (size_t)id];
      for (int p = 0; p < fgs::TILE_PIX; ++p) {
        const int px = tx * st->tw + (p % st->tw), py = ty * st->th + p / st->tw;
        if (px >= st->W || py >= st->H) continue;
        const R a = fgs::splat_alpha(sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], (R)px, (R)py);
        if (a == R(0)) continue;  // alpha < 1/255
        const R w = a * T[p];
        T[p] = T[p] - w;
        wsum[id] += (double)w;
        wcount[id] += 1;
        if (wmax[p] < w) {
          wmax[p] = w;
          imax[p] = id;
          flag[p] = true;
        }
      }
    };
    uint32_t k = s;
    if ((e - s) % 2 == 1) shade(k++);
    for (; k < e; k += 2) {
      shade(k);
      shade(k + 1);
      bool done = true;
      for (int p = 0; p < fgs::TILE_PIX; ++p) done = done && T[p] < R(fgs::T_CUT_F);
      if (done) break;
    }
    for (int p = 0; p < fgs::TILE_PIX; ++p) {
      const int px = tx * st->tw + (p % st->tw), py = ty * st->th + p / st->tw;
      if (px < st->W && py < st->H && flag[p]) maxcount[imax[p]] += 1.0;
    }
  }
}

// Per-pixel depth of render_depth (render_depth.cu:92-290) over the sorted
// lists of a forward: each pixel's camera ray (through the inverse
// projection) walks its tile's list front to back until T would drop below
// 1e-4; among splats hit in front of the camera the one of largest blending
// weight gives the ray parameter of the midpoint of its 3-sigma ellipsoid
// (depth = t_mid / |ray|), the 3D point, the discriminant and its index.
// Outputs are written for every pixel inside the image.
template <typename R>
void render_depth(const State<R>* st, R* out_pts, R* out_depth, R* accum_alpha, int32_t* gidx,
                  R* disc) {
  const int W = st->W, H = st->H, P = st->P;
  const size_t HW = (size_t)W * H;
  R inv[16];
  fgs::inverse4(st->proj.data(), inv);
  const R* o = st->campos.data();
  std::vector<fgs::Ellipsoid<R>> ell(P);
  for (int i = 0; i < P; ++i)
    fgs::ellipsoid_setup(&st->means[3 * (size_t)i], &st->scales[3 * (size_t)i],
                         &st->rots[4 * (size_t)i], o, ell[i]);
  for (int t = 0; t < st->T; ++t) {
    const int tx = t % st->gx, ty = t / st->gx;
    const uint32_t s = st->ranges[2 * t], e = st->ranges[2 * t + 1];
    for (int p = 0; p < fgs::TILE_PIX; ++p) {
      const int px = tx * st->tw + (p % st->tw), py = ty * st->th + p / st->tw;
      if (px >= W || py >= H) continue;
      R nd[3], len;
      fgs::depth_ray(inv, o, px, py, W, H, nd, len);
      fgs::DepthPix<R> D;
      for (uint32_t k = s; k < e && !D.done; ++k) {
        const uint32_t id = st->vals[k];
        const R* sp = &st->splat[9 * (size_t)id];
        const R a = fgs::splat_alpha(sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], (R)px, (R)py);
        if (a == R(0)) continue;  // alpha < 1/255
        fgs::depth_pix_step(D, a, id, ell[id], o, nd, len);
      }
      const size_t pix = (size_t)py * W + px;
      for (int c = 0; c < 3; ++c) out_pts[c * HW + pix] = D.pt[c];
      out_depth[pix] = D.depth;
      accum_alpha[pix] = D.T;
      gidx[pix] = D.idx;
      disc[pix] = D.disc;
    }
  }
}

template <typename R>
void add_stats(State<R>* st, R* accum, R* accum_abs, R* max_radii, R* denom) {
  for (int i = 0; i < st->P; ++i) {
    if (!(st->radii[i] > 0)) continue;
    const R* g = &st->dmean2d[4 * (size_t)i];
    if (denom) denom[i] += R(1);
    if (max_radii) max_radii[i] = std::max(max_radii[i], (R)st->radii[i]);
    if (accum) accum[i] += fgs::sqrt_(g[0] * g[0] + g[1] * g[1]);
    if (accum_abs) accum_abs[i] += fgs::sqrt_(g[2] * g[2] + g[3] * g[3]);
  }
}

}  // namespace

extern "C" {

// float64 instances of render_depth's geometry, for pinning it on its own.
void orc_inverse4_f64(const double* m, double* inv) { fgs::inverse4(m, inv); }
void orc_depth_ray_f64(const double* proj, const double* campos, int px, int py, int W, int H,
                       double* out4) {
  double inv[16];
  fgs::inverse4(proj, inv);
  fgs::depth_ray(inv, campos, px, py, W, H, out4, out4[3]);
}
void orc_depth_ray_f32(const float* proj, const float* campos, int px, int py, int W, int H,
                       float* out4) {
  float inv[16];
  fgs::inverse4(proj, inv);
  fgs::depth_ray(inv, campos, px, py, W, H, out4, out4[3]);
}
void orc_ellipsoid_f64(const double* mean, const double* scale, const double* q,
                       const double* origin, const double* ndir, double* out2) {
  fgs::Ellipsoid<double> E;
  fgs::ellipsoid_setup(mean, scale, q, origin, E);
  fgs::ellipsoid_hit(E, ndir, out2[0], out2[1]);
}
void orc_render_depth_f32(void* h, float* out_pts, float* out_depth, float* accum_alpha,
                          int32_t* gidx, float* disc) {
  render_depth((const State<float>*)h, out_pts, out_depth, accum_alpha, gidx, disc);
}
void orc_render_simp_f32(void* h, double* wsum, int32_t* wcount, double* maxcount) {
  render_simp((const State<float>*)h, wsum, wcount, maxcount);
}


// ---- float32 oracle
void* orc_create_f32(int P, const float* means, const float* opac, const float* scales,
                     const float* rots, const float* dc, const float* shs, int M,
                     const float* colors, const uint8_t* cull, const float* view,
                     const float* proj, const float* campos, const float* bg, double tanfovx,
                     double tanfovy, double scale_mod, double mult, int W, int H, int deg,
                     int bucket, int mode, int tile_w) {
  return create<float>(P, means, opac, scales, rots, dc, shs, M, colors, cull, view, proj, campos,
                       bg, tanfovx, tanfovy, scale_mod, mult, W, H, deg, bucket, mode, tile_w);
}
void orc_forward_f32(void* h) { forward(*(State<float>*)h); }
void orc_backward_f32(void* h, const float* dimg) { backward(*(State<float>*)h, dimg); }
void orc_gauss_grads_from_raw_f32(void* h, const float* raw12) {
  gauss_grads(*(State<float>*)h, raw12);
}
long long orc_get_f32(void* h, const char* name, void* dst) {
  return get_named((State<float>*)h, name, dst);
}
void orc_stats_f32(void* h, float* accum, float* accum_abs, float* max_radii, float* denom) {
  add_stats((State<float>*)h, accum, accum_abs, max_radii, denom);
}
void orc_destroy_f32(void* h) { delete (State<float>*)h; }

// ---- float64 oracle (finite-difference checks)
void* orc_create_f64(int P, const double* means, const double* opac, const double* scales,
                     const double* rots, const double* dc, const double* shs, int M,
                     const double* colors, const uint8_t* cull, const double* view,
                     const double* proj, const double* campos, const double* bg, double tanfovx,
                     double tanfovy, double scale_mod, double mult, int W, int H, int deg,
                     int bucket, int mode, int tile_w) {
  return create<double>(P, means, opac, scales, rots, dc, shs, M, colors, cull, view, proj,
                        campos, bg, tanfovx, tanfovy, scale_mod, mult, W, H, deg, bucket, mode,
                        tile_w);
}
void orc_forward_f64(void* h) { forward(*(State<double>*)h); }
void orc_backward_f64(void* h, const double* dimg) { backward(*(State<double>*)h, dimg); }
long long orc_get_f64(void* h, const char* name, void* dst) {
  return get_named((State<double>*)h, name, dst);
}
void orc_destroy_f64(void* h) { delete (State<double>*)h; }

// scalar helpers exposed for unit tests of the shared arithmetic
float orc_exp2f(float x) { return fgs::exp2_(x); }
float orc_log2f(float x) { return fgs::log2_(x); }
int orc_tile_hit_f32(int tx, int ty, float x, float y, float ca, float cb, float cc, float power,
                     int tw) {
  const float conic[3] = {ca, cb, cc};
  return fgs::tile_hit(tx, ty, tw, fgs::tile_h_of(tw), x, y, conic, power) ? 1 : 0;
}
// Checks tile_hit's factored form (tile_col / tile_row / tile_hit_terms,
// used by the GPU's walks over large rects) against tile_hit on every tile
// of a gx x gy grid for one ellipse; returns the number of disagreements
// and the number of hits.
int orc_tile_terms_check(float x, float y, float ca, float cb, float cc, float power, int gx,
                         int gy, int* hits, int tw) {
  const float conic[3] = {ca, cb, cc};
  const int th = fgs::tile_h_of(tw);
  int bad = 0, h = 0;
  for (int ty = 0; ty < gy; ++ty) {
    const fgs::TileRow r = fgs::tile_row(ty, th, y, conic, power);
    for (int tx = 0; tx < gx; ++tx) {
      const bool a = fgs::tile_hit(tx, ty, tw, th, x, y, conic, power);
      const bool b = fgs::tile_hit_terms(fgs::tile_col(tx, tw, x, conic, power), r, conic);
      bad += a != b;
      h += a;
    }
  }
  *hits = h;
  return bad;
}
int orc_num_threads() { return omp_get_max_threads(); }

}  // extern "C"
