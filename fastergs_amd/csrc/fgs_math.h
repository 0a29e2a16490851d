// Single-source per-Gaussian and per-pixel arithmetic of the rasterizer.
//
// This header is compiled twice: by hipcc into the gfx950 kernels and by g++
// into the CPU oracle. Both builds use -ffp-contract=off, IEEE-correct
// division and square root, and the project's own exp2/log2 below, so every
// decision (culling, rectangle, tile test, depth key, alpha threshold, bucket
// cutoff) comes out bit-identical on CPU and GPU. Fused multiply-adds appear
// only where written explicitly with fma_().
//
// Everything is templated on the real type R: R=float is the product path,
// R=double is the FP64 oracle used for finite-difference checks (it uses the
// C library exp2/log2 instead of the polynomials).
//
// Reference semantics restated here (paths relative to the FasterGS repo,
// FGS = gs_submodules/diff-gaussian-rasterization_faster):
//   projection / culling / radius / power / rect : FGS/csrc/cuda_rasterizer/preprocess.cu:450-499
//   3D covariance from unnormalised quaternion   : preprocess.cu:364-400
//   EWA 2D covariance with 1.3 tanfov clamp      : preprocess.cu:79-117
//   SH degree<=3 to RGB with clamping            : preprocess.cu:119-166, constants FGS/csrc/ops.h:16-32
//   exact tile / ellipse test                    : preprocess.cu:168-218
//   per-pixel alpha and front-to-back blend      : FGS/csrc/cuda_rasterizer/render.cu:99-123
//   bucket merge (T,C) (+) (Tb,Cb)               : render.cu:329-362
//   per-pixel reverse step of the backward       : FGS/csrc/cuda_rasterizer/backward.cu:549-632
//   cov2D / projection / SH / cov3D adjoints     : backward.cu:121-245, 250-403, 407-483, 489-539
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FGS_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define FGS_HD inline
#endif

namespace fgs {

// ---------------------------------------------------------------- constants
// Tiles hold TILE_PIX = 256 pixels in one of three shapes, tile_w x
// (256 / tile_w) with tile_w 8, 16 or 32; 16 x 16 is the reference's
// (BLOCK_X x BLOCK_Y, render_16x16, __init__.py:109). The shape is a
// property of the camera (CamRef) and of each view record.
constexpr int TILE_PIX = 256;
constexpr int TILE_W_REF = 16;
FGS_HD bool tile_w_ok(int tw) { return tw == 8 || tw == 16 || tw == 32; }
FGS_HD int tile_h_of(int tw) { return TILE_PIX / tw; }
FGS_HD int tile_w_log2(int tw) { return tw == 8 ? 3 : (tw == 32 ? 5 : 4); }
FGS_HD int grid_dim(int n, int tile) { return (n + tile - 1) / tile; }
constexpr float LOG2E_F = 1.4426950216293334961f;   // rasterizer_imp.h:43
constexpr float LN2_F = 0.69314718055f;             // rasterizer_imp.h:44
constexpr float ALPHA_CAP_F = 0.99f;                // render.cu:109
constexpr float T_CUT_F = 0.0001f;                  // render.cu:355
constexpr float NEAR_Z_F = 0.2f;                    // preprocess.cu:456
constexpr float LOWPASS_F = 0.3f;                   // preprocess.cu:114-115
constexpr float W_EPS_F = 0.0000001f;               // preprocess.cu:464
constexpr float EXTENT_CAP_F = 3.33f;               // preprocess.cu:483
constexpr float EIG_EPS_F = 0.01f;                  // preprocess.cu:484
constexpr float SQRT2_F = 1.414214f;                // preprocess.cu:491

// Real spherical-harmonics constants (FGS/csrc/ops.h:16-32).
constexpr float SH_C0 = 0.28209479177387814f;
constexpr float SH_C1 = 0.4886025119029199f;
constexpr float SH_C2_0 = 1.0925484305920792f;
constexpr float SH_C2_1 = -1.0925484305920792f;
constexpr float SH_C2_2 = 0.31539156525252005f;
constexpr float SH_C2_3 = -1.0925484305920792f;
constexpr float SH_C2_4 = 0.5462742152960396f;
constexpr float SH_C3_0 = -0.5900435899266435f;
constexpr float SH_C3_1 = 2.890611442640554f;
constexpr float SH_C3_2 = -0.4570457994644658f;
constexpr float SH_C3_3 = 0.3731763325901154f;
constexpr float SH_C3_4 = -0.4570457994644658f;
constexpr float SH_C3_5 = 1.445305721320277f;
constexpr float SH_C3_6 = -0.5900435899266435f;

// ------------------------------------------------------ scalar primitives
FGS_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
FGS_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

FGS_HD float fma_(float a, float b, float c) { return fmaf(a, b, c); }
FGS_HD double fma_(double a, double b, double c) { return fma(a, b, c); }
FGS_HD float sqrt_(float a) { return sqrtf(a); }
FGS_HD double sqrt_(double a) { return sqrt(a); }
FGS_HD float min_(float a, float b) { return fminf(a, b); }
FGS_HD double min_(double a, double b) { return fmin(a, b); }
FGS_HD float max_(float a, float b) { return fmaxf(a, b); }
FGS_HD double max_(double a, double b) { return fmax(a, b); }
FGS_HD float abs_(float a) { return fabsf(a); }
FGS_HD double abs_(double a) { return fabs(a); }
FGS_HD float ceil_(float a) { return ceilf(a); }
FGS_HD double ceil_(double a) { return ceil(a); }

// float -> int truncation with the argument clamped to +-1e9 first, so the
// conversion is defined (and identical on x86 and gfx950) for any input; NaN
// maps to -1e9 because fmaxf returns the non-NaN operand.
FGS_HD int trunc_i(float v) { return (int)fminf(fmaxf(v, -1e9f), 1e9f); }
FGS_HD int trunc_i(double v) { return (int)fmin(fmax(v, -1e9), 1e9); }

// Exact scaling by 2^e for a normal result (v_ldexp_f32 on the GPU).
#if defined(__HIP_DEVICE_COMPILE__)
FGS_HD float ldexp_(float v, int e) { return __builtin_amdgcn_ldexpf(v, e); }
#else
FGS_HD float ldexp_(float v, int e) { return ldexpf(v, e); }
#endif

// 2^x for float. Method: Cody-Waite style reduction x = n + f with
// n = rint(x), f in [-1/2, 1/2] (the subtraction is exact), then a degree-6
// polynomial for 2^f whose coefficients were fitted by iteratively
// re-weighted least squares at Chebyshev nodes (Lawson's algorithm, which
// converges to the minimax fit), evaluated by Horner's rule with fused
// multiply-adds, and an exact scaling by 2^n. Maximum error on the reduced
// range is below 1 ulp. exp2_core needs x in [-125, 127]; exp2_ handles the
// rest (flush to 0 below 2^-125, inf above 2^127, NaN propagates).
FGS_HD float exp2_core(float x) {
  const float n = rintf(x);
  const float f = x - n;
  float p = 1.534580806037411e-4f;
  p = fma_(p, f, 1.3399930903688073e-3f);
  p = fma_(p, f, 9.618489071726799e-3f);
  p = fma_(p, f, 5.550328642129898e-2f);
  p = fma_(p, f, 2.4022646248340607e-1f);
  p = fma_(p, f, 6.931471824645996e-1f);
  p = fma_(p, f, 1.0f);
  return ldexp_(p, (int)n);
}
FGS_HD float exp2_(float x) {
  if (!(x >= -125.0f)) return (x != x) ? x : 0.0f;
  if (x > 127.0f) return u2f(0x7f800000u);
  return exp2_core(x);
}
FGS_HD double exp2_(double x) { return exp2(x); }

// log2(x) for float. Method: x = m * 2^e with m in [sqrt(1/2), sqrt(2)],
// s = (m-1)/(m+1) (correctly rounded division), log2(m) = s * P(s^2) where
// P approximates 2/ln2 * atanh(s)/s (degree 3 in s^2, fitted like exp2_).
// Error <= 3 ulp. Non-positive input gives -inf / NaN.
FGS_HD float log2_(float x) {
  if (!(x > 0.0f)) return (x == 0.0f) ? u2f(0xff800000u) : u2f(0x7fc00000u);
  if (x == u2f(0x7f800000u)) return x;
  uint32_t b = f2u(x);
  int e = 0;
  if (b < 0x00800000u) {  // subnormal: scale into the normal range
    b = f2u(x * 8388608.0f);
    e = -23;
  }
  e += (int)(b >> 23) - 127;
  float m = u2f((b & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  float p = 0.43174833059310913f;
  p = fma_(p, s2, 0.5767138600349426f);
  p = fma_(p, s2, 0.9617988467216492f);
  p = fma_(p, s2, 2.885390043258667f);
  return fma_(s, p, (float)e);
}
FGS_HD double log2_(double x) { return log2(x); }

// ------------------------------------------------------------ camera view
// Matrices are 4x4 column-major exactly as the reference kernels read them
// (FGS/csrc/cuda_rasterizer/rasterizer_imp.h:502-520): out_r = sum_c m[4c+r] p_c.
template <typename R>
struct CamRef {
  const R* view;    // world->view, 16
  const R* proj;    // world->clip, 16
  const R* campos;  // 3
  R tanfovx, tanfovy, focal_x, focal_y, scale_mod, mult;
  int W, H, grid_x, grid_y, deg, max_coeffs;
  int tile_w, tile_h;  // tile shape in pixels (tile_w * tile_h = TILE_PIX)
};

template <typename R>
struct GaussRef {
  const R* means;      // P x 3
  const R* opac;       // P
  const R* scales;     // P x 3
  const R* rots;       // P x 4, (r, x, y, z), not normalised
  const R* dc;         // P x 3 (may be null when colors is given)
  const R* shs;        // P x max_coeffs x 3 (may be null)
  const R* colors;     // P x 3 precomputed RGB (null -> SH)
  const uint8_t* culling;  // P (null -> none culled)
};

// Result of preprocessing one Gaussian.
template <typename R>
struct Proj {
  int valid;        // passed both culls and has a non-empty tile rect
  int radius;       // radii output (0 when not valid)
  int rect[4];      // tile rect: min x, min y, max x, max y (exclusive)
  int clamp;        // bit k set when colour channel k was clamped at 0
  R x, y;           // pixel-space mean
  R depth;          // view-space z (its float bits are the depth key)
  R conic[3];       // inverse 2D covariance (a, b, c)
  R power;          // (8 ln2 + ln2 log2 opacity) * mult
  R log2op;         // log2(opacity)
  R rgb[3];
};

template <typename R>
FGS_HD void xform_point(const R* m, R x, R y, R z, R out[3]) {
  out[0] = m[0] * x + m[4] * y + m[8] * z + m[12];
  out[1] = m[1] * x + m[5] * y + m[9] * z + m[13];
  out[2] = m[2] * x + m[6] * y + m[10] * z + m[14];
}
template <typename R>
FGS_HD R xform_w(const R* m, R x, R y, R z) {
  return m[3] * x + m[7] * y + m[11] * z + m[15];
}

// ndc -> pixel, evaluated in double like the reference's double literals
// (rasterizer_imp.h:493-495), then rounded to R.
template <typename R>
FGS_HD R ndc2pix(R v, int S) {
  return (R)((((double)v + 1.0) * (double)S - 1.0) * 0.5);
}

// Standard rotation matrix of the (unnormalised) quaternion (r,x,y,z).
// The reference builds its transpose in glm and uses M = S R_glm,
// Sigma = M^T M (preprocess.cu:364-400); that equals Rq S^2 Rq^T used here.
template <typename R>
FGS_HD void quat_rot(const R q[4], R Rm[9]) {
  const R r = q[0], x = q[1], y = q[2], z = q[3];
  Rm[0] = R(1) - R(2) * (y * y + z * z);
  Rm[1] = R(2) * (x * y - r * z);
  Rm[2] = R(2) * (x * z + r * y);
  Rm[3] = R(2) * (x * y + r * z);
  Rm[4] = R(1) - R(2) * (x * x + z * z);
  Rm[5] = R(2) * (y * z - r * x);
  Rm[6] = R(2) * (x * z - r * y);
  Rm[7] = R(2) * (y * z + r * x);
  Rm[8] = R(1) - R(2) * (x * x + y * y);
}

// Sigma = Rq diag(s)^2 Rq^T, upper triangle (00,01,02,11,12,22).
template <typename R>
FGS_HD void cov3d_of(const R s[3], const R q[4], R cov[6]) {
  R Rm[9];
  quat_rot(q, Rm);
  R M[9];  // M = Rq diag(s)
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M[3 * i + j] = Rm[3 * i + j] * s[j];
  cov[0] = M[0] * M[0] + M[1] * M[1] + M[2] * M[2];
  cov[1] = M[0] * M[3] + M[1] * M[4] + M[2] * M[5];
  cov[2] = M[0] * M[6] + M[1] * M[7] + M[2] * M[8];
  cov[3] = M[3] * M[3] + M[4] * M[4] + M[5] * M[5];
  cov[4] = M[3] * M[6] + M[4] * M[7] + M[5] * M[8];
  cov[5] = M[6] * M[6] + M[7] * M[7] + M[8] * M[8];
}

// EWA Jacobian-times-view T (2x3) at view-space point t (t is clamped in x/y
// to 1.3 tanfov like preprocess.cu:88-95). Returns the clamped tx, ty.
template <typename R>
FGS_HD void ewa_T(const CamRef<R>& cam, const R t_in[3], R T[6], R& tx, R& ty,
                  R& xmul, R& ymul) {
  const R tz = t_in[2];
  const R limx = R(1.3f) * cam.tanfovx;
  const R limy = R(1.3f) * cam.tanfovy;
  const R txtz = t_in[0] / tz;
  const R tytz = t_in[1] / tz;
  tx = min_(limx, max_(-limx, txtz)) * tz;
  ty = min_(limy, max_(-limy, tytz)) * tz;
  xmul = (txtz < -limx || txtz > limx) ? R(0) : R(1);
  ymul = (tytz < -limy || tytz > limy) ? R(0) : R(1);
  const R J00 = cam.focal_x / tz;
  const R J02 = -(cam.focal_x * tx) / (tz * tz);
  const R J11 = cam.focal_y / tz;
  const R J12 = -(cam.focal_y * ty) / (tz * tz);
  const R* v = cam.view;  // V(r,c) = v[4c+r]
  for (int c = 0; c < 3; ++c) {
    T[c] = J00 * v[4 * c + 0] + J02 * v[4 * c + 2];
    T[3 + c] = J11 * v[4 * c + 1] + J12 * v[4 * c + 2];
  }
}

// 2D covariance (a, b, c) = T Sigma T^T plus the 0.3 low-pass on the diagonal.
template <typename R>
FGS_HD void cov2d_of(const R T[6], const R cov[6], R& a, R& b, R& c) {
  const R S00 = cov[0], S01 = cov[1], S02 = cov[2], S11 = cov[3], S12 = cov[4], S22 = cov[5];
  const R u0x = S00 * T[0] + S01 * T[1] + S02 * T[2];
  const R u0y = S01 * T[0] + S11 * T[1] + S12 * T[2];
  const R u0z = S02 * T[0] + S12 * T[1] + S22 * T[2];
  const R u1x = S00 * T[3] + S01 * T[4] + S02 * T[5];
  const R u1y = S01 * T[3] + S11 * T[4] + S12 * T[5];
  const R u1z = S02 * T[3] + S12 * T[4] + S22 * T[5];
  a = T[0] * u0x + T[1] * u0y + T[2] * u0z + R(LOWPASS_F);
  b = T[0] * u1x + T[1] * u1y + T[2] * u1z;
  c = T[3] * u1x + T[4] * u1y + T[5] * u1z + R(LOWPASS_F);
}

// Real SH basis (degree <= 3) at unit direction (x,y,z); basis[0] is the DC
// term. Order and signs follow preprocess.cu:132-155.
template <typename R>
FGS_HD void sh_basis(R x, R y, R z, R bas[16]) {
  const R xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
  bas[0] = R(SH_C0);
  bas[1] = -R(SH_C1) * y;
  bas[2] = R(SH_C1) * z;
  bas[3] = -R(SH_C1) * x;
  bas[4] = R(SH_C2_0) * xy;
  bas[5] = R(SH_C2_1) * yz;
  bas[6] = R(SH_C2_2) * (R(2) * zz - xx - yy);
  bas[7] = R(SH_C2_3) * xz;
  bas[8] = R(SH_C2_4) * (xx - yy);
  bas[9] = R(SH_C3_0) * y * (R(3) * xx - yy);
  bas[10] = R(SH_C3_1) * xy * z;
  bas[11] = R(SH_C3_2) * y * (R(4) * zz - xx - yy);
  bas[12] = R(SH_C3_3) * z * (R(2) * zz - R(3) * xx - R(3) * yy);
  bas[13] = R(SH_C3_4) * x * (R(4) * zz - xx - yy);
  bas[14] = R(SH_C3_5) * z * (xx - yy);
  bas[15] = R(SH_C3_6) * x * (xx - R(3) * yy);
}

FGS_HD int sh_count(int deg) { return deg <= 0 ? 1 : (deg == 1 ? 4 : (deg == 2 ? 9 : 16)); }

// View-dependent colour: sum_k basis_k coef_k + 0.5, clamped at 0. d0 is the
// Gaussian's DC row (3), sh its rest rows (3 per coefficient); loops have
// static bounds so the arrays stay in registers.
template <typename R>
FGS_HD void sh_color_rows(R ox, R oy, R oz, const R* d0, const R* sh, int deg, R rgb[3],
                          int& clamp) {
  const R len = sqrt_(ox * ox + oy * oy + oz * oz);
  R bas[16];
  sh_basis(ox / len, oy / len, oz / len, bas);
  const int n = sh_count(deg);
  clamp = 0;
  for (int ch = 0; ch < 3; ++ch) {
    R acc = bas[0] * d0[ch];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int k = 1; k < 16; ++k)
      if (k < n) acc = acc + bas[k] * sh[3 * (k - 1) + ch];
    acc = acc + R(0.5f);
    if (acc < R(0)) clamp |= 1 << ch;
    rgb[ch] = max_(acc, R(0));
  }
}

template <typename R>
FGS_HD void sh_color(int i, const GaussRef<R>& G, const CamRef<R>& cam, R rgb[3], int& clamp) {
  const R* m = G.means + 3 * i;
  sh_color_rows(m[0] - cam.campos[0], m[1] - cam.campos[1], m[2] - cam.campos[2], G.dc + 3 * i,
                G.shs + (size_t)i * cam.max_coeffs * 3, cam.deg, rgb, clamp);
}

// Tile rect of a splat centred at (x,y) with integer half-extent (w,h)
// (rasterizer_imp.h:532-541) on tw x th tiles.
template <typename R>
FGS_HD void tile_rect(R x, R y, int w, int h, int gx, int gy, int tw, int th, int rect[4]) {
  const R bx = R(tw), by = R(th);
  int v;
  v = trunc_i((x - (R)w) / bx); rect[0] = v < 0 ? 0 : (v > gx ? gx : v);
  v = trunc_i((y - (R)h) / by); rect[1] = v < 0 ? 0 : (v > gy ? gy : v);
  v = trunc_i((x + (R)w) / bx) + 1; rect[2] = v < 0 ? 0 : (v > gx ? gx : v);
  v = trunc_i((y + (R)h) / by) + 1; rect[3] = v < 0 ? 0 : (v > gy ? gy : v);
}

// Full per-Gaussian forward (preprocess.cu:450-499 and :698-720). With
// color == false the colour is left at zero (a separate pass evaluates it).
template <typename R>
FGS_HD void preprocess_gauss(int i, const GaussRef<R>& G, const CamRef<R>& cam, Proj<R>& o,
                             const R* dc_row = nullptr, const R* sh_row = nullptr,
                             bool color = true) {
  o.valid = 0;
  o.radius = 0;
  o.clamp = 0;
  o.rect[0] = o.rect[1] = o.rect[2] = o.rect[3] = 0;
  if (G.culling != nullptr && G.culling[i]) return;
  const R* pm = G.means + 3 * i;
  R t[3];
  xform_point(cam.view, pm[0], pm[1], pm[2], t);
  if (t[2] <= R(NEAR_Z_F)) return;
  const R op = G.opac[i];
  if (R(255) * op < R(1)) return;
  R hom[3];
  xform_point(cam.proj, pm[0], pm[1], pm[2], hom);
  const R pw = R(1) / (xform_w(cam.proj, pm[0], pm[1], pm[2]) + R(W_EPS_F));
  const R ndx = hom[0] * pw, ndy = hom[1] * pw;

  R s[3], q[4], cov[6];
  for (int k = 0; k < 3; ++k) s[k] = cam.scale_mod * G.scales[3 * i + k];
  for (int k = 0; k < 4; ++k) q[k] = G.rots[4 * i + k];
  cov3d_of(s, q, cov);
  R T[6], tx, ty, xm, ym, a, b, c;
  ewa_T(cam, t, T, tx, ty, xm, ym);
  cov2d_of(T, cov, a, b, c);

  const R det = a * c - b * b;
  const R det_inv = R(1) / det;
  o.conic[0] = c * det_inv;
  o.conic[1] = -b * det_inv;
  o.conic[2] = a * det_inv;

  const R mid = R(0.5f) * (a + c);
  const R ln255op = log2_(op * R(255)) * R(LN2_F);
  const R extent = min_(R(EXTENT_CAP_F), sqrt_(R(2) * ln255op));
  const R lambda = mid + sqrt_(max_(R(EIG_EPS_F), mid * mid - det));
  const R radius_f = ceil_(extent * sqrt_(lambda));
  if (radius_f < R(0)) return;

  const R l2op = log2_(op);
  R power = R(LN2_F) * R(8) + R(LN2_F) * l2op;
  power = power * cam.mult;
  const int hw = trunc_i(R(SQRT2_F) * sqrt_(a * power) + R(1));
  const int hh = trunc_i(R(SQRT2_F) * sqrt_(c * power) + R(1));
  const R px = ndc2pix(ndx, cam.W), py = ndc2pix(ndy, cam.H);
  tile_rect(px, py, hw, hh, cam.grid_x, cam.grid_y, cam.tile_w, cam.tile_h, o.rect);
  const int tc = (o.rect[2] - o.rect[0]) * (o.rect[3] - o.rect[1]);
  if (tc <= 0) return;

  o.valid = 1;
  o.radius = (radius_f == radius_f) ? trunc_i(radius_f) : 0;
  o.x = px;
  o.y = py;
  o.depth = t[2];
  o.power = power;
  o.log2op = l2op;
  if (!color) {
    o.rgb[0] = o.rgb[1] = o.rgb[2] = R(0);
  } else if (G.colors != nullptr) {
    for (int k = 0; k < 3; ++k) o.rgb[k] = G.colors[3 * i + k];
  } else if (dc_row != nullptr) {
    sh_color_rows(pm[0] - cam.campos[0], pm[1] - cam.campos[1], pm[2] - cam.campos[2], dc_row,
                  sh_row, cam.deg, o.rgb, o.clamp);
  } else {
    sh_color(i, G, cam, o.rgb, o.clamp);
  }
}

// Does the ellipse {d^T conic d <= 2 power} meet one edge segment?
// (preprocess.cu:168-177)
template <typename R>
FGS_HD bool seg_hits(R a, R b, R c, R d, R l, R r) {
  const R delta = b * b - R(4) * a * c;
  const R t1 = (l - d) * (R(2) * a) + b;
  const R t2 = (r - d) * (R(2) * a) + b;
  // every term evaluated (no short circuit): the same value, branch-free on the GPU
  return (delta >= R(0)) & ((t1 <= R(0)) | (t1 * t1 <= delta)) & ((t2 >= R(0)) | (t2 * t2 <= delta));
}

// Exact tile test of preprocess.cu:179-218 for tile (tx,ty) of tw x th
// tiles: true when the pixel-centre box of the tile contains the centre or
// the ellipse meets the nearer vertical edge or the nearer horizontal edge.
template <typename R>
FGS_HD bool tile_hit(int tx, int ty, int tw, int th, R x, R y, const R conic[3], R power) {
  const int x0 = tx * tw, y0 = ty * th;
  const int x1 = x0 + tw - 1, y1 = y0 + th - 1;
  const bool inside = (x >= (R)x0) & (x <= (R)x1) & (y >= (R)y0) & (y <= (R)y1);
  const R w = R(2) * power;
  const R dx = (x * R(2) < (R)(x0 + x1)) ? x - (R)x0 : x - (R)x1;
  const bool h1 = seg_hits(conic[2], R(-2) * conic[1] * dx, conic[0] * dx * dx - w, y, (R)y0, (R)y1);
  const R dy = (y * R(2) < (R)(y0 + y1)) ? y - (R)y0 : y - (R)y1;
  const bool h2 = seg_hits(conic[0], R(-2) * conic[1] * dy, conic[2] * dy * dy - w, x, (R)x0, (R)x1);
  return inside | h1 | h2;
}

// tile_hit factored for walks over many tiles of one ellipse: the terms
// that depend on the tile column only (TileCol) or on the tile row only
// (TileRow) are formed once, and tile_hit_terms finishes the test with the
// remaining per-tile operations. Every float operation has the operands and
// order of tile_hit, so the result is identical.
struct TileCol {
  float x0f, x1f;  // pixel-centre box columns
  bool xin;        // x0 <= x <= x1
  float b1, d1;    // nearer vertical edge: b and discriminant of seg_hits
  float l2, r2;    // x0 - x, x1 - x (horizontal-edge segment ends)
};
struct TileRow {
  float l1, r1;    // y0 - y, y1 - y (vertical-edge segment ends)
  bool yin;        // y0 <= y <= y1
  float b2, d2;    // nearer horizontal edge: b and discriminant of seg_hits
};
FGS_HD TileCol tile_col(int tx, int tw, float x, const float conic[3], float power) {
  TileCol c;
  const int x0 = tx * tw, x1 = x0 + tw - 1;
  c.x0f = (float)x0;
  c.x1f = (float)x1;
  c.xin = (x >= c.x0f) & (x <= c.x1f);
  const float w = 2.0f * power;
  const float dx = (x * 2.0f < (float)(x0 + x1)) ? x - c.x0f : x - c.x1f;
  c.b1 = -2.0f * conic[1] * dx;
  const float cc = conic[0] * dx * dx - w;
  c.d1 = c.b1 * c.b1 - 4.0f * conic[2] * cc;
  c.l2 = c.x0f - x;
  c.r2 = c.x1f - x;
  return c;
}
FGS_HD TileRow tile_row(int ty, int th, float y, const float conic[3], float power) {
  TileRow r;
  const int y0 = ty * th, y1 = y0 + th - 1;
  const float y0f = (float)y0, y1f = (float)y1;
  r.yin = (y >= y0f) & (y <= y1f);
  const float w = 2.0f * power;
  const float dy = (y * 2.0f < (float)(y0 + y1)) ? y - y0f : y - y1f;
  r.b2 = -2.0f * conic[1] * dy;
  const float cc = conic[2] * dy * dy - w;
  r.d2 = r.b2 * r.b2 - 4.0f * conic[0] * cc;
  r.l1 = y0f - y;
  r.r1 = y1f - y;
  return r;
}
// The end tests of seg_hits for a segment [l, r] relative to the centre.
FGS_HD bool seg_ends(float two_a, float b, float delta, float l, float r) {
  const float t1 = l * two_a + b;
  const float t2 = r * two_a + b;
  return (delta >= 0.0f) & ((t1 <= 0.0f) | (t1 * t1 <= delta)) & ((t2 >= 0.0f) | (t2 * t2 <= delta));
}
FGS_HD bool tile_hit_terms(const TileCol& c, const TileRow& r, const float conic[3]) {
  const bool h1 = seg_ends(2.0f * conic[2], c.b1, c.d1, r.l1, r.r1);
  const bool h2 = seg_ends(2.0f * conic[0], r.b2, r.d2, c.l2, c.r2);
  return (c.xin & r.yin) | h1 | h2;
}

// ------------------------------------------------------------ per pixel
// Log2-domain conic coefficients stored per splat (preprocess.cu:709-710):
// ca = -1/2 log2e a, cb = -log2e b, cc = -1/2 log2e c, w = log2(opacity).
template <typename R>
FGS_HD void log2_conic(const R conic[3], R log2op, R& ca, R& cb, R& cc, R& w) {
  ca = R(-0.5f * LOG2E_F) * conic[0];
  cb = R(-LOG2E_F) * conic[1];
  cc = R(-0.5f * LOG2E_F) * conic[2];
  w = log2op;
}

// Alpha of a splat at integer pixel (px,py) (render.cu:99-110).
// The exponent is clamped to [-9, 0] before exp2: every p > 0 gives alpha =
// 0.99 and every p < -9 gives alpha < 1/255 -> 0 either way, so the result
// equals min(0.99, 2^p) thresholded at 1/255 for all p (NaN included, which
// becomes 0.99 like fminf(0.99, NaN)), and exp2 needs no range checks.
// alpha_capped is min(0.99, 2^p) before the threshold (never NaN).
FGS_HD float alpha_capped(float p) {
  const float pc = fmaxf(fminf(p, 0.0f), -9.0f);
  return fminf(ALPHA_CAP_F, exp2_core(pc));
}
FGS_HD float alpha_of(float p) {
  const float a = alpha_capped(p);
  return a < 1.0f / 255.0f ? 0.0f : a;
}
FGS_HD double alpha_of(double p) {
  const double a = fmin(0.99, exp2(p));
  return a < 1.0 / 255.0 ? 0.0 : a;
}

// Exponent of a splat at pixel (px,py): log2 opacity plus the log2-domain
// quadratic form of the offset.
template <typename R>
FGS_HD R splat_power(R gx, R gy, R ca, R cb, R cc, R w, R px, R py) {
  const R dx = gx - px, dy = gy - py;
  return fma_(cb * dx, dy, fma_(cc * dy, dy, fma_(ca * dx, dx, w)));
}
template <typename R>
FGS_HD R splat_alpha(R gx, R gy, R ca, R cb, R cc, R w, R px, R py) {
  return alpha_of(splat_power(gx, gy, ca, cb, cc, w, px, py));
}

// Front-to-back blend of one splat (render.cu:112-115).
template <typename R>
FGS_HD void blend(R a, R r, R g, R b, R& T, R& C0, R& C1, R& C2) {
  const R wgt = a * T;
  C0 = fma_(r, wgt, C0);
  C1 = fma_(g, wgt, C1);
  C2 = fma_(b, wgt, C2);
  T = T - wgt;
}

// Associative merge (T,C) <- (T,C) (+) (Tb,Cb) = (T Tb, C + T Cb)
// (render.cu:338-342).
template <typename R>
FGS_HD void merge(R Tb, R Cb0, R Cb1, R Cb2, R& T, R& C0, R& C1, R& C2) {
  C0 = fma_(T, Cb0, C0);
  C1 = fma_(T, Cb1, C1);
  C2 = fma_(T, Cb2, C2);
  T = T * Tb;
}

// Raw per-(pixel, splat) gradient sums. The constant factors (W/2, H/2,
// -1/2 log2e, log2e/opacity) are applied once per Gaussian afterwards.
enum RawField {
  RAW_COL0 = 0, RAW_COL1, RAW_COL2,   // sum (alpha T) dL/dpix_k
  RAW_MX, RAW_MY,                     // sum dL/dp * dp/d(dx), dp/d(dy)
  RAW_KXX, RAW_KXY, RAW_KYY,          // sum dL/dp * dx^2, dx dy, dy^2
  RAW_POW,                            // sum dL/dp
  RAW_AX, RAW_AY,                     // sum |dL/dp dp/d(dx)|, |dL/dp dp/d(dy)|
  RAW_N = 11
};
constexpr int RAW_STRIDE = 12;
// Device layout of the raw sums, two arrays of rows per record: the conic
// sums RAW_KXX..RAW_KYY in double (padded to 4), summed order-independently
// because the conic -> 2D covariance adjoint amplifies their rounding, and
// the other eight fields in float, in RawField order without the conic ones.
constexpr int RAW_CON_STRIDE = 4;
constexpr int RAW_LIN_STRIDE = 8;
FGS_HD constexpr int raw_lin_index(int f) { return f < RAW_KXX ? f : f - 3; }

// One reverse step of the backward for one pixel (backward.cu:549-632):
// recovers T before the splat and accumulates the raw sums into acc[RAW_N].
// The pixel's residual colour (what later splats and the background add)
// enters dL/dalpha only through its dot product with dL/dpixel, so the state
// is the scalar Q = residual . dL/dpixel + T_final (bg . dL/dpixel):
//   dL/dalpha = T (colour . dL/dpixel) - Q / (1 - alpha),  Q += w (colour . dL/dpixel).
// Reciprocal of the per-pixel backward: IEEE division (the GPU backward
// passes a reciprocal that equals it for every argument the walk can see).
struct IeeeRcp {
  template <typename R>
  FGS_HD R operator()(R x) const { return R(1) / x; }
};

// Initial Q of a reverse walk that starts from a checkpoint with residual
// colour (r0, r1, r2); tb = T_final (bg . dL/dpixel).
template <typename R>
FGS_HD R residual_dot(R r0, R r1, R r2, R dp0, R dp1, R dp2, R tb) {
  return fma_(r2, dp2, fma_(r1, dp1, r0 * dp0)) + tb;
}

// ABS = false leaves acc[RAW_AX], acc[RAW_AY] untouched.
template <typename R, typename Rcp = IeeeRcp, bool ABS = true>
FGS_HD void pix_bwd(R gx, R gy, R ca, R cb, R cc, R w, R r, R g, R b, R px, R py,
                    R dp0, R dp1, R dp2, R& T, R& Q, R* acc) {
  const R dx = gx - px, dy = gy - py;
  const R p = fma_(cb * dx, dy, fma_(cc * dy, dy, fma_(ca * dx, dx, w)));
  const R a = alpha_of(p);
  const R inv = Rcp()(R(1) - a);
  T = min_(T * inv, R(1));
  const R wgt = a * T;
  const R cd = fma_(b, dp2, fma_(g, dp1, r * dp0));
  const R dLda = fma_(-inv, Q, T * cd);
  Q = fma_(wgt, cd, Q);
  acc[RAW_COL0] = fma_(wgt, dp0, acc[RAW_COL0]);
  acc[RAW_COL1] = fma_(wgt, dp1, acc[RAW_COL1]);
  acc[RAW_COL2] = fma_(wgt, dp2, acc[RAW_COL2]);
  const R dLdp = dLda * (a * R(LN2_F));
  // d power / d(gx, gy) = (2 ca dx + cb dy, 2 cc dy + cb dx) and d power /
  // d(ca, cb, cc) = (dx^2, dx dy, dy^2), through u = dLdp dx, v = dLdp dy
  const R u = dLdp * dx, v = dLdp * dy;
  const R mx = fma_(R(2) * ca, u, cb * v);
  const R my = fma_(R(2) * cc, v, cb * u);
  acc[RAW_MX] += mx;
  acc[RAW_MY] += my;
  if (ABS) {
    acc[RAW_AX] += abs_(mx);
    acc[RAW_AY] += abs_(my);
  }
  acc[RAW_KXX] = fma_(u, dx, acc[RAW_KXX]);
  acc[RAW_KXY] = fma_(u, dy, acc[RAW_KXY]);
  acc[RAW_KYY] = fma_(v, dy, acc[RAW_KYY]);
  acc[RAW_POW] += dLdp;
}

// ------------------------------------------------- per-Gaussian backward
// Gradients a Gaussian receives from the rasterizer in the reference's
// tensor conventions (pybind.cpp:258-267): mean2D (P,4) holds NDC gradients
// (x, y, |x|, |y|); conic (P,2,2) holds (x, y, -, w) with the off-diagonal
// entry carrying half the derivative w.r.t. conic b (backward.cu:624-628).
template <typename R>
struct Grad2D {
  R mean2d[4];
  R conic[4];
  R opacity;
  R color[3];
};

template <typename R>
FGS_HD void finalize_raw(const R* raw, R opacity, int W, int H, Grad2D<R>& g) {
  const R hx = R(0.5f) * (R)W, hy = R(0.5f) * (R)H;
  const R kc = R(-0.5f * LOG2E_F);
  g.mean2d[0] = raw[RAW_MX] * hx;
  g.mean2d[1] = raw[RAW_MY] * hy;
  g.mean2d[2] = raw[RAW_AX] * hx;
  g.mean2d[3] = raw[RAW_AY] * hy;
  g.conic[0] = kc * raw[RAW_KXX];
  g.conic[1] = kc * raw[RAW_KXY];
  g.conic[2] = R(0);
  g.conic[3] = kc * raw[RAW_KYY];
  g.opacity = raw[RAW_POW] * (R(LOG2E_F) / opacity);
  g.color[0] = raw[RAW_COL0];
  g.color[1] = raw[RAW_COL1];
  g.color[2] = raw[RAW_COL2];
}

// Outputs of the per-Gaussian backward (pointers into P-sized arrays, already
// offset to Gaussian i). dsh may be null (colour given precomputed).
template <typename R>
struct GaussGrad {
  R* dmean3d;  // 3
  R* dcov3d;   // 6
  R* dscale;   // 3
  R* drot;     // 4
  R* ddc;      // 3
  R* dsh;      // max_coeffs x 3
};

// Derivative of real SH basis function k (1..15) w.r.t. the unit direction
// (x, y, z) (backward.cu:121-245); k is a compile-time constant in the
// unrolled loop of sh_jacobian, so each call folds to its own expressions.
template <typename R>
FGS_HD void sh_dbasis_k(int k, R x, R y, R z, R d[3]) {
  switch (k) {
    case 1: d[0] = R(0); d[1] = -R(SH_C1); d[2] = R(0); break;
    case 2: d[0] = R(0); d[1] = R(0); d[2] = R(SH_C1); break;
    case 3: d[0] = -R(SH_C1); d[1] = R(0); d[2] = R(0); break;
    case 4: d[0] = R(SH_C2_0) * y; d[1] = R(SH_C2_0) * x; d[2] = R(0); break;
    case 5: d[0] = R(0); d[1] = R(SH_C2_1) * z; d[2] = R(SH_C2_1) * y; break;
    case 6:
      d[0] = R(-2) * R(SH_C2_2) * x; d[1] = R(-2) * R(SH_C2_2) * y; d[2] = R(4) * R(SH_C2_2) * z;
      break;
    case 7: d[0] = R(SH_C2_3) * z; d[1] = R(0); d[2] = R(SH_C2_3) * x; break;
    case 8: d[0] = R(2) * R(SH_C2_4) * x; d[1] = R(-2) * R(SH_C2_4) * y; d[2] = R(0); break;
    case 9:
      d[0] = R(6) * R(SH_C3_0) * (x * y); d[1] = R(SH_C3_0) * (R(3) * (x * x) - R(3) * (y * y));
      d[2] = R(0);
      break;
    case 10:
      d[0] = R(SH_C3_1) * (y * z); d[1] = R(SH_C3_1) * (x * z); d[2] = R(SH_C3_1) * (x * y);
      break;
    case 11:
      d[0] = R(-2) * R(SH_C3_2) * (x * y);
      d[1] = R(SH_C3_2) * (R(4) * (z * z) - (x * x) - R(3) * (y * y));
      d[2] = R(8) * R(SH_C3_2) * (y * z);
      break;
    case 12:
      d[0] = R(-6) * R(SH_C3_3) * (x * z); d[1] = R(-6) * R(SH_C3_3) * (y * z);
      d[2] = R(SH_C3_3) * (R(6) * (z * z) - R(3) * (x * x) - R(3) * (y * y));
      break;
    case 13:
      d[0] = R(SH_C3_4) * (R(4) * (z * z) - R(3) * (x * x) - (y * y));
      d[1] = R(-2) * R(SH_C3_4) * (x * y); d[2] = R(8) * R(SH_C3_4) * (x * z);
      break;
    case 14:
      d[0] = R(2) * R(SH_C3_5) * (x * z); d[1] = R(-2) * R(SH_C3_5) * (y * z);
      d[2] = R(SH_C3_5) * ((x * x) - (y * y));
      break;
    default:
      d[0] = R(SH_C3_6) * (R(3) * (x * x) - R(3) * (y * y)); d[1] = R(-6) * R(SH_C3_6) * (x * y);
      d[2] = R(0);
      break;
  }
}

// Jacobian of the SH colour w.r.t. the unit view direction (x, y, z):
// J[ch][j] = sum over k = 1 .. n-1 in order of sh_k[ch] d basis_k / d dir_j
// (the reference forms the same three colour vectors dRGBdx, dRGBdy, dRGBdz,
// backward.cu:121-245). The forward's colour pass stores it for the
// backward, which then needs no coefficient reads for the mean gradient.
template <typename R>
FGS_HD void sh_jacobian(R x, R y, R z, const R* sh, int deg, R J[3][3]) {
  const int n = sh_count(deg);
  for (int ch = 0; ch < 3; ++ch)
    for (int j = 0; j < 3; ++j) J[ch][j] = R(0);
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int k = 1; k < 16; ++k)
    if (k < n) {
      R d[3];
      sh_dbasis_k(k, x, y, z, d);
      for (int ch = 0; ch < 3; ++ch) {
        const R c = sh[3 * (k - 1) + ch];
        for (int j = 0; j < 3; ++j) J[ch][j] = J[ch][j] + c * d[j];
      }
    }
}
// dL/d(direction) from the Jacobian: ddir_j = sum_ch dRGB[ch] J[ch][j].
template <typename R>
FGS_HD void sh_ddir(const R J[3][3], const R dRGB[3], R ddir[3]) {
  for (int j = 0; j < 3; ++j) ddir[j] = dRGB[0] * J[0][j] + dRGB[1] * J[1][j] + dRGB[2] * J[2][j];
}

// Colour gradient seen by the SH coefficients: zero on clamped channels.
template <typename R>
FGS_HD void sh_drgb(int clamp, const R color[3], R dRGB[3]) {
  for (int ch = 0; ch < 3; ++ch) dRGB[ch] = (clamp >> ch) & 1 ? R(0) : color[ch];
}

// Mean gradient through the normalisation dir = o / |o| (backward.cu:88-103).
template <typename R>
FGS_HD void sh_dmean(R ox, R oy, R oz, const R ddir[3], R dm[3]) {
  const R s2 = ox * ox + oy * oy + oz * oz;
  const R inv32 = R(1) / sqrt_(s2 * s2 * s2);
  dm[0] = ((s2 - ox * ox) * ddir[0] - oy * ox * ddir[1] - oz * ox * ddir[2]) * inv32;
  dm[1] = (-ox * oy * ddir[0] + (s2 - oy * oy) * ddir[1] - oz * oy * ddir[2]) * inv32;
  dm[2] = (-ox * oz * ddir[0] - oy * oz * ddir[1] + (s2 - oz * oz) * ddir[2]) * inv32;
}

// Per-Gaussian adjoint of preprocess for a Gaussian with radius > 0:
// cov2D path (backward.cu:250-403), projection path (:504-526), SH path
// (:121-245) and cov3D path (:407-483). Like the reference, the scale
// gradient is taken w.r.t. s = scale_modifier * scale (it equals the
// gradient w.r.t. scale when the modifier is 1), and the rotation gradient
// is w.r.t. the unnormalised quaternion.
// With sh_dm given, the SH path is left out (a separate pass computed its
// coefficient gradients) and sh_dm is added as its mean-gradient term.
template <typename R>
FGS_HD void gauss_bwd(int i, const GaussRef<R>& G, const CamRef<R>& cam, int clamp,
                      const Grad2D<R>& g2, const GaussGrad<R>& out, const R* sh_row = nullptr,
                      const R* sh_dm = nullptr) {
  const R* pm = G.means + 3 * i;
  R t[3];
  xform_point(cam.view, pm[0], pm[1], pm[2], t);
  R s[3], q[4], cov[6];
  for (int k = 0; k < 3; ++k) s[k] = cam.scale_mod * G.scales[3 * i + k];
  for (int k = 0; k < 4; ++k) q[k] = G.rots[4 * i + k];
  cov3d_of(s, q, cov);
  R T[6], tx, ty, xm, ym, a, b, c;
  ewa_T(cam, t, T, tx, ty, xm, ym);
  cov2d_of(T, cov, a, b, c);

  // conic -> 2D covariance
  const R Gx = g2.conic[0], Gy = g2.conic[1], Gz = g2.conic[3];
  const R den = a * c - b * b;
  const R inv2 = R(1) / (den * den + R(W_EPS_F));
  R dLda = R(0), dLdb = R(0), dLdc = R(0);
  R dS[6] = {R(0), R(0), R(0), R(0), R(0), R(0)};
  if (inv2 != R(0)) {
    dLda = inv2 * (-c * c * Gx + R(2) * b * c * Gy + (den - a * c) * Gz);
    dLdc = inv2 * (-a * a * Gz + R(2) * a * b * Gy + (den - a * c) * Gx);
    dLdb = inv2 * R(2) * (b * c * Gx - (den + R(2) * b * b) * Gy + a * b * Gz);
    // 2D covariance -> Sigma (symmetric storage, off-diagonals counted twice)
    dS[0] = T[0] * T[0] * dLda + T[0] * T[3] * dLdb + T[3] * T[3] * dLdc;
    dS[3] = T[1] * T[1] * dLda + T[1] * T[4] * dLdb + T[4] * T[4] * dLdc;
    dS[5] = T[2] * T[2] * dLda + T[2] * T[5] * dLdb + T[5] * T[5] * dLdc;
    dS[1] = R(2) * T[0] * T[1] * dLda + (T[0] * T[4] + T[1] * T[3]) * dLdb + R(2) * T[3] * T[4] * dLdc;
    dS[2] = R(2) * T[0] * T[2] * dLda + (T[0] * T[5] + T[2] * T[3]) * dLdb + R(2) * T[3] * T[5] * dLdc;
    dS[4] = R(2) * T[1] * T[2] * dLda + (T[1] * T[5] + T[2] * T[4]) * dLdb + R(2) * T[4] * T[5] * dLdc;
  }
  for (int k = 0; k < 6; ++k) out.dcov3d[k] = dS[k];

  // 2D covariance -> T = J V3 (T Sigma T^T): dT0 = 2 dLda Sigma T0 + dLdb Sigma T1, etc.
  const R S00 = cov[0], S01 = cov[1], S02 = cov[2], S11 = cov[3], S12 = cov[4], S22 = cov[5];
  const R u0[3] = {S00 * T[0] + S01 * T[1] + S02 * T[2], S01 * T[0] + S11 * T[1] + S12 * T[2],
                   S02 * T[0] + S12 * T[1] + S22 * T[2]};
  const R u1[3] = {S00 * T[3] + S01 * T[4] + S02 * T[5], S01 * T[3] + S11 * T[4] + S12 * T[5],
                   S02 * T[3] + S12 * T[4] + S22 * T[5]};
  R dT0[3], dT1[3];
  for (int k = 0; k < 3; ++k) {
    dT0[k] = R(2) * dLda * u0[k] + dLdb * u1[k];
    dT1[k] = R(2) * dLdc * u1[k] + dLdb * u0[k];
  }
  const R* v = cam.view;
  const R dJ00 = dT0[0] * v[0] + dT0[1] * v[4] + dT0[2] * v[8];
  const R dJ02 = dT0[0] * v[2] + dT0[1] * v[6] + dT0[2] * v[10];
  const R dJ11 = dT1[0] * v[1] + dT1[1] * v[5] + dT1[2] * v[9];
  const R dJ12 = dT1[0] * v[2] + dT1[1] * v[6] + dT1[2] * v[10];
  const R tz = R(1) / t[2];
  const R tz2 = tz * tz, tz3 = tz2 * tz;
  const R fx = cam.focal_x, fy = cam.focal_y;
  const R dtx = xm * -fx * tz2 * dJ02;
  const R dty = ym * -fy * tz2 * dJ12;
  const R dtz = -fx * tz2 * dJ00 - fy * tz2 * dJ11 + (R(2) * fx * tx) * tz3 * dJ02 +
                (R(2) * fy * ty) * tz3 * dJ12;
  R dm[3];
  dm[0] = v[0] * dtx + v[1] * dty + v[2] * dtz;
  dm[1] = v[4] * dtx + v[5] * dty + v[6] * dtz;
  dm[2] = v[8] * dtx + v[9] * dty + v[10] * dtz;

  // projection path: NDC mean gradient -> mean3D
  const R* P = cam.proj;
  R hom[3];
  xform_point(P, pm[0], pm[1], pm[2], hom);
  const R mw = R(1) / (xform_w(P, pm[0], pm[1], pm[2]) + R(W_EPS_F));
  const R mul1 = hom[0] * mw * mw, mul2 = hom[1] * mw * mw;
  const R gx = g2.mean2d[0], gy = g2.mean2d[1];
  dm[0] += (P[0] * mw - P[3] * mul1) * gx + (P[1] * mw - P[3] * mul2) * gy;
  dm[1] += (P[4] * mw - P[7] * mul1) * gx + (P[5] * mw - P[7] * mul2) * gy;
  dm[2] += (P[8] * mw - P[11] * mul1) * gx + (P[9] * mw - P[11] * mul2) * gy;

  // SH path (dc alone at degree 0; rest coefficients for k < sh_count(deg))
  if (sh_dm != nullptr) {
    for (int k = 0; k < 3; ++k) dm[k] += sh_dm[k];
  } else if (G.colors == nullptr && G.dc != nullptr) {
    const R ox = pm[0] - cam.campos[0], oy = pm[1] - cam.campos[1], oz = pm[2] - cam.campos[2];
    const R len = sqrt_(ox * ox + oy * oy + oz * oz);
    const R x = ox / len, y = oy / len, z = oz / len;
    R bas[16];
    sh_basis(x, y, z, bas);
    R dRGB[3];
    sh_drgb(clamp, g2.color, dRGB);
    const int n = sh_count(cam.deg);
    const R* sh = sh_row ? sh_row : G.shs + (size_t)i * cam.max_coeffs * 3;
    R J[3][3], ddir[3];
    sh_jacobian(x, y, z, sh, cam.deg, J);
    sh_ddir(J, dRGB, ddir);
    // SH coefficient gradients (written after sh has been read, so sh and dsh
    // may share storage)
    for (int ch = 0; ch < 3; ++ch) out.ddc[ch] = bas[0] * dRGB[ch];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int k = 1; k < 16; ++k)
      if (k < n)
        for (int ch = 0; ch < 3; ++ch) out.dsh[3 * (k - 1) + ch] = bas[k] * dRGB[ch];
    R dms[3];
    sh_dmean(ox, oy, oz, ddir, dms);
    for (int k = 0; k < 3; ++k) dm[k] += dms[k];
  }
  for (int k = 0; k < 3; ++k) out.dmean3d[k] = dm[k];

  // cov3D path: Sigma = M M^T with M = Rq diag(s); dL/dM = 2 Gs M where Gs
  // is the symmetric matrix gradient (off-diagonals halved).
  R Rm[9];
  quat_rot(q, Rm);
  R M[9];
  for (int r_ = 0; r_ < 3; ++r_)
    for (int c_ = 0; c_ < 3; ++c_) M[3 * r_ + c_] = Rm[3 * r_ + c_] * s[c_];
  const R Gs[9] = {dS[0], R(0.5f) * dS[1], R(0.5f) * dS[2],
                   R(0.5f) * dS[1], dS[3], R(0.5f) * dS[4],
                   R(0.5f) * dS[2], R(0.5f) * dS[4], dS[5]};
  R dM[9];
  for (int r_ = 0; r_ < 3; ++r_)
    for (int c_ = 0; c_ < 3; ++c_)
      dM[3 * r_ + c_] = R(2) * (Gs[3 * r_ + 0] * M[c_] + Gs[3 * r_ + 1] * M[3 + c_] +
                                Gs[3 * r_ + 2] * M[6 + c_]);
  for (int c_ = 0; c_ < 3; ++c_)
    out.dscale[c_] = dM[c_] * Rm[c_] + dM[3 + c_] * Rm[3 + c_] + dM[6 + c_] * Rm[6 + c_];
  R dR[9];
  for (int r_ = 0; r_ < 3; ++r_)
    for (int c_ = 0; c_ < 3; ++c_) dR[3 * r_ + c_] = dM[3 * r_ + c_] * s[c_];
  const R qr = q[0], qx = q[1], qy = q[2], qz = q[3];
  out.drot[0] = R(2) * qz * (dR[3] - dR[1]) + R(2) * qy * (dR[2] - dR[6]) + R(2) * qx * (dR[7] - dR[5]);
  out.drot[1] = R(2) * qy * (dR[1] + dR[3]) + R(2) * qz * (dR[2] + dR[6]) +
                R(2) * qr * (dR[7] - dR[5]) - R(4) * qx * (dR[4] + dR[8]);
  out.drot[2] = R(2) * qx * (dR[1] + dR[3]) + R(2) * qr * (dR[2] - dR[6]) +
                R(2) * qz * (dR[5] + dR[7]) - R(4) * qy * (dR[0] + dR[8]);
  out.drot[3] = R(2) * qr * (dR[3] - dR[1]) + R(2) * qx * (dR[2] + dR[6]) +
                R(2) * qy * (dR[5] + dR[7]) - R(4) * qz * (dR[0] + dR[4]);
}

// Depth key bits of a visible Gaussian: its depth is > 0.2, so the IEEE bit
// pattern orders like the value (preprocess.cu:618-620).
FGS_HD uint32_t depth_bits(float z) { return f2u(z); }

// ------------------------------------------------------------ render_depth
// Inverse of a 4x4 matrix (column-major) by cofactor expansion; zeros when
// singular. Used once per tile for the pixel rays of render_depth.
template <typename R>
FGS_HD void inverse4(const R* m, R* inv) {
  R c[16];
  c[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
         m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  c[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
         m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  c[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
         m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  c[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
          m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  c[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
         m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  c[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
         m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  c[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
         m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  c[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
          m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  c[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
         m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  c[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
         m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  c[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
          m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  c[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
          m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  c[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
         m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  c[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
         m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  c[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
          m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  c[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
          m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  const R det = m[0] * c[0] + m[1] * c[4] + m[2] * c[8] + m[3] * c[12];
  const R id = det != R(0) ? R(1) / det : R(0);
  for (int k = 0; k < 16; ++k) inv[k] = c[k] * id;
}

// Camera ray of pixel (px, py) as render_depth builds it
// (render_depth.cu:158-186): the pixel's NDC point (scaled by 1.0000001, at
// the normalised depth (100 - 100 * 0.01) / (100 - 0.01) of a 0.01 / 100
// frustum) mapped through the inverse projection without a perspective
// divide, minus the camera centre. ndir is the unit direction, len the
// length before normalisation.
template <typename R>
FGS_HD void depth_ray(const R* proj_inv, const R* campos, int px, int py, int W, int H, R ndir[3],
                      R& len) {
  const R nx = (R(2) * (R)px + R(1)) / (R)W - R(1);
  const R ny = (R(2) * (R)py + R(1)) / (R)H - R(1);
  const R hx = (R)((double)nx * 1.0000001), hy = (R)((double)ny * 1.0000001);
  const R hz = (R)((100.0 - 100.0 * 0.01) / (100.0 - 0.01));
  R o[3];
  xform_point(proj_inv, hx, hy, hz, o);
  const R d0 = o[0] - campos[0], d1 = o[1] - campos[1], d2 = o[2] - campos[2];
  len = sqrt_(d0 * d0 + d1 * d1 + d2 * d2);
  ndir[0] = d0 / len;
  ndir[1] = d1 / len;
  ndir[2] = d2 / len;
}

// Ray against the 3-sigma ellipsoid of a Gaussian (render_depth.cu:220-256),
// in the frame rotated by the given (unnormalised) quaternion: the quadratic
// a t^2 + b t + c in the parameter t along the unit direction. The parts
// that depend on the Gaussian and the ray origin only (rotation, rotated
// origin, 3 sigma, c) are set up once per Gaussian; ellipsoid_hit adds the
// direction and returns the midpoint t_mid = -b / 2 / a and the
// discriminant b^2 - 4ac.
template <typename R>
struct Ellipsoid {
  R Rm[9];  // quat_rot of the quaternion; applied as sum_c Rm[3c + r] v[c]
  R ro[3];  // rotated (origin - mean)
  R s3[3];  // 3 * scale
  R c;
};

template <typename R>
FGS_HD void ellipsoid_setup(const R* mean, const R* scale, const R* q, const R* origin,
                            Ellipsoid<R>& E) {
  quat_rot(q, E.Rm);
  const R tmp[3] = {origin[0] - mean[0], origin[1] - mean[1], origin[2] - mean[2]};
  for (int r = 0; r < 3; ++r)
    E.ro[r] = E.Rm[r] * tmp[0] + E.Rm[3 + r] * tmp[1] + E.Rm[6 + r] * tmp[2];
  R ct[3];
  for (int k = 0; k < 3; ++k) {
    E.s3[k] = scale[k] * R(3);
    ct[k] = E.ro[k] / E.s3[k] * E.ro[k] / E.s3[k];
  }
  E.c = ct[0] + ct[1] + ct[2] - R(1);
}

template <typename R>
FGS_HD void ellipsoid_hit(const Ellipsoid<R>& E, const R* ndir, R& t_mid, R& disc) {
  R at[3], bt[3];
  for (int k = 0; k < 3; ++k) {
    const R rd = E.Rm[k] * ndir[0] + E.Rm[3 + k] * ndir[1] + E.Rm[6 + k] * ndir[2];
    const R q = rd / E.s3[k];
    at[k] = q * rd / E.s3[k];
    bt[k] = q * E.ro[k] / E.s3[k];
  }
  const R a = at[0] + at[1] + at[2];
  const R b = R(2) * (bt[0] + bt[1] + bt[2]);
  disc = b * b - R(4) * a * E.c;
  t_mid = -b / R(2) / a;
}

// One pixel of render_depth (render_depth.cu:207-288), updated with one
// splat of its list that has alpha >= 1/255: a splat that would bring T
// below 1e-4 ends the pixel; a hit behind the camera (depth < 0) is skipped
// without changing T; otherwise the splat with the largest weight alpha*T
// so far (the first of equals) sets depth, discriminant, index and the
// point origin + t_mid * ndir, and T becomes T (1 - alpha).
template <typename R>
struct DepthPix {
  R T = R(1), wmax = R(0), depth = R(0), disc = R(0);
  R pt[3] = {R(0), R(0), R(0)};
  int32_t idx = 0;
  bool done = false;
};

template <typename R>
FGS_HD void depth_pix_step(DepthPix<R>& D, R alpha, uint32_t id, const Ellipsoid<R>& E,
                           const R* origin, const R* ndir, R len) {
  const R test_T = D.T * (R(1) - alpha);
  if (test_T < R(T_CUT_F)) {
    D.done = true;
    return;
  }
  R t_mid, disc;
  ellipsoid_hit(E, ndir, t_mid, disc);
  const R depth = t_mid / len;
  if (depth < R(0)) return;
  const R w = alpha * D.T;
  if (D.wmax < w) {
    D.wmax = w;
    D.depth = depth;
    D.disc = disc;
    D.idx = (int32_t)id;
    for (int k = 0; k < 3; ++k) D.pt[k] = origin[k] + t_mid * ndir[k];
  }
  D.T = test_T;
}

}  // namespace fgs
