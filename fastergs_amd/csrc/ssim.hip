// Fused SSIM loss (forward map + partial derivatives, and its backward) for
// the image-space loss that feeds the rasterizer's backward in the FasterGS
// training loops (fused-ssim: gs_submodules/fused-ssim/ssim.cu:186-441,
// fused_ssim/__init__.py; used by FastGS/train_faster.py:18).
//
// Semantics: 11-tap separable Gaussian window (sigma 1.5), zero padding
// outside the image, per channel
//   mu = G*x, sigma^2 = G*(x^2) - mu^2, sigma12 = G*(x1 x2) - mu1 mu2,
//   m = (2 mu1 mu2 + C1)(2 sigma12 + C2) / ((mu1^2 + mu2^2 + C1)(sigma1^2 + sigma2^2 + C2))
// plus dm/dmu1, dm/dsigma1^2, dm/dsigma12 for the backward,
//   dL/dx1 = G*(g dm/dmu1) + 2 x1 G*(g dm/dsigma1^2) + x2 G*(g dm/dsigma12).
// Every convolution accumulates its 11 taps in order from zero, rows first,
// with no fused multiply-adds, so the CPU restatement reproduces the maps
// bit for bit.
//
// Layout: one 256-thread workgroup per 64 x 16 pixel tile and image of the
// batch; the tile plus a 5-pixel halo is staged in LDS, the horizontal pass
// writes its sums for the 26 halo rows x 64 columns to LDS, and the vertical
// pass produces 4 pixels per thread. Channels are processed in turn.
#include <hip/hip_runtime.h>

#include "fgs_device.h"
#include "fgs_kernels.h"
#include "fgs_prof.h"

namespace fgs {

constexpr int SS_TX = 64, SS_TY = 16, SS_HALO = 5;
constexpr int SS_LX = SS_TX + 2 * SS_HALO;  // 74
constexpr int SS_LY = SS_TY + 2 * SS_HALO;  // 26
constexpr int SS_THREADS = 256;

// Window taps (a normalised 11-tap Gaussian, sigma 1.5), as float32.
__constant__ float k_ssim_g[11] = {0.001028380123898387f, 0.0075987582094967365f,
                                   0.036000773310661316f, 0.10936068743467331f,
                                   0.21300552785396576f,  0.26601171493530273f,
                                   0.21300552785396576f,  0.10936068743467331f,
                                   0.036000773310661316f, 0.0075987582094967365f,
                                   0.001028380123898387f};

__device__ __forceinline__ float pix_or_zero(const float* img, int x, int y, int W, int H) {
  return (x >= 0 && y >= 0 && x < W && y < H) ? img[(size_t)y * W + x] : 0.0f;
}

// Forward. Maps (B x CH x H x W): ssim, and when dmu1 != nullptr the three
// partial derivatives.
__global__ __launch_bounds__(SS_THREADS) void k_ssim_fwd(int CH, int H, int W, float C1, float C2,
                                                         const float* __restrict__ img1,
                                                         const float* __restrict__ img2,
                                                         float* __restrict__ map,
                                                         float* __restrict__ dmu1,
                                                         float* __restrict__ ds1,
                                                         float* __restrict__ ds12) {
  __shared__ float in1[SS_LY][SS_LX];
  __shared__ float in2[SS_LY][SS_LX];
  __shared__ float hz[5][SS_LY][SS_TX];  // mu1, mu2, x1^2, x2^2, x1 x2 (horizontal)
  const int t = threadIdx.x;
  const int x0 = blockIdx.x * SS_TX - SS_HALO, y0 = blockIdx.y * SS_TY - SS_HALO;
  const size_t plane = (size_t)H * W;
  for (int c = 0; c < CH; ++c) {
    const size_t base = ((size_t)blockIdx.z * CH + c) * plane;
    for (int e = t; e < SS_LY * SS_LX; e += SS_THREADS) {
      const int ly = e / SS_LX, lx = e - ly * SS_LX;
      in1[ly][lx] = pix_or_zero(img1 + base, x0 + lx, y0 + ly, W, H);
      in2[ly][lx] = pix_or_zero(img2 + base, x0 + lx, y0 + ly, W, H);
    }
    __syncthreads();
    for (int e = t; e < SS_LY * SS_TX; e += SS_THREADS) {
      const int ly = e / SS_TX, lx = e - ly * SS_TX;
      float m1 = 0.0f, m2 = 0.0f, q1 = 0.0f, q2 = 0.0f, q12 = 0.0f;
#pragma unroll
      for (int k = 0; k < 11; ++k) {
        const float a = in1[ly][lx + k], b = in2[ly][lx + k];
        const float g = k_ssim_g[k];
        m1 = m1 + g * a;
        m2 = m2 + g * b;
        q1 = q1 + g * (a * a);
        q2 = q2 + g * (b * b);
        q12 = q12 + g * (a * b);
      }
      hz[0][ly][lx] = m1;
      hz[1][ly][lx] = m2;
      hz[2][ly][lx] = q1;
      hz[3][ly][lx] = q2;
      hz[4][ly][lx] = q12;
    }
    __syncthreads();
    for (int e = t; e < SS_TY * SS_TX; e += SS_THREADS) {
      const int oy = e / SS_TX, ox = e - oy * SS_TX;
      float v[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int k = 0; k < 11; ++k) {
        const float g = k_ssim_g[k];
#pragma unroll
        for (int f = 0; f < 5; ++f) v[f] = v[f] + g * hz[f][oy + k][ox];
      }
      const int px = x0 + SS_HALO + ox, py = y0 + SS_HALO + oy;
      if (px < W && py < H) {
        const float mu1 = v[0], mu2 = v[1];
        const float s1 = v[2] - mu1 * mu1, s2 = v[3] - mu2 * mu2, s12 = v[4] - mu1 * mu2;
        const float mu1_sq = mu1 * mu1, mu2_sq = mu2 * mu2, mu1_mu2 = mu1 * mu2;
        const float Cn = 2.0f * mu1_mu2 + C1;
        const float Dn = 2.0f * s12 + C2;
        const float An = mu1_sq + mu2_sq + C1;
        const float Bn = s1 + s2 + C2;
        const size_t gi = base + (size_t)py * W + px;
        map[gi] = (Cn * Dn) / (An * Bn);
        if (dmu1) {
          dmu1[gi] = (mu2 * 2.0f * Dn) / (An * Bn) - (mu2 * 2.0f * Cn) / (An * Bn) -
                     (mu1 * 2.0f * Cn * Dn) / (An * An * Bn) +
                     (mu1 * 2.0f * Cn * Dn) / (An * Bn * Bn);
          ds1[gi] = (-Cn * Dn) / (An * Bn * Bn);
          ds12[gi] = (2.0f * Cn) / (An * Bn);
        }
      }
    }
    __syncthreads();
  }
}

// Backward: dL/dimg1 from dL/dmap and the stored partial derivatives.
__global__ __launch_bounds__(SS_THREADS) void k_ssim_bwd(int CH, int H, int W,
                                                         const float* __restrict__ img1,
                                                         const float* __restrict__ img2,
                                                         const float* __restrict__ dmap,
                                                         const float* __restrict__ dmu1,
                                                         const float* __restrict__ ds1,
                                                         const float* __restrict__ ds12,
                                                         float* __restrict__ dimg1) {
  __shared__ float pr[3][SS_LY][SS_LX];  // g dm/dmu1, g dm/dsigma1^2, g dm/dsigma12
  __shared__ float hz[3][SS_LY][SS_TX];
  const int t = threadIdx.x;
  const int x0 = blockIdx.x * SS_TX - SS_HALO, y0 = blockIdx.y * SS_TY - SS_HALO;
  const size_t plane = (size_t)H * W;
  for (int c = 0; c < CH; ++c) {
    const size_t base = ((size_t)blockIdx.z * CH + c) * plane;
    for (int e = t; e < SS_LY * SS_LX; e += SS_THREADS) {
      const int ly = e / SS_LX, lx = e - ly * SS_LX;
      const int x = x0 + lx, y = y0 + ly;
      const float g = pix_or_zero(dmap + base, x, y, W, H);
      pr[0][ly][lx] = pix_or_zero(dmu1 + base, x, y, W, H) * g;
      pr[1][ly][lx] = pix_or_zero(ds1 + base, x, y, W, H) * g;
      pr[2][ly][lx] = pix_or_zero(ds12 + base, x, y, W, H) * g;
    }
    __syncthreads();
    for (int e = t; e < SS_LY * SS_TX; e += SS_THREADS) {
      const int ly = e / SS_TX, lx = e - ly * SS_TX;
      float h0 = 0.0f, h1 = 0.0f, h2 = 0.0f;
#pragma unroll
      for (int k = 0; k < 11; ++k) {
        const float g = k_ssim_g[k];
        h0 = h0 + g * pr[0][ly][lx + k];
        h1 = h1 + g * pr[1][ly][lx + k];
        h2 = h2 + g * pr[2][ly][lx + k];
      }
      hz[0][ly][lx] = h0;
      hz[1][ly][lx] = h1;
      hz[2][ly][lx] = h2;
    }
    __syncthreads();
    for (int e = t; e < SS_TY * SS_TX; e += SS_THREADS) {
      const int oy = e / SS_TX, ox = e - oy * SS_TX;
      float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
#pragma unroll
      for (int k = 0; k < 11; ++k) {
        const float g = k_ssim_g[k];
        v0 = v0 + g * hz[0][oy + k][ox];
        v1 = v1 + g * hz[1][oy + k][ox];
        v2 = v2 + g * hz[2][oy + k][ox];
      }
      const int px = x0 + SS_HALO + ox, py = y0 + SS_HALO + oy;
      if (px < W && py < H) {
        const size_t gi = base + (size_t)py * W + px;
        float d = 0.0f;
        d = d + v0;
        d = d + img1[gi] * 2.0f * v1;
        d = d + img2[gi] * v2;
        dimg1[gi] = d;
      }
    }
    __syncthreads();
  }
}

hipError_t launch_ssim_fwd(int B, int CH, int H, int W, float C1, float C2, const float* img1,
                           const float* img2, float* map, float* dmu1, float* ds1, float* ds12,
                           hipStream_t s) {
  if (B <= 0 || CH <= 0 || H <= 0 || W <= 0) return hipSuccess;
  const dim3 grid((W + SS_TX - 1) / SS_TX, (H + SS_TY - 1) / SS_TY, B);
  FGS_TIMED("ssim_fwd", s,
            (k_ssim_fwd<<<grid, SS_THREADS, 0, s>>>(CH, H, W, C1, C2, img1, img2, map, dmu1, ds1,
                                                     ds12)));
  return hipGetLastError();
}

hipError_t launch_ssim_bwd(int B, int CH, int H, int W, const float* img1, const float* img2,
                           const float* dmap, const float* dmu1, const float* ds1,
                           const float* ds12, float* dimg1, hipStream_t s) {
  if (B <= 0 || CH <= 0 || H <= 0 || W <= 0) return hipSuccess;
  const dim3 grid((W + SS_TX - 1) / SS_TX, (H + SS_TY - 1) / SS_TY, B);
  FGS_TIMED("ssim_bwd", s,
            (k_ssim_bwd<<<grid, SS_THREADS, 0, s>>>(CH, H, W, img1, img2, dmap, dmu1, ds1, ds12,
                                                     dimg1)));
  return hipGetLastError();
}

}  // namespace fgs
