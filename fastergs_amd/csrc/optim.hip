// Visibility-masked Adam of the FasterGS training loop
// (SparseGaussianAdam.step, diff_gaussian_rasterization_faster/__init__.py:267-298;
// adamUpdateCUDA, csrc/cuda_rasterizer/adam.cu:9-38). For every Gaussian g
// with visible[g] and each of its M parameters:
//   m = b1 m + (1 - b1) g;  v = b2 v + (1 - b2) g g;  p += -lr m / (sqrt(v) + eps)
// with lr already bias-compensated by the caller. Invisible Gaussians keep
// their parameters and moments.
//
// One launch updates all parameter tensors of a step (the reference launches
// once per tensor): each tensor owns a contiguous range of workgroups, and a
// workgroup handles 1024 consecutive elements of its tensor with float4
// accesses where the row length M allows.
#include <hip/hip_runtime.h>

#include "fgs_device.h"
#include "fgs_kernels.h"
#include "fgs_prof.h"

namespace fgs {

constexpr int ADAM_THREADS = 256;
constexpr int ADAM_ITEMS = 4;
constexpr int ADAM_TILE = ADAM_THREADS * ADAM_ITEMS;

__device__ __forceinline__ void adam_elem(float& p, float g, float& m, float& v, float lr,
                                          float b1, float b2, float eps) {
  m = b1 * m + (1.0f - b1) * g;
  v = b2 * v + (1.0f - b2) * g * g;
  p += -lr * m / (sqrt_(v) + eps);
}

__global__ __launch_bounds__(ADAM_THREADS) void k_sparse_adam(AdamArgs a) {
  int t = 0;
  while (t + 1 < a.n && (int64_t)blockIdx.x >= a.block_off[t + 1]) ++t;
  const AdamTensor& d = a.t[t];
  const int64_t n = (int64_t)a.N * d.M;
  const int64_t e0 = ((int64_t)blockIdx.x - a.block_off[t]) * ADAM_TILE + (int64_t)threadIdx.x * ADAM_ITEMS;
  if (e0 >= n) return;
  if (d.vec && e0 + ADAM_ITEMS <= n) {
    // four consecutive elements; they may span two Gaussians when M is not 4
    float4 p = *(const float4*)(d.param + e0);
    const float4 g = *(const float4*)(d.grad + e0);
    float4 m = *(const float4*)(d.exp_avg + e0);
    float4 v = *(const float4*)(d.exp_avg_sq + e0);
    float* pp = &p.x;
    const float* gg = &g.x;
    float* mm = &m.x;
    float* vv = &v.x;
    bool any = false;
#pragma unroll
    for (int k = 0; k < ADAM_ITEMS; ++k) {
      if (a.visible[(e0 + k) / d.M]) {
        adam_elem(pp[k], gg[k], mm[k], vv[k], d.lr, a.b1, a.b2, d.eps);
        any = true;
      }
    }
    if (any) {
      *(float4*)(d.param + e0) = p;
      *(float4*)(d.exp_avg + e0) = m;
      *(float4*)(d.exp_avg_sq + e0) = v;
    }
    return;
  }
  for (int k = 0; k < ADAM_ITEMS && e0 + k < n; ++k) {
    const int64_t e = e0 + k;
    if (!a.visible[e / d.M]) continue;
    float p = d.param[e], m = d.exp_avg[e], v = d.exp_avg_sq[e];
    adam_elem(p, d.grad[e], m, v, d.lr, a.b1, a.b2, d.eps);
    d.param[e] = p;
    d.exp_avg[e] = m;
    d.exp_avg_sq[e] = v;
  }
}

hipError_t launch_sparse_adam(AdamArgs& a, hipStream_t s) {
  int64_t blocks = 0;
  for (int t = 0; t < a.n; ++t) {
    a.block_off[t] = blocks;
    const int64_t n = (int64_t)a.N * a.t[t].M;
    a.t[t].vec = ((((uintptr_t)a.t[t].param | (uintptr_t)a.t[t].grad | (uintptr_t)a.t[t].exp_avg |
                    (uintptr_t)a.t[t].exp_avg_sq) & 15) == 0) ? 1 : 0;
    blocks += (n + ADAM_TILE - 1) / ADAM_TILE;
  }
  if (blocks == 0) return hipSuccess;
  FGS_TIMED("sparse_adam", s, (k_sparse_adam<<<(unsigned)blocks, ADAM_THREADS, 0, s>>>(a)));
  return hipGetLastError();
}

}  // namespace fgs
