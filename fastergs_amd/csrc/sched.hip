// Tile scheduling for the FAST compositing pass and the cost model that
// steers it.
//
// The compositing pass runs one workgroup per work item (a tile, or one of
// its two 16x8 row bands). Which tiles are split and in which order the
// items are dispatched changes nothing in the outputs (pixels are
// independent and each keeps its own bucket order), only how evenly the
// work lands on the 256 CUs. Three schedules:
//   raster     one item per tile, tile order (the reference's render.cu:418)
//   heuristic  tile order, tiles whose list has >= split_len entries split
//              (a fixed threshold, like the reference's fixed 128-tile split
//              of large Gaussians, preprocess.cu:502-504); decided inside
//              the compositing kernel, two grid slots per tile
//   learned    a small MLP predicts each tile's cost (log2 of its measured
//              duration) and a split flag from 8 tile features; split flags
//              applied in tile order like the heuristic, or (order) items
//              dispatched longest-predicted first from a work list
// The MLP (8 -> 64 -> 64 -> 2, ReLU) is trained on the durations the
// compositing pass logs for every item (wall clock, 100 MHz ticks), with the
// kernels below: per-row forward/backward, fixed-order weight-gradient
// reductions and Adam. Inference rounds weights and activations to bfloat16
// (round to nearest even) and accumulates in fp32.
#include <hip/hip_runtime.h>

#include "fgs_device.h"
#include "fgs_kernels.h"
#include "fgs_layout.h"
#include "fgs_prof.h"
#include "fgs_sched.h"

namespace fgs {

// ------------------------------------------------------------ features
// Features of the first FEAT_SAMPLE entries of each tile's depth-ordered
// list (the front of the list decides when the tile's pixels stop):
//   0 log2(1 + list length)
//   1 mean and 2 max of log2(rect area in tiles) of the sampled splats
//   3..6 opacity quantiles 0.25 / 0.5 / 0.75 / 1.0 (16-bin histogram)
//   7 log2(max depth / min depth) of the sample
constexpr int FEAT_SAMPLE = 128;

__global__ __launch_bounds__(FEAT_SAMPLE) void k_tile_features(
    int T, const uint32_t* __restrict__ ranges, const uint32_t* __restrict__ vals,
    const float* __restrict__ geom, const float* __restrict__ splat, float* __restrict__ feat) {
  __shared__ uint32_t hist[16];
  __shared__ float red[4][FEAT_SAMPLE / WAVE];
  const int tile = blockIdx.x;
  const int t = threadIdx.x;
  const uint32_t start = ranges[2 * tile], end = ranges[2 * tile + 1];
  const uint32_t len = end - start;
  const int n = (int)min(len, (uint32_t)FEAT_SAMPLE);
  if (t < 16) hist[t] = 0;
  __syncthreads();
  float la = 0.0f, lmax = 0.0f, dmin = 3.0e38f, dmax = 0.0f;
  if (t < n) {
    const uint32_t g = vals[start + t];
    const float* gr = geom + 8 * (size_t)g;
    const uint32_t rmin = f2u(gr[6]), rmax = f2u(gr[7]);
    const int area = (int)((rmax & 0xffffu) - (rmin & 0xffffu)) * (int)((rmax >> 16) - (rmin >> 16));
    la = log2f((float)max(area, 1));
    lmax = la;
    const float* sp = splat + 12 * (size_t)g;
    const float op = exp2f(sp[5]);
    const float d = sp[6];
    dmin = d;
    dmax = d;
    atomicAdd(&hist[min((int)(op * 16.0f), 15)], 1u);
  }
  // block reductions: sum / max of log-areas, min / max depth
  const float s1 = wave_sum(la);
  float m1 = lmax, m2 = dmin, m3 = dmax;
  for (int o = 32; o > 0; o >>= 1) {
    m1 = fmaxf(m1, __shfl_xor(m1, o, WAVE));
    m2 = fminf(m2, __shfl_xor(m2, o, WAVE));
    m3 = fmaxf(m3, __shfl_xor(m3, o, WAVE));
  }
  const int w = t / WAVE;
  if ((t & (WAVE - 1)) == 0) {
    red[0][w] = s1;
    red[1][w] = m1;
    red[2][w] = m2;
    red[3][w] = m3;
  }
  __syncthreads();
  if (t == 0) {
    float S = 0.0f, M = 0.0f, dn = 3.0e38f, dx = 0.0f;
    for (int q = 0; q < FEAT_SAMPLE / WAVE; ++q) {
      S += red[0][q];
      M = fmaxf(M, red[1][q]);
      dn = fminf(dn, red[2][q]);
      dx = fmaxf(dx, red[3][q]);
    }
    float* f = feat + 8 * (size_t)tile;
    f[0] = log2f(1.0f + (float)len);
    f[1] = n > 0 ? S / (float)n : 0.0f;
    f[2] = M;
    const float qs[4] = {0.25f, 0.5f, 0.75f, 1.0f};
    int bin = 0;
    uint32_t cum = 0;
    for (int k = 0; k < 4; ++k) {
      const uint32_t need = n > 0 ? (uint32_t)ceilf(qs[k] * (float)n) : 0u;
      while (bin < 15 && cum + hist[bin] < need) cum += hist[bin++];
      f[3 + k] = n > 0 ? ((float)bin + 0.5f) / 16.0f : 0.0f;
    }
    f[7] = (n > 0 && dn > 0.0f) ? log2f(dx / dn) : 0.0f;
  }
}

hipError_t launch_tile_features(int T, const uint32_t* ranges, const uint32_t* vals,
                                const float* geom, const float* splat, float* feat,
                                hipStream_t s) {
  if (T > 0)
    FGS_TIMED("tile_features", s,
              (k_tile_features<<<T, FEAT_SAMPLE, 0, s>>>(T, ranges, vals, geom, splat, feat)));
  return hipGetLastError();
}

// ---------------------------------------------------------------- MLP
__device__ __forceinline__ float bf16_round(float x) {
  uint32_t u = f2u(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u2f(u & 0xffff0000u);
}

// Inference. One wave evaluates a stream of rows: lane j owns hidden unit j
// of both layers (its rows of W1 and W2 and its two output weights stay in
// registers), the first layer's outputs reach every lane through readlane,
// and the two outputs are wave sums. BF16: weights, inputs and hidden
// activations rounded to bfloat16, fp32 accumulation; else plain fp32.
template <bool BF16>
__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ params,
                                                   const float* __restrict__ x, int64_t n,
                                                   float* __restrict__ y) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t wave = ((int64_t)blockIdx.x * 256 + threadIdx.x) / WAVE;
  const int64_t nwaves = (int64_t)gridDim.x * (256 / WAVE);
  auto rnd = [](float v) { return BF16 ? bf16_round(v) : v; };
  float w1[MLP_IN], w2[MLP_HID];
#pragma unroll
  for (int k = 0; k < MLP_IN; ++k) w1[k] = rnd(params[MLP_W1 + lane * MLP_IN + k]);
#pragma unroll
  for (int k = 0; k < MLP_HID; ++k) w2[k] = rnd(params[MLP_W2 + lane * MLP_HID + k]);
  const float b1 = rnd(params[MLP_B1 + lane]), b2 = rnd(params[MLP_B2 + lane]);
  const float w30 = rnd(params[MLP_W3 + lane]), w31 = rnd(params[MLP_W3 + MLP_HID + lane]);
  const float b30 = rnd(params[MLP_B3]), b31 = rnd(params[MLP_B3 + 1]);
  for (int64_t r = wave; r < n; r += nwaves) {
    float z1 = b1;
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) z1 = fmaf(w1[k], rnd(x[r * MLP_IN + k]), z1);
    const float h1 = rnd(fmaxf(z1, 0.0f));
    float z2 = b2;
#pragma unroll
    for (int k = 0; k < MLP_HID; ++k)
      z2 = fmaf(w2[k], __int_as_float(__builtin_amdgcn_readlane(__float_as_int(h1), k)), z2);
    const float h2 = rnd(fmaxf(z2, 0.0f));
    const float o0 = wave_sum(w30 * h2), o1 = wave_sum(w31 * h2);
    if (lane == 0) {
      y[2 * r] = o0 + b30;
      y[2 * r + 1] = o1 + b31;
    }
  }
}

hipError_t launch_mlp_infer(const float* params, const float* x, int64_t n, float* y, bool bf16,
                            hipStream_t s) {
  if (n <= 0) return hipSuccess;
  // up to 4096 waves (one row each for a frame's tiles): the rows of a wave
  // are sequential, so spreading them wins over reloading the weights
  const int64_t want = (n + 3) / 4;
  const unsigned blocks = (unsigned)(want < 1024 ? want : 1024);
  if (bf16)
    FGS_TIMED("mlp_infer", s, (k_mlp_infer<true><<<blocks, 256, 0, s>>>(params, x, n, y)));
  else
    FGS_TIMED("mlp_infer", s, (k_mlp_infer<false><<<blocks, 256, 0, s>>>(params, x, n, y)));
  return hipGetLastError();
}

// Features and bf16 inference of the learned schedule in one launch: one
// wave per tile (grid-stride), 16 waves per workgroup, so that a frame's
// tiles are all in flight at once and the dependent gathers of their
// features (range -> list entries -> records) overlap. The workgroup stages
// the bf16-rounded weights once in LDS (W2 rows padded to 65 floats, so lane
// j reading row j meets no bank conflicts) and each lane then keeps hidden
// unit j's weights in registers, as k_mlp_infer does. The features are those of k_tile_features (two sample entries per
// lane; sums in a different order, so equal to float rounding), the network
// is evaluated with k_mlp_infer's arithmetic (same rounding points and
// accumulation order).
constexpr int PRED_THREADS = 1024;
constexpr int PRED_BLOCKS_MAX = 1024;
constexpr int W2_PAD = MLP_HID + 1;

__global__ __launch_bounds__(PRED_THREADS) void k_tile_predict(
    int T, const uint32_t* __restrict__ ranges, const uint32_t* __restrict__ vals,
    const float* __restrict__ geom, const float* __restrict__ splat,
    const float* __restrict__ params, float* __restrict__ feat, float* __restrict__ pred) {
  __shared__ float sW2[MLP_HID * W2_PAD];
  __shared__ float sW1[MLP_HID * (MLP_IN + 1)];
  __shared__ float sB[MLP_HID * 2 + MLP_OUT * MLP_HID + MLP_OUT];  // b1 | b2 | w3 | b3
  const int t = threadIdx.x;
  for (int k = t; k < MLP_HID * MLP_HID; k += PRED_THREADS)
    sW2[(k / MLP_HID) * W2_PAD + (k % MLP_HID)] = bf16_round(params[MLP_W2 + k]);
  for (int k = t; k < MLP_HID * MLP_IN; k += PRED_THREADS)
    sW1[(k / MLP_IN) * (MLP_IN + 1) + (k % MLP_IN)] = bf16_round(params[MLP_W1 + k]);
  for (int k = t; k < MLP_HID; k += PRED_THREADS) {
    sB[k] = bf16_round(params[MLP_B1 + k]);
    sB[MLP_HID + k] = bf16_round(params[MLP_B2 + k]);
  }
  for (int k = t; k < MLP_OUT * MLP_HID + MLP_OUT; k += PRED_THREADS)
    sB[2 * MLP_HID + k] = bf16_round(params[MLP_W3 + k]);  // w3 rows, then b3
  __syncthreads();
  const int lane = t & (WAVE - 1);
  float w1[MLP_IN], w2[MLP_HID];
#pragma unroll
  for (int k = 0; k < MLP_IN; ++k) w1[k] = sW1[lane * (MLP_IN + 1) + k];
#pragma unroll
  for (int k = 0; k < MLP_HID; ++k) w2[k] = sW2[lane * W2_PAD + k];
  const float b1 = sB[lane], b2 = sB[MLP_HID + lane];
  const float w30 = sB[2 * MLP_HID + lane], w31 = sB[2 * MLP_HID + MLP_HID + lane];
  const float b30 = sB[2 * MLP_HID + 2 * MLP_HID], b31 = sB[2 * MLP_HID + 2 * MLP_HID + 1];
  const int nwaves = gridDim.x * (PRED_THREADS / WAVE);
  for (int tile = blockIdx.x * (PRED_THREADS / WAVE) + t / WAVE; tile < T; tile += nwaves) {
    const uint32_t start = ranges[2 * tile], end = ranges[2 * tile + 1];
    const uint32_t len = end - start;
    const int n = (int)min(len, (uint32_t)FEAT_SAMPLE);
    float la[2], d[2];
    int bin[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int e = lane + h * WAVE;
      la[h] = 0.0f;
      d[h] = 0.0f;
      bin[h] = -1;
      if (e < n) {
        const uint32_t g = vals[start + e];
        const float* gr = geom + 8 * (size_t)g;
        const uint32_t rmin = f2u(gr[6]), rmax = f2u(gr[7]);
        const int area = (int)((rmax & 0xffffu) - (rmin & 0xffffu)) * (int)((rmax >> 16) - (rmin >> 16));
        la[h] = log2f((float)max(area, 1));
        const float* sp = splat + 12 * (size_t)g;
        bin[h] = min((int)(exp2f(sp[5]) * 16.0f), 15);
        d[h] = sp[6];
      }
    }
    const bool v0 = lane < n, v1 = lane + WAVE < n;
    const float S = wave_sum(la[0] + la[1]);
    float M = fmaxf(la[0], la[1]);
    float dn = fminf(v0 ? d[0] : 3.0e38f, v1 ? d[1] : 3.0e38f);
    float dx = fmaxf(d[0], d[1]);
    for (int o = 32; o > 0; o >>= 1) {
      M = fmaxf(M, __shfl_xor(M, o, WAVE));
      dn = fminf(dn, __shfl_xor(dn, o, WAVE));
      dx = fmaxf(dx, __shfl_xor(dx, o, WAVE));
    }
    float x[MLP_IN];
    x[0] = log2f(1.0f + (float)len);
    x[1] = n > 0 ? S / (float)n : 0.0f;
    x[2] = M;
    // opacity quantiles from the 16-bin histogram (ballot counts)
    const float qs[4] = {0.25f, 0.5f, 0.75f, 1.0f};
    uint32_t cum = 0, hb = 0;
    {
      int b = 0;
      hb = (uint32_t)(__popcll(__ballot(bin[0] == 0)) + __popcll(__ballot(bin[1] == 0)));
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t need = n > 0 ? (uint32_t)ceilf(qs[k] * (float)n) : 0u;
        while (b < 15 && cum + hb < need) {
          cum += hb;
          ++b;
          hb = (uint32_t)(__popcll(__ballot(bin[0] == b)) + __popcll(__ballot(bin[1] == b)));
        }
        x[3 + k] = n > 0 ? ((float)b + 0.5f) / 16.0f : 0.0f;
      }
    }
    x[7] = (n > 0 && dn > 0.0f) ? log2f(dx / dn) : 0.0f;
    if (feat && lane < MLP_IN) {
      float fv = x[0];
#pragma unroll
      for (int k = 1; k < MLP_IN; ++k) fv = lane == k ? x[k] : fv;
      feat[8 * (size_t)tile + lane] = fv;
    }
    // the network (k_mlp_infer<true>)
    float z1 = b1;
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) z1 = fmaf(w1[k], bf16_round(x[k]), z1);
    const float h1 = bf16_round(fmaxf(z1, 0.0f));
    float z2 = b2;
#pragma unroll
    for (int k = 0; k < MLP_HID; ++k)
      z2 = fmaf(w2[k], __int_as_float(__builtin_amdgcn_readlane(__float_as_int(h1), k)), z2);
    const float h2 = bf16_round(fmaxf(z2, 0.0f));
    const float o0 = wave_sum(w30 * h2), o1 = wave_sum(w31 * h2);
    if (lane == 0) {
      pred[2 * tile] = o0 + b30;
      pred[2 * tile + 1] = o1 + b31;
    }
  }
}

hipError_t launch_tile_predict(int T, const uint32_t* ranges, const uint32_t* vals,
                               const float* geom, const float* splat, const float* params,
                               float* feat, float* pred, hipStream_t s) {
  if (T <= 0) return hipSuccess;
  const int want = (T + PRED_THREADS / WAVE - 1) / (PRED_THREADS / WAVE);  // a tile per wave
  const unsigned blocks = (unsigned)(want > PRED_BLOCKS_MAX ? PRED_BLOCKS_MAX : want);
  FGS_TIMED("tile_predict", s,
            (k_tile_predict<<<blocks, PRED_THREADS, 0, s>>>(T, ranges, vals, geom, splat, params,
                                                            feat, pred)));
  return hipGetLastError();
}

// Training rows: fp32 forward, loss and backward to the pre-activation
// gradients, one thread per row in one-wave workgroups (the row's h1 and gz2
// live in LDS, column-major per wave, so no per-thread arrays spill). Per row
// r the workspace keeps [h1 | h2 | gz1 | gz2 | gy] (post-ReLU activations,
// pre-activation gradients, output gradients) for the weight-gradient
// reductions. Loss per row:
//   (y0 - t0)^2 + split_weight * BCE(sigmoid(y1), t1),  averaged over rows.
constexpr int ROWS_THREADS = WAVE;
__global__ __launch_bounds__(ROWS_THREADS) void k_mlp_rows(const float* __restrict__ params,
                                                           const float* __restrict__ x,
                                                           const float* __restrict__ target,
                                                           int64_t n, float split_weight,
                                                           float* __restrict__ act,
                                                           float* __restrict__ row_loss) {
  __shared__ float w[MLP_NPARAMS];
  __shared__ float h1s[MLP_HID][ROWS_THREADS];
  __shared__ float gz2s[MLP_HID][ROWS_THREADS];
  const int t = threadIdx.x;
  for (int k = t; k < MLP_NPARAMS; k += ROWS_THREADS) w[k] = params[k];
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * ROWS_THREADS + t;
  if (r >= n) return;  // no barrier below
  float* out = act + r * MLP_ACT_STRIDE;
  float xi[MLP_IN];
#pragma unroll
  for (int k = 0; k < MLP_IN; ++k) xi[k] = x[r * MLP_IN + k];
  for (int j = 0; j < MLP_HID; ++j) {
    float z = w[MLP_B1 + j];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) z = fmaf(w[MLP_W1 + j * MLP_IN + k], xi[k], z);
    z = fmaxf(z, 0.0f);
    h1s[j][t] = z;
    out[j] = z;
  }
  float o0 = w[MLP_B3], o1 = w[MLP_B3 + 1];
  for (int j = 0; j < MLP_HID; ++j) {
    float z = w[MLP_B2 + j];
#pragma unroll 8
    for (int k = 0; k < MLP_HID; ++k) z = fmaf(w[MLP_W2 + j * MLP_HID + k], h1s[k][t], z);
    z = fmaxf(z, 0.0f);
    gz2s[j][t] = z;  // h2 for now
    out[MLP_HID + j] = z;
    o0 = fmaf(w[MLP_W3 + j], z, o0);
    o1 = fmaf(w[MLP_W3 + MLP_HID + j], z, o1);
  }
  const float t0 = target[2 * r], t1 = target[2 * r + 1];
  const float inv_n = 1.0f / (float)n;
  const float d0 = o0 - t0;
  const float sig = 1.0f / (1.0f + expf(-o1));
  const float bce = fmaxf(o1, 0.0f) - o1 * t1 + log1pf(expf(-fabsf(o1)));
  row_loss[r] = d0 * d0 + split_weight * bce;
  const float g0 = 2.0f * d0 * inv_n, g1 = split_weight * (sig - t1) * inv_n;
  for (int j = 0; j < MLP_HID; ++j) {
    const float g = fmaf(w[MLP_W3 + j], g0, w[MLP_W3 + MLP_HID + j] * g1);
    const float gz = gz2s[j][t] > 0.0f ? g : 0.0f;
    gz2s[j][t] = gz;
    out[3 * MLP_HID + j] = gz;
  }
  for (int k = 0; k < MLP_HID; ++k) {
    float g = 0.0f;
#pragma unroll 8
    for (int j = 0; j < MLP_HID; ++j) g = fmaf(w[MLP_W2 + j * MLP_HID + k], gz2s[j][t], g);
    out[2 * MLP_HID + k] = h1s[k][t] > 0.0f ? g : 0.0f;  // gz1
  }
  out[4 * MLP_HID] = g0;
  out[4 * MLP_HID + 1] = g1;
}

// Weight gradients: parameter p sums its products over one chunk of rows
// (partial[chunk][p]); a second pass adds the chunks in order. Fixed order
// throughout, so the gradients are deterministic.
constexpr int MLP_CHUNK = 1024;
__device__ __forceinline__ void param_factors(int p, int& a_off, int& b_off, bool& bias) {
  // returns offsets inside a row's [x | act] view: grad = sum_r A[r] * B[r]
  // (B == 1 for biases)
  bias = false;
  if (p < MLP_B1) {  // W1[j][k] = gz1[j] * x[k]
    a_off = MLP_XOFF_GZ1 + p / MLP_IN;
    b_off = MLP_XOFF_X + p % MLP_IN;
  } else if (p < MLP_W2) {
    a_off = MLP_XOFF_GZ1 + (p - MLP_B1);
    bias = true;
  } else if (p < MLP_B2) {  // W2[j][k] = gz2[j] * h1[k]
    const int q = p - MLP_W2;
    a_off = MLP_XOFF_GZ2 + q / MLP_HID;
    b_off = MLP_XOFF_H1 + q % MLP_HID;
  } else if (p < MLP_W3) {
    a_off = MLP_XOFF_GZ2 + (p - MLP_B2);
    bias = true;
  } else if (p < MLP_B3) {  // W3[o][k] = gy[o] * h2[k]
    const int q = p - MLP_W3;
    a_off = MLP_XOFF_GY + q / MLP_HID;
    b_off = MLP_XOFF_H2 + q % MLP_HID;
  } else {
    a_off = MLP_XOFF_GY + (p - MLP_B3);
    bias = true;
  }
}

__device__ __forceinline__ float row_value(const float* __restrict__ x,
                                           const float* __restrict__ act, int64_t r, int off) {
  return off < MLP_XOFF_H1 ? x[r * MLP_IN + (off - MLP_XOFF_X)]
                           : act[r * MLP_ACT_STRIDE + (off - MLP_XOFF_H1)];
}

__global__ __launch_bounds__(256) void k_mlp_wgrad_partial(const float* __restrict__ x,
                                                           const float* __restrict__ act,
                                                           int64_t n,
                                                           float* __restrict__ partial) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= MLP_NPARAMS) return;
  const int64_t r0 = (int64_t)blockIdx.y * MLP_CHUNK;
  const int64_t r1 = min(n, r0 + MLP_CHUNK);
  int a_off, b_off = 0;
  bool bias;
  param_factors(p, a_off, b_off, bias);
  float s = 0.0f;
  for (int64_t r = r0; r < r1; ++r) {
    const float av = row_value(x, act, r, a_off);
    s = bias ? s + av : fmaf(av, row_value(x, act, r, b_off), s);
  }
  partial[(size_t)blockIdx.y * MLP_NPARAMS + p] = s;
}

__global__ __launch_bounds__(256) void k_mlp_wgrad_final(const float* __restrict__ partial,
                                                         int chunks, const float* __restrict__ row_loss,
                                                         int64_t n, float* __restrict__ grad,
                                                         float* __restrict__ loss) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p < MLP_NPARAMS) {
    float s = 0.0f;
    for (int c = 0; c < chunks; ++c) s += partial[(size_t)c * MLP_NPARAMS + p];
    grad[p] = s;
  }
  if (blockIdx.x == 0 && loss) {  // mean row loss, fixed-order block sum
    float s = 0.0f;
    for (int64_t r = threadIdx.x; r < n; r += 256) s += row_loss[r];
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = red[0] / (float)(n > 0 ? n : 1);
  }
}

size_t mlp_grad_ws_bytes(int64_t n) {
  const int64_t chunks = (n + MLP_CHUNK - 1) / MLP_CHUNK;
  return align_up((size_t)(n > 0 ? n : 1) * MLP_ACT_STRIDE * 4) +
         align_up((size_t)(n > 0 ? n : 1) * 4) +
         align_up((size_t)(chunks > 0 ? chunks : 1) * MLP_NPARAMS * 4);
}

hipError_t launch_mlp_loss_grad(const float* params, const float* x, const float* target,
                                int64_t n, float split_weight, float* grad, float* loss, void* ws,
                                hipStream_t s) {
  if (n <= 0) return hipSuccess;
  char* base = (char*)ws;
  float* act = (float*)base;
  float* row_loss = (float*)(base + align_up((size_t)n * MLP_ACT_STRIDE * 4));
  float* partial = (float*)((char*)row_loss + align_up((size_t)n * 4));
  const int chunks = (int)((n + MLP_CHUNK - 1) / MLP_CHUNK);
  FGS_TIMED("mlp_rows", s,
            (k_mlp_rows<<<(unsigned)((n + ROWS_THREADS - 1) / ROWS_THREADS), ROWS_THREADS, 0,
                          s>>>(params, x, target, n, split_weight, act, row_loss)));
  const unsigned pb = (MLP_NPARAMS + 255) / 256;
  FGS_TIMED("mlp_wgrad", s,
            (k_mlp_wgrad_partial<<<dim3(pb, chunks), 256, 0, s>>>(x, act, n, partial)));
  FGS_TIMED("mlp_wgrad_final", s,
            (k_mlp_wgrad_final<<<pb, 256, 0, s>>>(partial, chunks, row_loss, n, grad, loss)));
  return hipGetLastError();
}

// Adam with bias correction (the update rule of torch.optim.Adam, no weight
// decay).
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = fmaf(b1, m[i], (1.0f - b1) * gi);
  const float vi = fmaf(b2, v[i], (1.0f - b2) * gi * gi);
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr,
                       float b1, float b2, float eps, int step, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  FGS_TIMED("adam", s,
            (k_adam<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, g, m, v, n, lr, b1, b2, eps,
                                                                 c1, c2)));
  return hipGetLastError();
}

// ------------------------------------------------------------ work list
// Work list of the learned schedule, built by one 1024-thread workgroup:
// tiles binned by predicted log2 cost (bins of 1/8, 256 bins), highest bin
// first (longest predicted first); a tile is split into its two row bands
// when its split logit is > 0 and it has more than one bucket (ranges NULL:
// a list built before the frame's lists exist, on the logit alone). The order
// inside a bin is arbitrary (LDS atomics); outputs never depend on it.
constexpr int ITEM_THREADS = 1024;
__global__ __launch_bounds__(ITEM_THREADS) void k_sched_items(
    int T, const uint32_t* __restrict__ ranges, int bucket, const float* __restrict__ pred,
    uint32_t* __restrict__ items, uint32_t* __restrict__ n_items) {
  __shared__ uint32_t cnt[256];
  __shared__ uint32_t off[256];
  const int t = threadIdx.x;
  auto split_of = [&](int tile) -> bool {  // without ranges: the split flag alone
    const uint32_t len = ranges ? ranges[2 * tile + 1] - ranges[2 * tile] : 0xffffffffu;
    return pred[2 * tile + 1] > 0.0f && len > (uint32_t)bucket;
  };
  auto bin_of = [&](int tile) -> int {
    const float c = pred[2 * tile] * 8.0f;
    const int b = c != c ? 0 : (int)fminf(fmaxf(c, 0.0f), 255.0f);
    return 255 - b;  // descending cost
  };
  if (t < 256) cnt[t] = 0;
  __syncthreads();
  for (int tile = t; tile < T; tile += ITEM_THREADS)
    atomicAdd(&cnt[bin_of(tile)], split_of(tile) ? 2u : 1u);
  __syncthreads();
  {
    uint32_t tot;
    const uint32_t v = t < 256 ? cnt[t] : 0u;
    const uint32_t ex = block_excl_scan<ITEM_THREADS>(v, tot);  // every thread takes part
    if (t < 256) off[t] = ex;
    if (t == 0) n_items[0] = tot;
  }
  __syncthreads();
  for (int tile = t; tile < T; tile += ITEM_THREADS) {
    const bool sp = split_of(tile);
    const uint32_t pos = atomicAdd(&off[bin_of(tile)], sp ? 2u : 1u);
    if (sp) {
      items[pos] = ((uint32_t)tile << 2) | 1u;
      items[pos + 1] = ((uint32_t)tile << 2) | 2u;
    } else {
      items[pos] = (uint32_t)tile << 2;
    }
  }
}

hipError_t launch_sched_items(int T, const uint32_t* ranges, int bucket, const float* pred,
                              uint32_t* items, uint32_t* n_items, hipStream_t s) {
  if (T <= 0) return hipSuccess;
  FGS_TIMED("sched_items", s,
            (k_sched_items<<<1, ITEM_THREADS, 0, s>>>(T, ranges, bucket, pred, items, n_items)));
  return hipGetLastError();
}

}  // namespace fgs
