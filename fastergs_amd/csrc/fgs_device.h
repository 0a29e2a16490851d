// Device-side helpers for gfx950 (wave64): DPP reductions, lane masks,
// block scans. Only included by .hip translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fgs_math.h"

namespace fgs {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Number of set bits of `mask` in lanes below the calling lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ float dpp_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL,
                                                               ROW_MASK, BANK_MASK, false));
}
template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, BANK_MASK, false);
}

// Sum over the 64 lanes with DPP (quad xor 1, quad xor 2, half-row mirror,
// row mirror, then row_bcast 15 / 31 across rows). The full sum lands in
// lane 63; wave_sum() broadcasts it as a wave-uniform value.
__device__ __forceinline__ float wave_sum_to_63(float v) {
  v += dpp_f<0xb1>(v);        // quad_perm [1,0,3,2]
  v += dpp_f<0x4e>(v);        // quad_perm [2,3,0,1]
  v += dpp_f<0x141>(v);       // row_half_mirror
  v += dpp_f<0x140>(v);       // row_mirror
  v += dpp_f<0x142, 0xa>(v);  // row_bcast:15 -> rows 1,3
  v += dpp_f<0x143, 0xc>(v);  // row_bcast:31 -> rows 2,3
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_readlane(__builtin_bit_cast(int, wave_sum_to_63(v)), 63));
}

__device__ __forceinline__ int wave_max_i(int v) {
  v = max(v, dpp_i<0xb1>(v));
  v = max(v, dpp_i<0x4e>(v));
  v = max(v, dpp_i<0x141>(v));
  v = max(v, dpp_i<0x140>(v));
  v = max(v, dpp_i<0x142, 0xa>(v));
  v = max(v, dpp_i<0x143, 0xc>(v));
  return __builtin_amdgcn_readlane(v, 63);
}

// Transposing sum of 32 per-lane values over the wave: v[k] of every lane
// is summed into one lane pair, and lane l returns the total of v[l >> 1]
// (both lanes of the pair hold it). Each step exchanges half of the values
// with a partner lane and adds the halves that stay: lane ^ 32 and lane ^ 16
// with v_permlane32_swap / v_permlane16_swap (gfx950), then lane ^ 15
// (row_mirror), lane ^ 7 (row_half_mirror) and lane ^ 2 (quad_perm) with a
// select of the kept half, and a final add over lane ^ 1. The partner masks
// span all six lane bits, so every total covers all 64 lanes; 70 VALU
// operations per call instead of 32 wave sums.
__device__ __forceinline__ float f_of(uint32_t u) { return __builtin_bit_cast(float, u); }
__device__ __forceinline__ uint32_t u_of(float f) { return __builtin_bit_cast(uint32_t, f); }

__device__ __forceinline__ float wave_transpose_sum32(const float (&v)[32]) {
  const int l = lane_id();
  float a[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {  // lanes < 32 keep v[j], lanes >= 32 keep v[j + 16]
    const auto s = __builtin_amdgcn_permlane32_swap(u_of(v[j]), u_of(v[j + 16]), false, false);
    a[j] = f_of(s[0]) + f_of(s[1]);
  }
  float b[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {  // even rows keep a[j], odd rows a[j + 8]
    const auto s = __builtin_amdgcn_permlane16_swap(u_of(a[j]), u_of(a[j + 8]), false, false);
    b[j] = f_of(s[0]) + f_of(s[1]);
  }
  const bool h3 = (l & 8) != 0, h2 = (l & 4) != 0, h1 = (l & 2) != 0;
  float c[4];
#pragma unroll
  for (int j = 0; j < 4; ++j)  // bit 3 clear keeps b[j], set keeps b[j + 4]
    c[j] = (h3 ? b[j + 4] : b[j]) + dpp_f<0x140>(h3 ? b[j] : b[j + 4]);
  float d[2];
#pragma unroll
  for (int j = 0; j < 2; ++j)  // bit 2 clear keeps c[j], set keeps c[j + 2]
    d[j] = (h2 ? c[j + 2] : c[j]) + dpp_f<0x141>(h2 ? c[j] : c[j + 2]);
  const float e = (h1 ? d[1] : d[0]) + dpp_f<0x4e>(h1 ? d[0] : d[1]);
  return e + dpp_f<0xb1>(e);
}

// Inclusive prefix sum over the wave (shuffle based; used by scans only).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const uint32_t u = __shfl_up(v, d, WAVE);
    if (l >= d) v += u;
  }
  return v;
}

// Exclusive scan over a block of NT threads (NT a multiple of 64, <= 1024).
// Returns the exclusive prefix of `v`; `total` receives the block sum.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t& total) {
  __shared__ uint32_t wsum[NT / WAVE];
  const int w = threadIdx.x / WAVE;
  const uint32_t inc = wave_incl_scan(v);
  if (lane_id() == WAVE - 1) wsum[w] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < NT / WAVE; ++k) {
    const uint32_t s = wsum[k];
    before += (k < w) ? s : 0u;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return before + inc - v;
}

__device__ __forceinline__ void atomic_add_f(float* p, float v) { unsafeAtomicAdd(p, v); }

// Words handed from every workgroup of a grid to its last one travel
// through agent-scope atomics (performed where all XCDs see them; a fence
// here would write back the whole L2 of the XCD, which is slow after large
// writes). publish_u32 swaps the word in and waits until it is done;
// read_published_u32 reads such a word.
__device__ __forceinline__ void publish_u32(uint32_t* p, uint32_t v) {
  uint32_t old = __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  asm volatile("" : "+v"(old));  // the swap has returned, so it has been performed
}
__device__ __forceinline__ uint32_t read_published_u32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// True (in every thread) in the last workgroup of the grid to call it, once
// every workgroup has; the words every workgroup published before its call
// are then visible to read_published_u32. The counter is zero on entry and
// reset by the last workgroup, so the next launch can use it again. Every
// workgroup of the grid must call it once.
__device__ __forceinline__ bool last_workgroup(uint32_t* counter) {
  __shared__ uint32_t s_last;
  __syncthreads();  // every thread's publishes have been performed
  if (threadIdx.x == 0) {
    const uint32_t n = gridDim.x * gridDim.y * gridDim.z;
    const uint32_t prev = atomicAdd(counter, 1u);
    s_last = prev == n - 1 ? 1u : 0u;
    if (s_last) atomicExch(counter, 0u);
  }
  __syncthreads();
  return s_last != 0u;
}

// Exclusive prefix of f(i) over i < n into out[0..n), total into out[n], by
// one workgroup of NT threads (each sums a contiguous chunk).
template <int NT, class F>
__device__ __forceinline__ void block_scan_into(int n, F f, uint32_t* out) {
  const int per = (n + NT - 1) / NT;
  const int i0 = min((int)threadIdx.x * per, n), i1 = min(i0 + per, n);
  uint32_t s = 0;
  for (int i = i0; i < i1; ++i) s += f(i);
  uint32_t tot;
  uint32_t run = block_excl_scan<NT>(s, tot);
  for (int i = i0; i < i1; ++i) {
    const uint32_t v = f(i);
    out[i] = run;
    run += v;
  }
  if (threadIdx.x == 0) out[n] = tot;
}

// Element count of a stage whose size is produced on the device: the value
// at n_dev (when given) bounded by the host-side capacity cap the launch
// grid and the buffers were sized for.
__device__ __forceinline__ int64_t dev_count(const uint32_t* n_dev, int64_t cap) {
  return n_dev ? min((int64_t)*n_dev, cap) : cap;
}

}  // namespace fgs
