// Device-wide exclusive scan and stable LSD radix sort for gfx950.
//
// Scan: reduce-then-scan in three launches (per-tile sums, one-workgroup scan
// of the sums, per-tile scan with the carried offset). Used for the pair
// offsets (replacing the reference's warp-aggregated atomics, which give a
// nondeterministic pair order), per-tile bucket offsets (render.cu:420) and
// the radix histogram.
//
// Radix sort: 8-bit digits, three launches per pass (per-tile digit
// histogram in LDS, scan of the digit-major histogram, stable scatter). The
// scatter ranks equal digits inside a wave64 with eight ballots (a wave-level
// multisplit) and across the four waves of a workgroup through LDS, so the
// sort is stable. It replaces cub::DeviceRadixSort::SortPairs on bits
// [0, 32 + msb(tiles)) (FGS/csrc/cuda_rasterizer/sort.cu:28-44).
#include <hip/hip_runtime.h>

#include "fgs_device.h"
#include "fgs_layout.h"
#include "fgs_kernels.h"
#include "fgs_prof.h"

namespace fgs {

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                              int64_t n,
                                                              uint32_t* __restrict__ partial) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t i = base + k * SCAN_THREADS + threadIdx.x;
    if (i < n) s += in[i];
  }
  uint32_t tot;
  block_excl_scan<SCAN_THREADS>(s, tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_partials(uint32_t* __restrict__ partial, int nb,
                                                        uint32_t* __restrict__ total) {
  uint32_t carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + threadIdx.x;
    const uint32_t v = i < nb ? partial[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan<1024>(v, tot);
    if (i < nb) partial[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    partial[nb] = carry;
    if (total) *total = carry;
  }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const uint32_t* in, uint32_t* out,
                                                             int64_t n,
                                                             const uint32_t* __restrict__ partial) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan<SCAN_THREADS>(s, tot) + partial[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

// Small inputs (per-tile counts): one 1024-thread workgroup, one launch
// instead of three. Limited to one chunk of 8192: a single workgroup walking
// longer inputs chunk by chunk is latency-bound on its one CU and slower
// than the three-launch scan.
constexpr int SCAN1_THREADS = 1024;
constexpr int SCAN1_ITEMS = 8;
constexpr int64_t SCAN1_MAX = SCAN1_THREADS * SCAN1_ITEMS;
__global__ __launch_bounds__(SCAN1_THREADS) void k_scan_single(const uint32_t* in, uint32_t* out,
                                                               int64_t n, uint32_t* total) {
  uint32_t carry = 0;
  for (int64_t base = 0; base < n; base += SCAN1_THREADS * SCAN1_ITEMS) {
    const int64_t b = base + (int64_t)threadIdx.x * SCAN1_ITEMS;
    uint32_t v[SCAN1_ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN1_ITEMS; ++k) {
      v[k] = b + k < n ? in[b + k] : 0u;
      sum += v[k];
    }
    uint32_t tot;
    uint32_t run = carry + block_excl_scan<SCAN1_THREADS>(sum, tot);
#pragma unroll
    for (int k = 0; k < SCAN1_ITEMS; ++k) {
      if (b + k < n) out[b + k] = run;
      run += v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

hipError_t scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, void* ws,
                    hipStream_t s) {
  if (n <= SCAN1_MAX) {
    k_scan_single<<<1, SCAN1_THREADS, 0, s>>>(in, out, n, total);
    return hipGetLastError();
  }
  const int nb = (int)((n + SCAN_TILE - 1) / SCAN_TILE);
  uint32_t* partial = (uint32_t*)ws;
  if (nb > 0) k_scan_reduce<<<nb, SCAN_THREADS, 0, s>>>(in, n, partial);
  k_scan_partials<<<1, 1024, 0, s>>>(partial, nb, total);
  if (nb > 0) k_scan_apply<<<nb, SCAN_THREADS, 0, s>>>(in, out, n, partial);
  return hipGetLastError();
}

// ------------------------------------------------------------------ radix
__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const uint64_t* __restrict__ keys,
                                                             int64_t n_cap,
                                                             const uint32_t* __restrict__ n_dev,
                                                             int shift, uint32_t mask,
                                                             uint32_t* __restrict__ hist, int nb) {
  __shared__ uint32_t cnt[256];
  const int64_t n = dev_count(n_dev, n_cap);
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * SORT_TILE;
#pragma unroll 4
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const int64_t i = base + r * SORT_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(uint32_t)(keys[i] >> shift) & mask], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter of one 4096-element tile: the block first sorts its
// elements by digit in LDS (stable ranks from wave ballots + per-wave digit
// counts), then writes each digit's run contiguously to its global offset,
// so stores are coalesced runs instead of one element per digit per wave.
// SPLIT (key-only sorts): write the low and high 32-bit halves of each key
// to two arrays instead of the key (used by the last pass of the tile sort to
// emit the per-tile lists and the tile ids directly).
template <bool VALS, bool SPLIT = false>
__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n_cap, const uint32_t* __restrict__ n_dev, int shift,
    uint32_t mask, const uint32_t* __restrict__ hist, int nb, uint32_t* __restrict__ lo_out = nullptr,
    uint32_t* __restrict__ hi_out = nullptr) {
  constexpr int NW = SORT_THREADS / WAVE;
  __shared__ uint64_t sk[SORT_TILE];
  __shared__ uint32_t sv[VALS ? SORT_TILE : 1];
  __shared__ uint32_t gofs[256];
  __shared__ uint32_t bstart[256];
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[NW][256];
  const int t = threadIdx.x;
  const int w = t / WAVE;
  const int64_t base = (int64_t)blockIdx.x * SORT_TILE;
  const int64_t n = dev_count(n_dev, n_cap);
  if (base >= n) return;  // whole workgroup: past the device-side count
  const int nvalid = (int)min((int64_t)SORT_TILE, n - base);
  uint64_t k[SORT_ITEMS];
  uint32_t v[SORT_ITEMS];
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const int e = r * SORT_THREADS + t;
    k[r] = e < nvalid ? kin[base + e] : 0ull;
    v[r] = (VALS && e < nvalid) ? vin[base + e] : 0u;
  }
  run[t] = 0;
#pragma unroll
  for (int q = 0; q < NW; ++q) wcnt[q][t] = 0;
  gofs[t] = hist[(size_t)t * nb + blockIdx.x];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r)
    if (r * SORT_THREADS + t < nvalid) atomicAdd(&run[(uint32_t)(k[r] >> shift) & mask], 1u);
  __syncthreads();
  uint32_t tot;
  const uint32_t my = run[t];
  const uint32_t ex = block_excl_scan<SORT_THREADS>(my, tot);
  bstart[t] = ex;
  run[t] = 0;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const bool valid = r * SORT_THREADS + t < nvalid;
    const uint32_t d = (uint32_t)(k[r] >> shift) & mask;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t wrank = lanes_below(peers);
    if (valid && wrank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = bstart[d] + run[d] + wrank;
#pragma unroll
      for (int q = 0; q < NW; ++q) pos += (q < w) ? wcnt[q][d] : 0u;
      sk[pos] = k[r];
      if (VALS) sv[pos] = v[r];
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      add += wcnt[q][t];
      wcnt[q][t] = 0;
    }
    run[t] += add;
    __syncthreads();
  }
  for (int e = t; e < nvalid; e += SORT_THREADS) {
    const uint64_t key = sk[e];
    const uint32_t d = (uint32_t)(key >> shift) & mask;
    const uint32_t dst = gofs[d] + (uint32_t)e - bstart[d];
    if (SPLIT) {
      lo_out[dst] = (uint32_t)key;
      hi_out[dst] = (uint32_t)(key >> 32);
    } else {
      kout[dst] = key;
    }
    if (VALS) vout[dst] = sv[e];
  }
}

hipError_t sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                      int64_t n, int begin_bit, int end_bit, void* ws, hipStream_t s,
                      bool& in_alt, uint32_t* last_lo, uint32_t* last_hi, const uint32_t* n_dev) {
  in_alt = false;
  // with a device-side count the passes always run (the count may be 0 or 1)
  if ((n <= 1 && !n_dev) || n <= 0 || end_bit <= begin_bit) return hipSuccess;
  const int nb = (int)((n + SORT_TILE - 1) / SORT_TILE);
  uint32_t* hist = (uint32_t*)ws;
  void* scan_ws = (char*)ws + align_up((size_t)256 * nb * sizeof(uint32_t));
  uint64_t *ka = keys, *kb = keys_alt;
  uint32_t *va = vals, *vb = vals_alt;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    const int bits = end_bit - shift < 8 ? end_bit - shift : 8;
    const uint32_t mask = (1u << bits) - 1u;
    FGS_TIMED(vals ? "radix_hist_kv" : "radix_hist_k", s,
              (k_radix_hist<<<nb, SORT_THREADS, 0, s>>>(ka, n, n_dev, shift, mask, hist, nb)));
    hipError_t e = scan_u32(hist, hist, (int64_t)256 * nb, nullptr, scan_ws, s);
    if (e != hipSuccess) return e;
    const bool last = shift + 8 >= end_bit;
    if (vals)
      FGS_TIMED("radix_scatter_kv", s,
                (k_radix_scatter<true><<<nb, SORT_THREADS, 0, s>>>(ka, va, kb, vb, n, n_dev, shift,
                                                                    mask, hist, nb)));
    else if (last && last_lo)
      FGS_TIMED("radix_scatter_k_last", s,
                (k_radix_scatter<false, true><<<nb, SORT_THREADS, 0, s>>>(
                    ka, nullptr, kb, nullptr, n, n_dev, shift, mask, hist, nb, last_lo, last_hi)));
    else
      FGS_TIMED("radix_scatter_k", s,
                (k_radix_scatter<false><<<nb, SORT_THREADS, 0, s>>>(ka, nullptr, kb, nullptr, n,
                                                                     n_dev, shift, mask, hist, nb)));
    uint64_t* tk = ka; ka = kb; kb = tk;
    uint32_t* tv = va; va = vb; vb = tv;
    in_alt = !in_alt;
  }
  return hipGetLastError();
}

}  // namespace fgs
