// Device-wide exclusive scan and stable LSD radix sorts for gfx950.
//
// Scan: reduce-then-scan in two launches (per-tile sums; per-tile scan
// whose carried offset each workgroup adds up from the tile sums before it),
// a third launch scanning the sums beyond 4096 tiles, one launch up to 8192
// entries. Used for the pair offsets (replacing the reference's
// warp-aggregated atomics, which give a nondeterministic pair order), the
// large-rect compaction and per-tile bucket offsets (render.cu:420).
//
// Radix sort of (key, u32 value) pairs with 16-, 32- or 64-bit keys
// (replaces cub::DeviceRadixSort::SortPairs, FGS/csrc/cuda_rasterizer/
// sort.cu:28-44): LSD, digits of up to 8 bits spread evenly, stable; the
// ranking is compiled per digit width. Up to 4M keys every pass is one
// onesweep launch after one histogram read for all passes; above that each
// pass is a per-tile digit histogram, its scan and a scatter. Keys are
// as narrow as the data allows: the FAST tile sort moves 16-bit tile ids
// with 32-bit record values (6 B per pair per pass instead of a packed
// 8 B key, 2 B per pair for the histograms).
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

// Up to PARTIALS_FUSED_MAX tiles, the apply launch adds up the tile sums
// before its own tile itself (a strided sum and one block reduction)
// instead of a separate launch scanning them.
constexpr int PARTIALS_FUSED_MAX = 4096;
__device__ __forceinline__ uint32_t partials_before(const uint32_t* __restrict__ partial, int b) {
  uint32_t s = 0;
  for (int j = threadIdx.x; j < b; j += SCAN_THREADS) s += partial[j];
  uint32_t tot;
  block_excl_scan<SCAN_THREADS>(s, tot);
  return tot;
}

// FUSED: partial holds the tile sums (k_scan_partials not run); the last
// workgroup writes the total. Else partial holds their exclusive scan.
template <bool FUSED>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const uint32_t* in, uint32_t* out,
                                                             int64_t n,
                                                             const uint32_t* __restrict__ partial,
                                                             uint32_t* __restrict__ total) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  const uint32_t carry = FUSED ? partials_before(partial, blockIdx.x) : partial[blockIdx.x];
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan<SCAN_THREADS>(s, tot) + carry;
  if (FUSED && total && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) *total = carry + tot;
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

// Active buckets of the backward in one launch (up to SCAN1_MAX tiles):
// out[t] = exclusive prefix of ceil(maxc[t] / bucket), out[T] = the total.
__global__ __launch_bounds__(SCAN1_THREADS) void k_active_scan_single(const uint32_t* maxc, int T,
                                                                      int bucket,
                                                                      uint32_t* out) {
  const int b = threadIdx.x * SCAN1_ITEMS;
  uint32_t v[SCAN1_ITEMS];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN1_ITEMS; ++k) {
    v[k] = b + k < T ? (maxc[b + k] + (uint32_t)bucket - 1u) / (uint32_t)bucket : 0u;
    sum += v[k];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan<SCAN1_THREADS>(sum, tot);
#pragma unroll
  for (int k = 0; k < SCAN1_ITEMS; ++k) {
    if (b + k < T) out[b + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) out[T] = tot;
}

hipError_t active_bucket_scan(const uint32_t* maxc, int T, int bucket, uint32_t* counts,
                              uint32_t* out, void* ws, hipStream_t s) {
  if (T <= 0) return hipSuccess;
  if (T <= SCAN1_MAX) {
    FGS_TIMED("active_scan", s,
              (k_active_scan_single<<<1, SCAN1_THREADS, 0, s>>>(maxc, T, bucket, out)));
    return hipGetLastError();
  }
  hipError_t e = launch_active_counts(T, maxc, bucket, counts, s);
  return e != hipSuccess ? e : scan_u32(counts, out, T, out + T, ws, s);
}

hipError_t scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, void* ws,
                    hipStream_t s) {
  if (n <= SCAN1_MAX) {
    k_scan_single<<<1, SCAN1_THREADS, 0, s>>>(in, out, n, total);
    return hipGetLastError();
  }
  const int nb = (int)((n + SCAN_TILE - 1) / SCAN_TILE);
  uint32_t* partial = (uint32_t*)ws;
  k_scan_reduce<<<nb, SCAN_THREADS, 0, s>>>(in, n, partial);
  if (nb <= PARTIALS_FUSED_MAX) {
    k_scan_apply<true><<<nb, SCAN_THREADS, 0, s>>>(in, out, n, partial, total);
  } else {
    k_scan_partials<<<1, 1024, 0, s>>>(partial, nb, total);
    k_scan_apply<false><<<nb, SCAN_THREADS, 0, s>>>(in, out, n, partial, nullptr);
  }
  return hipGetLastError();
}

// ------------------------------------------------------ pair offsets
// Pair offsets of the FAST preprocess by depth rank, from the depth sort's
// output: keys[r] = depth bits << 32 | record and vals[r] = the record's
// pair count, or PAIRS_LARGE for a large rect, whose count the large-rect
// pass added to count[record] and whose rank is recorded in lrank[record].
// off[r] = exclusive prefix over ranks, counts[0] = the total, counts[1] =
// ceil(total / bucket) + T. Everything is read and written in rank order
// except the few large rects' entries. The first pass resolves the counts
// in place (vals) and sums tiles; the second scans them (one launch up to
// SCAN1_MAX records).
__device__ __forceinline__ uint32_t rank_count(const uint64_t* __restrict__ keys, uint32_t* vals,
                                               const uint32_t* __restrict__ count,
                                               uint32_t* __restrict__ lrank, int64_t r) {
  uint32_t v = vals[r];
  if (v & PAIRS_LARGE) {
    const uint32_t rec = (uint32_t)keys[r];
    lrank[rec] = (uint32_t)r;
    v = count[rec];
  }
  return v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_pscan_reduce(const uint64_t* __restrict__ keys,
                                                               uint32_t* __restrict__ vals,
                                                               const uint32_t* __restrict__ count,
                                                               uint32_t* __restrict__ lrank,
                                                               int64_t n,
                                                               uint32_t* __restrict__ partial) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t i = base + k * SCAN_THREADS + threadIdx.x;
    if (i < n) {
      const uint32_t v = rank_count(keys, vals, count, lrank, i);
      vals[i] = v;
      s += v;
    }
  }
  uint32_t tot;
  block_excl_scan<SCAN_THREADS>(s, tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_pscan_partials(uint32_t* __restrict__ partial, int nb,
                                                         uint32_t* __restrict__ counts, int bucket,
                                                         int T) {
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
    counts[0] = carry;
    counts[1] = (carry + bucket - 1) / bucket + (uint32_t)T;
  }
}

// FUSED: partial holds the tile sums; the last workgroup writes the total
// and the bucket-table bound. Else partial holds their exclusive scan.
template <bool FUSED>
__global__ __launch_bounds__(SCAN_THREADS) void k_pscan_apply(const uint32_t* __restrict__ vals,
                                                              int64_t n,
                                                              const uint32_t* __restrict__ partial,
                                                              uint32_t* __restrict__ off,
                                                              uint32_t* __restrict__ counts,
                                                              int bucket, int T) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  const uint32_t carry = FUSED ? partials_before(partial, blockIdx.x) : partial[blockIdx.x];
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = (base + k < n) ? vals[base + k] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan<SCAN_THREADS>(s, tot) + carry;
  if (FUSED && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) {
    const uint32_t all = carry + tot;
    counts[0] = all;
    counts[1] = (all + bucket - 1) / bucket + (uint32_t)T;
  }
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) off[base + k] = run;
    run += v[k];
  }
}

__global__ __launch_bounds__(SCAN1_THREADS) void k_pscan_single(const uint64_t* __restrict__ keys,
                                                                uint32_t* __restrict__ vals,
                                                                const uint32_t* __restrict__ count,
                                                                uint32_t* __restrict__ lrank,
                                                                int64_t n,
                                                                uint32_t* __restrict__ off,
                                                                uint32_t* __restrict__ counts,
                                                                int bucket, int T) {
  const int64_t b = (int64_t)threadIdx.x * SCAN1_ITEMS;
  uint32_t v[SCAN1_ITEMS];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN1_ITEMS; ++k) {
    v[k] = b + k < n ? rank_count(keys, vals, count, lrank, b + k) : 0u;
    sum += v[k];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan<SCAN1_THREADS>(sum, tot);
#pragma unroll
  for (int k = 0; k < SCAN1_ITEMS; ++k) {
    if (b + k < n) off[b + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) {
    counts[0] = tot;
    counts[1] = (tot + bucket - 1) / bucket + (uint32_t)T;
  }
}

hipError_t scan_pair_offsets(const uint64_t* keys, uint32_t* vals, const uint32_t* count,
                             uint32_t* lrank, int64_t n, uint32_t* off, uint32_t* counts,
                             int bucket, int T, void* ws, hipStream_t s) {
  if (n <= SCAN1_MAX) {
    k_pscan_single<<<1, SCAN1_THREADS, 0, s>>>(keys, vals, count, lrank, n, off, counts, bucket,
                                                T);
    return hipGetLastError();
  }
  const int nb = (int)((n + SCAN_TILE - 1) / SCAN_TILE);
  uint32_t* partial = (uint32_t*)ws;
  k_pscan_reduce<<<nb, SCAN_THREADS, 0, s>>>(keys, vals, count, lrank, n, partial);
  if (nb <= PARTIALS_FUSED_MAX) {
    k_pscan_apply<true><<<nb, SCAN_THREADS, 0, s>>>(vals, n, partial, off, counts, bucket, T);
  } else {
    k_pscan_partials<<<1, 1024, 0, s>>>(partial, nb, counts, bucket, T);
    k_pscan_apply<false><<<nb, SCAN_THREADS, 0, s>>>(vals, n, partial, off, counts, bucket, T);
  }
  return hipGetLastError();
}

// ------------------------------------------------------------ expansion
// Entry i with flag[i] = c > 0 (c <= LARGE_BANDS) becomes the c consecutive
// ids i * LARGE_BANDS + b, b < c, in index order; count = total ids. A scan
// whose last launch writes the ids directly (one launch up to 8192 entries,
// two beyond).
// FUSED: partial holds the tile sums; the last workgroup writes the count.
template <bool FUSED>
__global__ __launch_bounds__(SCAN_THREADS) void k_compact_apply(const uint32_t* __restrict__ flag,
                                                                int64_t n,
                                                                const uint32_t* __restrict__ partial,
                                                                uint32_t* __restrict__ ids,
                                                                uint32_t* __restrict__ count) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  const uint32_t carry = FUSED ? partials_before(partial, blockIdx.x) : partial[blockIdx.x];
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = base + k < n ? flag[base + k] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan<SCAN_THREADS>(s, tot) + carry;
  if (FUSED && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) *count = carry + tot;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    for (uint32_t b = 0; b < v[k]; ++b) ids[run + b] = (uint32_t)(base + k) * LARGE_BANDS + b;
    run += v[k];
  }
}

__global__ __launch_bounds__(SCAN1_THREADS) void k_compact_single(const uint32_t* __restrict__ flag,
                                                                  int64_t n,
                                                                  uint32_t* __restrict__ ids,
                                                                  uint32_t* __restrict__ count) {
  const int64_t b0 = (int64_t)threadIdx.x * SCAN1_ITEMS;
  uint32_t v[SCAN1_ITEMS];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN1_ITEMS; ++k) {
    v[k] = b0 + k < n ? flag[b0 + k] : 0u;
    sum += v[k];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan<SCAN1_THREADS>(sum, tot);
#pragma unroll
  for (int k = 0; k < SCAN1_ITEMS; ++k) {
    for (uint32_t b = 0; b < v[k]; ++b) ids[run + b] = (uint32_t)(b0 + k) * LARGE_BANDS + b;
    run += v[k];
  }
  if (threadIdx.x == 0) *count = tot;
}

hipError_t expand_flags(const uint32_t* flag, int64_t n, uint32_t* ids, uint32_t* count, void* ws,
                         hipStream_t s) {
  if (n <= SCAN1_MAX) {
    k_compact_single<<<1, SCAN1_THREADS, 0, s>>>(flag, n, ids, count);
    return hipGetLastError();
  }
  const int nb = (int)((n + SCAN_TILE - 1) / SCAN_TILE);
  uint32_t* partial = (uint32_t*)ws;
  // the plain per-tile sums of the flags and their scan apply
  k_scan_reduce<<<nb, SCAN_THREADS, 0, s>>>(flag, n, partial);
  if (nb <= PARTIALS_FUSED_MAX) {
    k_compact_apply<true><<<nb, SCAN_THREADS, 0, s>>>(flag, n, partial, ids, count);
  } else {
    k_scan_partials<<<1, 1024, 0, s>>>(partial, nb, count);
    k_compact_apply<false><<<nb, SCAN_THREADS, 0, s>>>(flag, n, partial, ids, count);
  }
  return hipGetLastError();
}

// ------------------------------------------------------- radix, classic
// Keys per thread (4096-key tiles for every key width: measured on the
// 21M-pair tile sort and the 1M-record depth sort, 8192-key tiles halve
// the occupancy and 2048-key tiles double the histogram and look-back work,
// both slower).
template <typename K>
struct SortTiling {
  static constexpr int ITEMS = SORT_ITEMS;
  static constexpr int TILE = SORT_THREADS * ITEMS;
  static_assert(TILE == SORT_TILE, "workspace sizes assume SORT_TILE-key tiles");
  static_assert(TILE >= SORT_MIN_TILE, "workspace rows are sized for SORT_MIN_TILE");
};
// Digit layout: ceil(bits / 8) passes with the key bits spread evenly over
// them (13 tile-id bits -> 7 + 6), which keeps each digit's run within a
// 4096-key tile long enough for full-line stores.
__host__ __device__ inline int sort_passes(int begin_bit, int end_bit) {
  return (end_bit - begin_bit + 7) / 8;
}
__host__ __device__ inline void pass_digits(int begin_bit, int end_bit, int p, int& shift,
                                            int& bits) {
  const int total = end_bit - begin_bit, np = (total + 7) / 8;
  const int w = total / np, extra = total % np;
  shift = begin_bit + p * w + (p < extra ? p : extra);
  bits = w + (p < extra ? 1 : 0);
}

// Up to 16 / sizeof(K) consecutive keys from e (one 16-byte load when all
// of them are below n and the start is 16-byte aligned, else one by one).
template <typename K>
struct KeyVec {
  static constexpr int N = 16 / sizeof(K);
  union {
    uint4 v;
    K k[N];
  };
};
template <typename K>
__device__ __forceinline__ int load_keys(const K* __restrict__ keys, int64_t e, int64_t n,
                                         KeyVec<K>& out) {
  constexpr int N = KeyVec<K>::N;
  if (e + N <= n && ((uintptr_t)(keys + e) & 15) == 0) {
    out.v = *(const uint4*)(keys + e);
    return N;
  }
  int c = 0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    out.k[j] = e + j < n ? keys[e + j] : (K)0;
    c += e + j < n ? 1 : 0;
  }
  return c;
}

template <typename K>
__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const K* __restrict__ keys,
                                                             int64_t n_cap,
                                                             const uint32_t* __restrict__ n_dev,
                                                             int shift, uint32_t mask,
                                                             uint32_t* __restrict__ hist, int nb) {
  constexpr int N = KeyVec<K>::N;
  __shared__ uint32_t cnt[256];
  const int64_t n = dev_count(n_dev, n_cap);
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * SortTiling<K>::TILE;
  for (int c = threadIdx.x; c < SortTiling<K>::TILE / N; c += SORT_THREADS) {
    const int64_t e = base + (int64_t)c * N;
    if (e >= n) break;
    KeyVec<K> kv;
    const int m = load_keys(keys, e, n, kv);
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (j < m) atomicAdd(&cnt[(uint32_t)(kv.k[j] >> shift) & mask], 1u);
  }
  __syncthreads();
  // digit-major rows of the pass's (mask + 1) digits only
  if (threadIdx.x <= mask) hist[(size_t)threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
}

// Register form of a tile's keys: 16-bit keys sit in the low half of a
// 32-bit register whose high half receives the key's rank (ranks are below
// 4096), which saves a register per item; wider keys keep a separate rank
// array.
template <typename K>
struct KeyReg {
  using T = K;
  static constexpr bool PACK = false;
};
template <>
struct KeyReg<uint16_t> {
  using T = uint32_t;
  static constexpr bool PACK = true;
};
template <typename K>
__device__ __forceinline__ uint32_t reg_digit(typename KeyReg<K>::T k, int shift, uint32_t mask) {
  if constexpr (KeyReg<K>::PACK) return ((k & 0xffffu) >> shift) & mask;
  else return (uint32_t)(k >> shift) & mask;
}

// Stable ranking of one 4096-key tile in LDS: wave w owns the keys
// [1024 w, 1024 w + 1024) as 16 rows of 64; equal digits in a row are ranked
// with ballots over the digit bits, and each wave keeps its own running
// count per digit in LDS (no barrier per row). Afterwards, per digit, the
// waves' counts become prefixes and the tile's digit offsets an exclusive
// scan; (wave, row, lane) order is index order, so the ranks are stable.
// Returns (after a barrier) with wcnt[w][d] = keys of digit d in waves < w,
// bstart[d] = the tile's keys of digits < d, and rk[r] = rank of key r
// within its (wave, digit).
template <typename K, int BITS>
__device__ __forceinline__ void rank_tile(typename KeyReg<K>::T (&k)[SortTiling<K>::ITEMS], int nvalid, int shift,
                                          uint32_t (&wcnt)[SORT_THREADS / WAVE][256],
                                          uint32_t* bstart, uint32_t (&rk)[SortTiling<K>::ITEMS],
                                          uint32_t& count) {
  constexpr int NW = SORT_THREADS / WAVE;
  constexpr int WITEMS = SortTiling<K>::TILE / NW;
  const int t = threadIdx.x, w = t / WAVE, lane = lane_id();
  constexpr uint32_t mask = (1u << BITS) - 1u;
#pragma unroll
  for (int r = 0; r < SortTiling<K>::ITEMS; ++r) {
    const bool valid = w * WITEMS + r * WAVE + lane < nvalid;
    const uint32_t d = reg_digit<K>(k[r], shift, mask);
    // lanes with the same digit: per bit, the ballot of the bit where this
    // lane's bit is set, its complement where it is clear
    const uint64_t vm = __ballot(valid);
    uint32_t plo = (uint32_t)vm, phi = (uint32_t)(vm >> 32);
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
      const uint32_t sb = (uint32_t)__builtin_amdgcn_sbfe((int)d, b, 1);  // all ones when set
      const uint64_t m = __ballot(sb != 0u);
      plo &= ~((uint32_t)m ^ sb);
      phi &= ~((uint32_t)(m >> 32) ^ sb);
    }
    const uint64_t peers = ((uint64_t)phi << 32) | plo;
    const uint32_t below = lanes_below(peers);
    const uint32_t before = wcnt[w][d];  // d is in range for every lane
    if constexpr (KeyReg<K>::PACK) k[r] |= (before + below) << 16;
    else rk[r] = before + below;
    if (valid && below == 0) wcnt[w][d] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  uint32_t c = 0;
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    const uint32_t x = wcnt[q][t];
    wcnt[q][t] = c;
    c += x;
  }
  uint32_t tot;
  bstart[t] = block_excl_scan<SORT_THREADS>(c, tot);
  count = c;
  __syncthreads();
}

// Loads this wave's 16 rows (keys and values) of the tile at base.
template <typename K, bool VALS>
__device__ __forceinline__ void load_tile(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                          int64_t base, int nvalid, typename KeyReg<K>::T (&k)[SortTiling<K>::ITEMS],
                                          uint32_t (&v)[SortTiling<K>::ITEMS]) {
  constexpr int WITEMS = SortTiling<K>::TILE / (SORT_THREADS / WAVE);
  const int w = threadIdx.x / WAVE, lane = lane_id();
  // every load issued unconditionally (index clamped to the last valid
  // key; such items are masked by nvalid afterwards)
#pragma unroll
  for (int r = 0; r < SortTiling<K>::ITEMS; ++r) {
    const int e = min(w * WITEMS + r * WAVE + lane, nvalid - 1);
    k[r] = (typename KeyReg<K>::T)kin[base + e];
    v[r] = VALS ? vin[base + e] : 0u;
  }
}

// Places the ranked keys (and values) in LDS in digit order, then writes
// each digit's run contiguously from gofs[d].
template <typename K, bool VALS>
__device__ __forceinline__ void place_and_write(const typename KeyReg<K>::T (&k)[SortTiling<K>::ITEMS], const uint32_t (&v)[SortTiling<K>::ITEMS],
                                                const uint32_t (&rk)[SortTiling<K>::ITEMS], int nvalid,
                                                int shift, uint32_t mask,
                                                uint32_t (&wcnt)[SORT_THREADS / WAVE][256],
                                                const uint32_t* bstart, const uint32_t* gofs,
                                                K* sk, uint32_t* sv, K* __restrict__ kout,
                                                uint32_t* __restrict__ vout) {
  constexpr int WITEMS = SortTiling<K>::TILE / (SORT_THREADS / WAVE);
  const int t = threadIdx.x, w = t / WAVE, lane = lane_id();
#pragma unroll
  for (int r = 0; r < SortTiling<K>::ITEMS; ++r) {
    if (w * WITEMS + r * WAVE + lane < nvalid) {
      const uint32_t d = reg_digit<K>(k[r], shift, mask);
      uint32_t rank;
      if constexpr (KeyReg<K>::PACK) rank = k[r] >> 16;
      else rank = rk[r];
      const uint32_t pos = bstart[d] + wcnt[w][d] + rank;
      sk[pos] = (K)k[r];
      if (VALS) sv[pos] = v[r];
    }
  }
  __syncthreads();
  for (int e = t; e < nvalid; e += SORT_THREADS) {
    const K key = sk[e];
    const uint32_t d = (uint32_t)(key >> shift) & mask;
    const uint32_t dst = gofs[d] + (uint32_t)e - bstart[d];
    kout[dst] = key;
    if (VALS) vout[dst] = sv[e];
  }
}

// Workgroup b runs on XCD b % 8 (round-robin dispatch over the 8 XCDs):
// tile of workgroup b such that each XCD takes a contiguous range of tiles,
// so the digit runs that consecutive tiles write next to each other meet in
// the same L2 and leave it as whole lines. The grid is 8 * ceil(nb / 8).
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb) {
  const int64_t q = (nb + 7) / 8;
  return (b % 8) * q + b / 8;
}
inline unsigned xcd_grid(int64_t nb) { return (unsigned)(8 * ((nb + 7) / 8)); }

template <typename K, bool VALS, int BITS>
__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin, K* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n_cap, const uint32_t* __restrict__ n_dev, int shift,
    const uint32_t* __restrict__ hist, int nb) {
  constexpr int NW = SORT_THREADS / WAVE;
  __shared__ K sk[SortTiling<K>::TILE];
  __shared__ uint32_t sv[VALS ? SortTiling<K>::TILE : 1];
  __shared__ uint32_t wcnt[NW][256];
  __shared__ uint32_t bstart[256];
  __shared__ uint32_t gofs[256];
  const int t = threadIdx.x;
  const int64_t n = dev_count(n_dev, n_cap);
  // the tiles holding keys (the device-side count may be far below the
  // capacity the grid is sized for), spread over all eight XCDs
  const int64_t nbe = min((int64_t)nb, (n + SortTiling<K>::TILE - 1) / SortTiling<K>::TILE);
  if ((int64_t)blockIdx.x >= 8 * ((nbe + 7) / 8)) return;  // whole workgroup
  const int64_t tile = xcd_tile(blockIdx.x, nbe);
  const int64_t base = tile * SortTiling<K>::TILE;
  if (tile >= nbe) return;  // whole workgroup
  const int nvalid = (int)min((int64_t)SortTiling<K>::TILE, n - base);
#pragma unroll
  for (int q = 0; q < NW; ++q) wcnt[q][t] = 0u;
  gofs[t] = t < (1 << BITS) ? hist[(size_t)t * nb + tile] : 0u;  // rows of the pass's digits
  typename KeyReg<K>::T k[SortTiling<K>::ITEMS];
  uint32_t v[SortTiling<K>::ITEMS], rk[SortTiling<K>::ITEMS], cnt;
  load_tile<K, VALS>(kin, vin, base, nvalid, k, v);
  __syncthreads();
  rank_tile<K, BITS>(k, nvalid, shift, wcnt, bstart, rk, cnt);
  place_and_write<K, VALS>(k, v, rk, nvalid, shift, (1u << BITS) - 1u, wcnt, bstart, gofs, sk, sv,
                           kout, vout);
}

// Stage names for the per-kernel timing, by sort: pairs (PLAIN), depth
// (FAST records), tiles (FAST lists).
struct SortNames {
  const char *hist, *scatter, *hist_all, *onesweep;
};
static const SortNames k_sort_names[5] = {
    {"radix_hist_kv", "radix_scatter_kv", "sort_hist_kv", "onesweep_kv"},
    {"depth_hist", "depth_scatter", "depth_hist_all", "depth_onesweep"},
    {"tile_hist", "tile_scatter", "tile_hist_all", "tile_onesweep"},
    {"redo_hist", "redo_scatter", "redo_hist_all", "redo_onesweep"},
    {"tile_hist", "tile_scatter", "tile_hist_all", "tile_onesweep"},
};

// Runs BODY with the constexpr int B_ = bits (1..8), the digit width the
// ranking kernels are compiled for.
#define FGS_WITH_BITS(bits, BODY)                   \
  switch (bits) {                                   \
    case 1: { constexpr int B_ = 1; BODY; } break;  \
    case 2: { constexpr int B_ = 2; BODY; } break;  \
    case 3: { constexpr int B_ = 3; BODY; } break;  \
    case 4: { constexpr int B_ = 4; BODY; } break;  \
    case 5: { constexpr int B_ = 5; BODY; } break;  \
    case 6: { constexpr int B_ = 6; BODY; } break;  \
    case 7: { constexpr int B_ = 7; BODY; } break;  \
    case 8: { constexpr int B_ = 8; BODY; } break;  \
    default: return hipErrorInvalidValue;           \
  }

template <typename K>
static hipError_t sort_kv_classic(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt,
                                  int64_t n, int begin_bit, int end_bit, void* ws, hipStream_t s,
                                  bool& in_alt, const uint32_t* n_dev, uint32_t* last_vals,
                                  const SortNames& nm) {
  const int nb = (int)((n + SortTiling<K>::TILE - 1) / SortTiling<K>::TILE);
  uint32_t* hist = (uint32_t*)ws;
  void* scan_ws = (char*)ws + align_up((size_t)256 * nb * sizeof(uint32_t));
  K *ka = keys, *kb = keys_alt;
  uint32_t *va = vals, *vb = vals_alt;
  const int npass = sort_passes(begin_bit, end_bit);
  for (int p = 0; p < npass; ++p) {
    int shift, bits;
    pass_digits(begin_bit, end_bit, p, shift, bits);
    const uint32_t mask = (1u << bits) - 1u;
    const bool last = p == npass - 1;
    uint32_t* vout = (last && last_vals) ? last_vals : vb;
    FGS_TIMED(nm.hist, s,
              (k_radix_hist<K><<<nb, SORT_THREADS, 0, s>>>(ka, n, n_dev, shift, mask, hist, nb)));
    hipError_t e = scan_u32(hist, hist, (int64_t)(mask + 1) * nb, nullptr, scan_ws, s);
    if (e != hipSuccess) return e;
    if (vals)
      FGS_WITH_BITS(bits, FGS_TIMED(nm.scatter, s,
                                    (k_radix_scatter<K, true, B_><<<xcd_grid(nb), SORT_THREADS, 0, s>>>(
                                        ka, va, kb, vout, n, n_dev, shift, hist, nb))))
    else
      FGS_WITH_BITS(bits, FGS_TIMED(nm.scatter, s,
                                    (k_radix_scatter<K, false, B_><<<xcd_grid(nb), SORT_THREADS, 0, s>>>(
                                        ka, nullptr, kb, nullptr, n, n_dev, shift, hist, nb))))
    K* tk = ka; ka = kb; kb = tk;
    uint32_t* tv = va; va = vb; vb = tv;
    in_alt = !in_alt;
  }
  return hipGetLastError();
}

// ------------------------------------------------------ radix, onesweep
// Onesweep-style passes for inputs up to ONESWEEP_MAX_KEYS (where the
// three-launch passes are launch- and latency-bound):
//   k_sort_prepare   zeroes the digit histograms and tile counters;
//   k_sort_hist      one read of the keys builds the histograms of all passes
//                    (LDS counts, one global atomic per digit and workgroup);
//   k_onesweep       one launch per pass: a workgroup takes the next 4096-key
//                    tile from a counter, ranks it (rank_tile), publishes its
//                    digit counts and looks back over the earlier tiles'
//                    published counts for its global offsets (decoupled
//                    look-back), then writes each digit's run contiguously.
// Tiles are taken in launch order, so every tile a workgroup waits on
// belongs to a workgroup that started earlier and is resident; the wait is
// additionally bounded (a fault flag is raised instead of hanging).
//
// Status words: bits 31..30 = 0 (not yet), 1 (tile count), 2 (inclusive
// prefix); bits 29..0 the count (n < 2^30).
constexpr uint32_t ST_AGG = 1u << 30, ST_PRE = 2u << 30, ST_CNT = (1u << 30) - 1u;
constexpr uint32_t SPIN_LIMIT = 1u << 22;
constexpr int64_t HIST_ALL_BLOCKS = 1024;
constexpr int LOOKBACK = 8;  // status words read per look-back step (in flight together)

__global__ __launch_bounds__(256) void k_sort_prepare(uint32_t* __restrict__ hist,
                                                      uint32_t* __restrict__ counters) {
  for (int i = threadIdx.x; i < SORT_MAX_PASSES * 256; i += 256) hist[i] = 0u;
  if (threadIdx.x <= SORT_MAX_PASSES) counters[threadIdx.x] = 0u;  // tile counters + fault flag
}

// Histograms of all passes in one read (per_block sort tiles per workgroup);
// also clears the first pass's status rows of the tiles it covers.
template <typename K>
__global__ __launch_bounds__(SORT_THREADS) void k_sort_hist(const K* __restrict__ keys,
                                                            int64_t n_cap,
                                                            const uint32_t* __restrict__ n_dev,
                                                            int begin_bit, int end_bit,
                                                            uint32_t* __restrict__ hist,
                                                            uint32_t* __restrict__ status0,
                                                            int64_t nb, int per_block) {
  __shared__ uint32_t cnt[SORT_MAX_PASSES][256];
  const int npass = sort_passes(begin_bit, end_bit);
  const int t = threadIdx.x;
  for (int p = 0; p < npass; ++p) cnt[p][t] = 0u;
  const int64_t tile0 = (int64_t)blockIdx.x * per_block;
  for (int64_t tl = tile0; tl < min(nb, tile0 + per_block); ++tl) status0[tl * 256 + t] = 0u;
  __syncthreads();
  const int64_t n = dev_count(n_dev, n_cap);
  const int64_t e0 = tile0 * SortTiling<K>::TILE, e1 = min(n, e0 + (int64_t)per_block * SortTiling<K>::TILE);
  constexpr int N = KeyVec<K>::N;
  constexpr int U = 4;  // vector loads in flight per thread
  constexpr int64_t STEP = (int64_t)SORT_THREADS * N;
  for (int64_t e = e0 + (int64_t)t * N; e < e1; e += U * STEP) {
    KeyVec<K> kv[U];
    int m[U];
#pragma unroll
    for (int u = 0; u < U; ++u) m[u] = load_keys(keys, e + u * STEP, e1, kv[u]);
    for (int p = 0; p < npass; ++p) {
      int sh, bits;
      pass_digits(begin_bit, end_bit, p, sh, bits);
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int j = 0; j < N; ++j)
          if (j < m[u]) atomicAdd(&cnt[p][(uint32_t)(kv[u].k[j] >> sh) & ((1u << bits) - 1u)], 1u);
    }
  }
  __syncthreads();
  for (int p = 0; p < npass; ++p)
    if (cnt[p][t]) atomicAdd(&hist[p * 256 + t], cnt[p][t]);
}

__device__ __forceinline__ uint32_t status_load(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename K, bool VALS, int BITS>
__global__ __launch_bounds__(SORT_THREADS) void k_onesweep(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin, K* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n_cap, const uint32_t* __restrict__ n_dev, int shift,
    const uint32_t* __restrict__ hist, uint32_t* __restrict__ counter,
    uint32_t* __restrict__ status, uint32_t* __restrict__ status_next,
    uint32_t* __restrict__ fault) {
  constexpr int NW = SORT_THREADS / WAVE;
  __shared__ K sk[SortTiling<K>::TILE];
  __shared__ uint32_t sv[VALS ? SortTiling<K>::TILE : 1];
  __shared__ uint32_t wcnt[NW][256];
  __shared__ uint32_t bstart[256];
  __shared__ uint32_t gofs[256];
  __shared__ uint32_t s_tile;
  const int t = threadIdx.x;
  constexpr uint32_t mask = (1u << BITS) - 1u;
  const int64_t n = dev_count(n_dev, n_cap);
  // workgroups past the tiles that hold keys leave before taking one (the
  // device-side count may be far below the capacity the grid is sized for,
  // and every take is an atomic on one word); the others take exactly the
  // tiles that hold keys
  if ((int64_t)blockIdx.x * SortTiling<K>::TILE >= n) return;  // whole workgroup
  if (t == 0) s_tile = atomicAdd(counter, 1u);
#pragma unroll
  for (int q = 0; q < NW; ++q) wcnt[q][t] = 0u;
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t base = tile * SortTiling<K>::TILE;
  if (base >= n) return;  // whole workgroup; no later tile waits on this one
  if (status_next) status_next[tile * 256 + t] = 0u;
  const int nvalid = (int)min((int64_t)SortTiling<K>::TILE, n - base);
  typename KeyReg<K>::T k[SortTiling<K>::ITEMS];
  uint32_t v[SortTiling<K>::ITEMS], rk[SortTiling<K>::ITEMS], c;
  load_tile<K, VALS>(kin, vin, base, nvalid, k, v);
  rank_tile<K, BITS>(k, nvalid, shift, wcnt, bstart, rk, c);
  // global offset of digit t: digits below it plus this digit in earlier tiles
  uint32_t gd;
  {
    const uint32_t hv = (uint32_t)t <= mask ? hist[t] : 0u;
    uint32_t htot;
    gd = block_excl_scan<SORT_THREADS>(hv, htot);
  }
  uint32_t* my = status + tile * 256 + t;
  uint32_t excl = 0;
  if (tile == 0) {
    status_store(my, ST_PRE | c);
  } else {
    status_store(my, ST_AGG | c);
    // Walk back over the earlier tiles' words of digit t, LOOKBACK at a time
    // (independent loads, one latency per step): add counts up to the first
    // inclusive prefix; a word not yet published ends the step and is read
    // again in the next one.
    int64_t p = tile - 1;
    uint32_t spins = 0;
    bool done = false;
    while (!done) {
      uint32_t st[LOOKBACK];
#pragma unroll
      for (int j = 0; j < LOOKBACK; ++j)
        st[j] = p - j >= 0 ? status_load(status + (p - j) * 256 + t) : 0u;
      int used = 0;
      bool stop = false;
#pragma unroll
      for (int j = 0; j < LOOKBACK; ++j) {
        const uint32_t flag = st[j] & ~ST_CNT;
        if (!stop && flag != 0u) {
          excl += st[j] & ST_CNT;
          if (flag == ST_PRE) done = true;
          ++used;
        }
        stop = stop || flag == 0u || flag == ST_PRE;
      }
      p -= used;
      if (!done && used == 0) {
        if (++spins > SPIN_LIMIT) {  // cannot happen with in-order tiles; never hang
          atomicOr(fault, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    status_store(my, ST_PRE | (excl + c));
  }
  gofs[t] = gd + excl;
  __syncthreads();
  place_and_write<K, VALS>(k, v, rk, nvalid, shift, mask, wcnt, bstart, gofs, sk, sv, kout, vout);
}

template <typename K>
static hipError_t sort_kv_onesweep(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt,
                                   int64_t n, int begin_bit, int end_bit, void* ws, hipStream_t s,
                                   bool& in_alt, const uint32_t* n_dev, uint32_t* last_vals,
                                   const SortNames& nm, uint32_t* sync, uint32_t* fault_word) {
  const int64_t nb = (n + SortTiling<K>::TILE - 1) / SortTiling<K>::TILE;
  SortWs w = carve_sort(ws, nb);
  const int npass = sort_passes(begin_bit, end_bit);
  if (sync) {  // cleared by the caller
    w.hist = sync;
    w.counters = sync + SORT_MAX_PASSES * 256;
  } else {
    k_sort_prepare<<<1, 256, 0, s>>>(w.hist, w.counters);
  }
  // one sort tile per workgroup up to HIST_ALL_BLOCKS workgroups (measured
  // at 1M keys: fewer, longer workgroups are slower although they issue
  // fewer same-address atomics)
  const int per_block = (int)((nb + HIST_ALL_BLOCKS - 1) / HIST_ALL_BLOCKS);
  const unsigned hb = (unsigned)((nb + per_block - 1) / per_block);
  FGS_TIMED(nm.hist_all, s,
            (k_sort_hist<K><<<hb, SORT_THREADS, 0, s>>>(keys, n, n_dev, begin_bit, end_bit,
                                                         w.hist, w.status[0], nb, per_block)));
  K *ka = keys, *kb = keys_alt;
  uint32_t *va = vals, *vb = vals_alt;
  for (int p = 0; p < npass; ++p) {
    int shift, bits;
    pass_digits(begin_bit, end_bit, p, shift, bits);
    const bool last = p == npass - 1;
    uint32_t* vout = (last && last_vals) ? last_vals : vb;
    uint32_t* st = w.status[p & 1];
    uint32_t* st_next = last ? nullptr : w.status[(p + 1) & 1];
    const uint32_t* h = w.hist + p * 256;
    uint32_t* ctr = w.counters + p;
    uint32_t* fault = fault_word ? fault_word : w.counters + SORT_MAX_PASSES;
    if (vals)
      FGS_WITH_BITS(bits, FGS_TIMED(nm.onesweep, s,
                                    (k_onesweep<K, true, B_><<<(unsigned)nb, SORT_THREADS, 0, s>>>(
                                        ka, va, kb, vout, n, n_dev, shift, h, ctr, st, st_next,
                                        fault))))
    else
      FGS_WITH_BITS(bits, FGS_TIMED(nm.onesweep, s,
                                    (k_onesweep<K, false, B_><<<(unsigned)nb, SORT_THREADS, 0, s>>>(
                                        ka, nullptr, kb, nullptr, n, n_dev, shift, h, ctr, st,
                                        st_next, fault))))
    K* tk = ka; ka = kb; kb = tk;
    uint32_t* tv = va; va = vb; vb = tv;
    in_alt = !in_alt;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------- dispatch
template <typename K>
hipError_t sort_kv(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, int64_t n,
                   int begin_bit, int end_bit, void* ws, hipStream_t s, bool& in_alt,
                   const uint32_t* n_dev, uint32_t* last_vals, SortKind kind,
                   uint32_t* sync, uint32_t* fault_word) {
  in_alt = false;
  // with a device-side count the passes always run (the count may be 0 or 1)
  if ((n <= 1 && !n_dev) || n <= 0 || end_bit <= begin_bit) return hipSuccess;
  if (n > (int64_t)ST_CNT || end_bit - begin_bit > 8 * SORT_MAX_PASSES ||
      end_bit > (int)(8 * sizeof(K)))
    return hipErrorInvalidValue;
  const SortNames& nm = k_sort_names[kind];
  // the redo list build usually has no keys at all: the fewest launches
  if (n <= ONESWEEP_MAX_KEYS || kind == SORT_TILES_REDO)
    return sort_kv_onesweep<K>(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, ws, s,
                               in_alt, n_dev, last_vals, nm, sync, fault_word);
  return sort_kv_classic<K>(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, ws, s, in_alt,
                            n_dev, last_vals, nm);
}

template <typename K>
hipError_t sort_kv_passes(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, int64_t n,
                          int begin_bit, int end_bit, void* ws, hipStream_t s, bool& in_alt) {
  in_alt = false;
  if (n <= 1 || end_bit <= begin_bit) return hipSuccess;
  if (n > (int64_t)ST_CNT || end_bit - begin_bit > 8 * SORT_MAX_PASSES ||
      end_bit > (int)(8 * sizeof(K)))
    return hipErrorInvalidValue;
  return sort_kv_classic<K>(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, ws, s, in_alt,
                            nullptr, nullptr, k_sort_names[SORT_PAIRS]);
}
template hipError_t sort_kv_passes<uint64_t>(uint64_t*, uint32_t*, uint64_t*, uint32_t*, int64_t,
                                             int, int, void*, hipStream_t, bool&);

template hipError_t sort_kv<uint16_t>(uint16_t*, uint32_t*, uint16_t*, uint32_t*, int64_t, int,
                                      int, void*, hipStream_t, bool&, const uint32_t*, uint32_t*,
                                      SortKind, uint32_t*, uint32_t*);
template hipError_t sort_kv<uint32_t>(uint32_t*, uint32_t*, uint32_t*, uint32_t*, int64_t, int,
                                      int, void*, hipStream_t, bool&, const uint32_t*, uint32_t*,
                                      SortKind, uint32_t*, uint32_t*);
template hipError_t sort_kv<uint64_t>(uint64_t*, uint32_t*, uint64_t*, uint32_t*, int64_t, int,
                                      int, void*, hipStream_t, bool&, const uint32_t*, uint32_t*,
                                      SortKind, uint32_t*, uint32_t*);

}  // namespace fgs
