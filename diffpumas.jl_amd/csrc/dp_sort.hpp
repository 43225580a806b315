// Stable LSD radix sort of (key, value) int32 pairs with keys in [0, 2^bits):
// 8-bit digits, one 64-lane wave per tile of 1024 entries.
//
//   sort_hist     per tile and digit value: count (LDS atomics; the counts
//                 do not depend on their order), written digit-major;
//   sort_scan_*   exclusive scan of the digit-major counts (chunk sums, their
//                 scan, chunk scans);
//   sort_scatter  each tile walks its entries in index order, 64 per round:
//                 the lanes holding the same digit find each other with 8
//                 ballots, take ranks from the lower lanes of that set, and
//                 the highest lane of the set advances the digit's base.
//
// Equal keys keep their input order, so the result is the unique stable
// order, whatever the grid.  Only kernels touch memory (no memset, no
// allocation): the sort can sit inside a captured graph.  Scratch:
// sort_hist_ints(n) ints of counts and two key/value pairs of n ints.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dps {
namespace {   // one copy of the kernels per translation unit that sorts

constexpr int SORT_ITEMS = 16;                 // rounds of 64 entries per tile
constexpr int SORT_TILE = 64 * SORT_ITEMS;     // 1024 entries per tile
constexpr int SORT_SCAN_THREADS = 1024;

inline int sort_tiles(int n) { return (n + SORT_TILE - 1) / SORT_TILE; }

__global__ __launch_bounds__(64) void sort_hist(const int32_t* __restrict__ key, int n, int shift,
                                                int n_tiles, int32_t* __restrict__ hist) {
  __shared__ int cnt[256];
  for (int d = threadIdx.x; d < 256; d += 64) cnt[d] = 0;
  __syncthreads();
  const int base = blockIdx.x * SORT_TILE;
  for (int k = 0; k < SORT_ITEMS; ++k) {
    const int i = base + k * 64 + (int)threadIdx.x;
    if (i < n) atomicAdd(&cnt[((unsigned)key[i] >> shift) & 255u], 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < 256; d += 64) hist[(size_t)d * n_tiles + blockIdx.x] = cnt[d];
}

// In-place exclusive scan of m ints in three launches: per-chunk sums,
// a scan of the chunk sums (one block), then each chunk scanned from its
// offset.  A chunk is SORT_SCAN_THREADS x 16 consecutive ints.
constexpr int SORT_SCAN_PER = 16;
constexpr int SORT_SCAN_CHUNK = SORT_SCAN_THREADS * SORT_SCAN_PER;

inline int scan_chunks(int m) { return (m + SORT_SCAN_CHUNK - 1) / SORT_SCAN_CHUNK; }
// Ints of the count buffer: 256 digit counts per tile, then the chunk sums.
inline size_t sort_hist_ints(int n) {
  const int m = 256 * sort_tiles(n);
  return (size_t)m + scan_chunks(m);
}

// Exclusive scan of one int per thread of a SORT_SCAN_THREADS block; returns
// the block total through *total.
__device__ __forceinline__ int block_exclusive(int v, int* sh, int* total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < SORT_SCAN_THREADS; off <<= 1) {
    const int add = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  const int incl = sh[t];
  *total = sh[SORT_SCAN_THREADS - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(SORT_SCAN_THREADS) void sort_scan_sums(const int32_t* __restrict__ a,
                                                                    int m,
                                                                    int32_t* __restrict__ sums) {
  __shared__ int sh[SORT_SCAN_THREADS];
  const int lo = blockIdx.x * SORT_SCAN_CHUNK + threadIdx.x * SORT_SCAN_PER;
  int v = 0;
  for (int i = 0; i < SORT_SCAN_PER; ++i)
    if (lo + i < m) v += a[lo + i];
  int total;
  (void)block_exclusive(v, sh, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SORT_SCAN_THREADS) void sort_scan_top(int32_t* __restrict__ sums,
                                                                   int nc) {
  __shared__ int sh[SORT_SCAN_THREADS];
  int carry = 0;
  for (int c0 = 0; c0 < nc; c0 += SORT_SCAN_THREADS) {
    const int c = c0 + (int)threadIdx.x;
    const int v = c < nc ? sums[c] : 0;
    int total;
    const int ex = block_exclusive(v, sh, &total);
    if (c < nc) sums[c] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(SORT_SCAN_THREADS) void sort_scan_apply(int32_t* __restrict__ a,
                                                                     int m,
                                                                     const int32_t* __restrict__ sums) {
  __shared__ int sh[SORT_SCAN_THREADS];
  const int lo = blockIdx.x * SORT_SCAN_CHUNK + threadIdx.x * SORT_SCAN_PER;
  int v = 0;
  for (int i = 0; i < SORT_SCAN_PER; ++i)
    if (lo + i < m) v += a[lo + i];
  int total;
  int run = sums[blockIdx.x] + block_exclusive(v, sh, &total);
  for (int i = 0; i < SORT_SCAN_PER; ++i)
    if (lo + i < m) {
      const int x = a[lo + i];
      a[lo + i] = run;
      run += x;
    }
}

__global__ __launch_bounds__(64) void sort_scatter(const int32_t* __restrict__ kin,
                                                   const int32_t* __restrict__ vin, int n,
                                                   int shift, int n_tiles,
                                                   const int32_t* __restrict__ offs,
                                                   int32_t* __restrict__ kout,
                                                   int32_t* __restrict__ vout) {
  __shared__ int base[256];
  for (int d = threadIdx.x; d < 256; d += 64) base[d] = offs[(size_t)d * n_tiles + blockIdx.x];
  __syncthreads();
  const int lane = threadIdx.x;
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int t0 = blockIdx.x * SORT_TILE;
  for (int k = 0; k < SORT_ITEMS; ++k) {
    const int i = t0 + k * 64 + lane;
    const bool live = i < n;
    const int32_t kv = live ? kin[i] : 0;
    const int32_t vv = live ? vin[i] : 0;
    const unsigned d = ((unsigned)kv >> shift) & 255u;
    unsigned long long same = __ballot(live);
    for (int b = 0; b < 8; ++b) {
      const unsigned long long ones = __ballot(live && ((d >> b) & 1u));
      same &= ((d >> b) & 1u) ? ones : ~ones;
    }
    const int pos = live ? base[d] + __popcll(same & below) : 0;
    // every lane has read its digit's base before the highest lane of each set moves it
    __builtin_amdgcn_wave_barrier();
    if (live) {
      kout[pos] = kv;
      vout[pos] = vv;
      if ((same >> lane) == 1ull) base[d] += __popcll(same);   // highest lane of the set
    }
    __syncthreads();
  }
}

// Sorts (k_in, v_in) into (k_out, v_out) stably by the low `bits` key bits.
// ping/pong: 2 * n ints each (keys then values); hist: sort_hist_ints(n)
// ints.  k_in/v_in are left unchanged.  Returns the buffers that hold the
// result through k_out/v_out.
inline void radix_sort_pairs(const int32_t* k_in, const int32_t* v_in, int n, int bits,
                             int32_t* ping, int32_t* pong, int32_t* hist, hipStream_t st,
                             const int32_t** k_out, const int32_t** v_out) {
  const int tiles = sort_tiles(n);
  const int32_t* kc = k_in;
  const int32_t* vc = v_in;
  int32_t* buf[2] = {ping, pong};
  int w = 0;
  for (int shift = 0; shift < bits; shift += 8) {
    int32_t* ko = buf[w];
    int32_t* vo = buf[w] + n;
    hipLaunchKernelGGL(sort_hist, dim3(tiles), dim3(64), 0, st, kc, n, shift, tiles, hist);
    const int m = 256 * tiles, nc = scan_chunks(m);
    int32_t* sums = hist + m;
    hipLaunchKernelGGL(sort_scan_sums, dim3(nc), dim3(SORT_SCAN_THREADS), 0, st, hist, m, sums);
    hipLaunchKernelGGL(sort_scan_top, dim3(1), dim3(SORT_SCAN_THREADS), 0, st, sums, nc);
    hipLaunchKernelGGL(sort_scan_apply, dim3(nc), dim3(SORT_SCAN_THREADS), 0, st, hist, m, sums);
    hipLaunchKernelGGL(sort_scatter, dim3(tiles), dim3(64), 0, st, kc, vc, n, shift, tiles, hist,
                       ko, vo);
    kc = ko;
    vc = vo;
    w ^= 1;
  }
  *k_out = kc;
  *v_out = vc;
}

}  // namespace
}  // namespace dps
