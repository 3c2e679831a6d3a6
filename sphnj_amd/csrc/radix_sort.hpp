// Stable LSD radix sort of (32-bit key, 32-bit value) pairs for gfx950, written for wave64.
//
// Replaces thrust::sort_by_key + reduce_by_key + exclusive_scan of the reference
// (src/sort.cu:10-45).  One pass per 8-bit digit, three kernels per pass:
//   upsweep    per-tile digit histogram: each wave finds the lanes sharing its digit with 8
//              ballots (peer groups), and one lane per group adds the group size to an LDS
//              counter, so equal digits (common: particles share cells) cost one LDS atomic
//              per group instead of one per lane;
//   rowscan    exclusive scan of each digit's row of tile counts (digit-major layout);
//   downsweep  re-reads the tile: wave w takes the tile's w-th quarter, 64 consecutive items
//              a round; an item's place in its wave's run of its digit is the run's count so
//              far (kept by the peer group's first lane in LDS) plus its rank among its peers
//              (popcount of the lower lanes); one block scan then turns the per-wave digit
//              counts into offsets, and the items, held in registers, are scattered.
// Items are ranked in index order (quarter by quarter, round by round, lane by lane), so the
// sort is stable.  Values may be implicit (value = input index) on the first pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sphnj {

constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;

__device__ __forceinline__ unsigned long long lanemask_lt() {
  const unsigned lane = threadIdx.x & 63u;
  return lane == 0 ? 0ull : (~0ull >> (64u - lane));
}

// Lanes of the wave whose digit equals this lane's digit (0 for invalid lanes).  Every lane
// of the wave must call it (the ballots are wave-wide).
__device__ __forceinline__ unsigned long long match_digit(unsigned d, bool valid) {
  unsigned long long m = __ballot(valid ? 1 : 0);
#pragma unroll
  for (int b = 0; b < kRadixBits; ++b) {
    const bool bit = (d >> b) & 1u;
    const unsigned long long bal = __ballot(bit ? 1 : 0);
    m &= bit ? bal : ~bal;
  }
  return valid ? m : 0ull;
}

// Exclusive scan over the 256 threads of a block; returns the exclusive prefix of v and
// writes the block total to *total.  tmp holds 4 ints (one per wave).
__device__ __forceinline__ int block_exscan256(int v, int *tmp, int *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) tmp[w] = inc;
  __syncthreads();
  int wbase = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int t = tmp[k];
    if (k < w) wbase += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return wbase + inc - v;
}

__global__ void __launch_bounds__(kSortThreads) radix_upsweep(const int *__restrict__ keys, int n, int shift,
                                                              int nblocks, int *__restrict__ hist) {
  __shared__ int cnt[kRadix];
  const int t = threadIdx.x;
  cnt[t] = 0;
  __syncthreads();
  const int base = blockIdx.x * kSortTile;
  const unsigned long long lt = lanemask_lt();
  for (int r = 0; r < kSortItems; ++r) {
    const int idx = base + r * kSortThreads + t;
    const bool valid = idx < n;
    const unsigned d = valid ? (((unsigned)keys[idx]) >> shift) & (kRadix - 1) : 0u;
    const unsigned long long peers = match_digit(d, valid);
    if (valid && (peers & lt) == 0ull) atomicAdd(&cnt[d], (int)__popcll(peers));
  }
  __syncthreads();
  hist[t * nblocks + blockIdx.x] = cnt[t];
}

__global__ void __launch_bounds__(kSortThreads) radix_rowscan(int *__restrict__ hist, int nblocks,
                                                              int *__restrict__ digit_total) {
  __shared__ int tmp[4];
  int *row = hist + blockIdx.x * nblocks;
  int carry = 0;
  for (int s = 0; s < nblocks; s += kSortThreads) {
    const int i = s + threadIdx.x;
    const int v = i < nblocks ? row[i] : 0;
    int tot;
    const int ex = block_exscan256(v, tmp, &tot);
    if (i < nblocks) row[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) digit_total[blockIdx.x] = carry;
}

__global__ void __launch_bounds__(kSortThreads) radix_downsweep(const int *__restrict__ kin,
                                                                const int *__restrict__ vin,
                                                                int *__restrict__ kout, int *__restrict__ vout,
                                                                int n, int shift, int nblocks,
                                                                const int *__restrict__ hist,
                                                                const int *__restrict__ digit_total) {
  constexpr int kWaves = kSortThreads / 64;
  __shared__ int base[kRadix];
  __shared__ int woff[kWaves][kRadix];  // per-wave digit counts, then the waves' offsets
  __shared__ int tmp[4];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
#pragma unroll
  for (int k = 0; k < kWaves; ++k) woff[k][t] = 0;
  int tot;
  const int ex = block_exscan256(digit_total[t], tmp, &tot);  // its barriers order the clears
  base[t] = ex + hist[t * nblocks + blockIdx.x];
  const int first = blockIdx.x * kSortTile + w * (kSortItems * 64);
  const unsigned long long lt = lanemask_lt();
  int key[kSortItems], val[kSortItems], loc[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int idx = first + r * 64 + lane;
    const bool valid = idx < n;
    key[r] = valid ? kin[idx] : 0;
    val[r] = valid ? (vin ? vin[idx] : idx) : 0;
    const unsigned d = (((unsigned)key[r]) >> shift) & (kRadix - 1);
    const unsigned long long peers = match_digit(d, valid);
    const int rank = (int)__popcll(peers & lt);
    int run = 0;
    if (valid && rank == 0) {  // the group's first lane owns the wave's count of digit d
      run = woff[w][d];
      woff[w][d] = run + (int)__popcll(peers);
    }
    run = __shfl(run, valid ? (int)__builtin_ctzll(peers) : lane, 64);
    loc[r] = run + rank;
  }
  __syncthreads();
  {
    int b = base[t];
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
      const int c = woff[k][t];
      woff[k][t] = b;
      b += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    if (first + r * 64 + lane < n) {
      const unsigned d = (((unsigned)key[r]) >> shift) & (kRadix - 1);
      const int pos = woff[w][d] + loc[r];
      kout[pos] = key[r];
      vout[pos] = val[r];
    }
  }
}

// The same stable passes for n <= kSmallSortN items in one 1024-thread block: wave w holds
// the w-th sixteenth of the items (16 rounds of 64, in index order) in registers, ranks them
// as the downsweep does, and the block exchanges them through LDS between passes, so a pass
// costs barriers instead of three launches (a few thousand solids: one launch instead of nine).
// The digits are taken from key - min(key), so a set of keys spanning few values (the cells of
// one solid body) needs fewer passes than the largest key would.
constexpr int kSmallSortThreads = 1024;
constexpr int kSmallSortItems = 16;
constexpr int kSmallSortN = kSmallSortThreads * kSmallSortItems;

__global__ void __launch_bounds__(kSmallSortThreads) radix_sort_small(const int *__restrict__ kin,
                                                                      int *__restrict__ kout,
                                                                      int *__restrict__ vout, int n, int passes,
                                                                      const int *only_if) {
  if (only_if != nullptr && *only_if == 0) return;  // the outputs of the last sort still hold
  constexpr int kWaves = kSmallSortThreads / 64;
  __shared__ int s_key[kSmallSortN];
  __shared__ int s_val[kSmallSortN];
  __shared__ int woff[kWaves][kRadix];  // per-wave digit counts, then the waves' offsets
  __shared__ int tmp[kRadix / 64];
  __shared__ int s_lo[kWaves], s_hi[kWaves];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int first = w * (kSmallSortItems * 64);
  const unsigned long long lt = lanemask_lt();
  int key[kSmallSortItems], val[kSmallSortItems], loc[kSmallSortItems];
  int lo = 0x7fffffff, hi = 0;
#pragma unroll
  for (int r = 0; r < kSmallSortItems; ++r) {
    const int idx = first + r * 64 + lane;
    key[r] = idx < n ? kin[idx] : 0;
    val[r] = idx;
    if (idx < n) {
      lo = min(lo, key[r]);
      hi = max(hi, key[r]);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o, 64));
    hi = max(hi, __shfl_xor(hi, o, 64));
  }
  if (lane == 0) {
    s_lo[w] = lo;
    s_hi[w] = hi;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kWaves; ++k) {
    lo = min(lo, s_lo[k]);
    hi = max(hi, s_hi[k]);
  }
  const unsigned span = (unsigned)(hi - lo);  // keys are non-negative, so hi - lo fits
  int need = 0;
  while (need < passes && (span >> (need * kRadixBits)) != 0u) ++need;
  for (int p = 0; p < need; ++p) {
    const int shift = p * kRadixBits;
    for (int k = t; k < kWaves * kRadix; k += kSmallSortThreads) woff[k / kRadix][k % kRadix] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSmallSortItems; ++r) {
      const bool valid = first + r * 64 + lane < n;
      const unsigned d = ((unsigned)(key[r] - lo) >> shift) & (kRadix - 1);
      const unsigned long long peers = match_digit(d, valid);
      const int rank = (int)__popcll(peers & lt);
      int run = 0;
      if (valid && rank == 0) {  // the group's first lane owns the wave's count of digit d
        run = woff[w][d];
        woff[w][d] = run + (int)__popcll(peers);
      }
      run = __shfl(run, valid ? (int)__builtin_ctzll(peers) : lane, 64);
      loc[r] = run + rank;
    }
    __syncthreads();
    // digit totals, their exclusive scan over the digits (threads 0..255), then each wave's
    // offset within its digit
    int dtot = 0, inc = 0;
    if (t < kRadix) {
#pragma unroll
      for (int k = 0; k < kWaves; ++k) dtot += woff[k][t];
      inc = dtot;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(inc, o, 64);
        if (lane >= o) inc += u;
      }
      if (lane == 63) tmp[w] = inc;
    }
    __syncthreads();
    if (t < kRadix) {
      int b = inc - dtot;
      for (int k = 0; k < w; ++k) b += tmp[k];
#pragma unroll
      for (int k = 0; k < kWaves; ++k) {
        const int c = woff[k][t];
        woff[k][t] = b;
        b += c;
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSmallSortItems; ++r) {
      if (first + r * 64 + lane < n) {
        const int pos = woff[w][((unsigned)(key[r] - lo) >> shift) & (kRadix - 1)] + loc[r];
        s_key[pos] = key[r];
        s_val[pos] = val[r];
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSmallSortItems; ++r) {
      const int idx = first + r * 64 + lane;
      if (idx < n) {
        key[r] = s_key[idx];
        val[r] = s_val[idx];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kSmallSortItems; ++r) {
    const int idx = first + r * 64 + lane;
    if (idx < n) {
      kout[idx] = key[r];
      vout[idx] = val[r];
    }
  }
}

// Work buffers for one sorter instance (sized for n items).
struct RadixSorter {
  int n = 0, nblocks = 0, passes = 0;
  int *k1 = nullptr, *v0 = nullptr, *v1 = nullptr, *hist = nullptr, *dtot = nullptr;
};

inline int radix_passes_for(int max_key) {
  int bits = 0;
  while (bits < 31 && (max_key >> bits) != 0) ++bits;
  int p = (bits + kRadixBits - 1) / kRadixBits;
  return p < 1 ? 1 : p;
}

// Sort (keys[i], i) for i < n.  keys is overwritten on even pass counts of the multi-block
// path.  On return *out_keys / *out_vals point at the sorted keys and values (inside keys / the
// sorter's buffers).  With only_if (n <= kSmallSortN only), the sort runs when *only_if is
// non-zero at that point of the stream, and otherwise leaves the last sort's outputs in place.
inline hipError_t radix_sort_pairs(const RadixSorter &S, int *keys, int **out_keys, int **out_vals,
                                   hipStream_t st, const int *only_if = nullptr) {
  if (S.n == 0) {
    *out_keys = keys;
    *out_vals = S.v0;
    return hipSuccess;
  }
  if (S.n <= kSmallSortN) {
    radix_sort_small<<<1, kSmallSortThreads, 0, st>>>(keys, S.k1, S.v0, S.n, S.passes, only_if);
    *out_keys = S.k1;
    *out_vals = S.v0;
    return hipGetLastError();
  }
  int *ka = keys, *kb = S.k1;
  int *va = nullptr, *vb = S.v1;
  for (int p = 0; p < S.passes; ++p) {
    const int shift = p * kRadixBits;
    radix_upsweep<<<S.nblocks, kSortThreads, 0, st>>>(ka, S.n, shift, S.nblocks, S.hist);
    radix_rowscan<<<kRadix, kSortThreads, 0, st>>>(S.hist, S.nblocks, S.dtot);
    radix_downsweep<<<S.nblocks, kSortThreads, 0, st>>>(ka, va, kb, vb, S.n, shift, S.nblocks, S.hist, S.dtot);
    int *tk = ka;
    ka = kb;
    kb = tk;
    int *tv = va;
    va = vb;
    vb = (p == 0) ? S.v0 : tv;
  }
  *out_keys = ka;
  *out_vals = va;
  return hipGetLastError();
}

}  // namespace sphnj
