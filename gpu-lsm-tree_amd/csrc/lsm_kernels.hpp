// Device side of the LSM dictionary, for (uint32 key, uint32 value) rows and
// (uint64 key, uint64 value) rows.  Every kernel is a template over the row
// traits KV<K> below.
//
// Data model in HBM.  The dictionary is a stack of immutable sorted runs,
// newest first.  A run of c rows is two allocations:
//   rows[c]  rows of (key word, value): the key word holds the low B-1 bits
//            of the key with the tombstone flag in its top bit, the value
//            is 0 for a tombstone (B = key bits)
//   dir      RunMeta header (count, split, directory bits, tombstones,
//            bucket window) followed by a radix directory of 2^b + 1 u32
//            row offsets: dir[j] = first row whose key >= j << (B - b),
//            valid for j in [lo_bucket, hi_bucket + 1], the buckets of the
//            run's first and last keys; a key below that window has
//            lower bound 0, a key above it has lower bound count, so the
//            empty stretches of key space below and above the run are
//            never written.
// The top bit of a key is implied by the row position (rows >= split have
// it set), which frees it for the tombstone flag, so a row is exactly
// 2 * sizeof(K) bytes like the reference's 8-byte Row (GPULSMOpt.cuh:
// 343-377), and key and value share one cache line, so a point hit costs
// one line of rows.  Every search enters through the directory (b >= 1);
// inside one directory bucket all keys share the top bit, so comparisons
// use the low B-1 bits only.
#pragma once

#include "device_utils.hpp"

namespace gdict {
namespace lsm {

constexpr int kMaxRuns = 48;
constexpr u32 kMaxDirBits = 24;

struct RunMeta {
  u32 count;
  u32 split;      // rows whose key has the top bit clear
  u32 dir_bits;   // b, 1 <= b <= kMaxDirBits
  u32 tombs;      // nonzero when the run holds tombstone rows
  u32 lo_bucket;  // directory window: buckets of the first and last key
  u32 hi_bucket;  // (lo_bucket = 1, hi_bucket = 0: no rows)
};
constexpr u64 kMetaWords = sizeof(RunMeta) / sizeof(u32);
constexpr u64 kMetaBytes = 256;  // header slot in front of the directory

struct alignas(16) Row64 {
  u64 k;  // key word
  u64 v;  // value
};

// Row traits.  kSmallTile: keys a fused one-workgroup admission sorts in
// LDS; kMergeItemsMax: merge items per thread (merge tile = 256 x items),
// both sized so the LDS staging stays below 64 KB.
template <class K>
struct KV;

template <>
struct KV<u32> {
  using K = u32;
  using V = u32;
  using Row = uint2;
  static constexpr int kBits = 32;
  static constexpr K kTomb = 0x80000000u;
  static constexpr K kLow = 0x7fffffffu;
  static constexpr K kAllOnes = 0xffffffffu;
  static constexpr int kSmallTile = 4096;
  static constexpr int kMergeItemsMax = 16;
  __device__ __forceinline__ static K kw(const Row &r) { return r.x; }
  __device__ __forceinline__ static V val(const Row &r) { return r.y; }
  __device__ __forceinline__ static Row make(K kw, V v) {
    return make_uint2(kw, v);
  }
};

template <>
struct KV<u64> {
  using K = u64;
  using V = u64;
  using Row = Row64;
  static constexpr int kBits = 64;
  static constexpr K kTomb = 0x8000000000000000ull;
  static constexpr K kLow = 0x7fffffffffffffffull;
  static constexpr K kAllOnes = ~0ull;
  static constexpr int kSmallTile = 2048;
  static constexpr int kMergeItemsMax = 8;
  __device__ __forceinline__ static K kw(const Row &r) { return r.k; }
  __device__ __forceinline__ static V val(const Row &r) { return r.v; }
  __device__ __forceinline__ static Row make(K kw, V v) {
    Row r;
    r.k = kw;
    r.v = v;
    return r;
  }
};

template <class T>
struct RunView {
  const typename T::Row *rows;
  const u32 *dir;
  const RunMeta *meta;
};

template <class T>
struct RunSet {
  u32 count;
  u32 pad;
  RunView<T> run[kMaxRuns];
};

template <class T>
__device__ __forceinline__ typename T::K decode_key(typename T::K w, u32 pos,
                                                    u32 split) {
  return (w & T::kLow) | (pos >= split ? T::kTomb : typename T::K(0));
}

// Directory bits for `count` rows: about 2^rows_log rows per bucket.
__host__ __device__ __forceinline__ u32 dir_bits_for(u64 count, u32 max_bits,
                                                     u32 rows_log = 3) {
  u32 lg = 0;
  while (lg < 63 && (u64(1) << (lg + 1)) <= count) ++lg;
  u32 b = lg > rows_log ? lg - rows_log : 1;
  if (b < 1) b = 1;
  if (b > max_bits) b = max_bits;
  return b;
}

// Directory bounds [lo, hi) of the rows that may hold key x; false when x
// lies outside the run's bucket window (then lo = hi = its lower bound).
template <class T>
__device__ __forceinline__ bool dir_window(const RunView<T> &r,
                                           const RunMeta &m, typename T::K x,
                                           u32 &lo, u32 &hi) {
  const u32 bucket = static_cast<u32>(x >> (T::kBits - m.dir_bits));
  if (bucket < m.lo_bucket) {
    lo = hi = 0;
    return false;
  }
  if (bucket > m.hi_bucket) {
    lo = hi = m.count;
    return false;
  }
  lo = r.dir[bucket];
  hi = r.dir[bucket + 1];
  return true;
}

// First row in [lo, hi) whose key's low bits are >= target (the rows of
// one directory bucket): bisection down to kFindWindow rows, then one batch
// of loads.
constexpr u32 kFindWindow = 8;

template <class T>
__device__ __forceinline__ u32 window_lower_bound(const typename T::Row *rows,
                                                  u32 lo, u32 hi,
                                                  typename T::K target) {
  while (hi - lo > kFindWindow) {
    const u32 mid = (lo + hi) >> 1;
    if ((T::kw(rows[mid]) & T::kLow) < target) lo = mid + 1;
    else hi = mid;
  }
  typename T::Row w[kFindWindow];
#pragma unroll
  for (u32 i = 0; i < kFindWindow; ++i)
    if (lo + i < hi) w[i] = rows[lo + i];
  u32 below = 0;
#pragma unroll
  for (u32 i = 0; i < kFindWindow; ++i)
    below += (lo + i < hi && (T::kw(w[i]) & T::kLow) < target) ? 1u : 0u;
  return lo + below;
}

// First row whose key is >= x.
template <class T>
__device__ __forceinline__ u32 run_lower_bound(const RunView<T> &r,
                                               const RunMeta &m,
                                               typename T::K x) {
  u32 lo, hi;
  if (!dir_window(r, m, x, lo, hi)) return lo;
  return window_lower_bound<T>(r.rows, lo, hi, x & T::kLow);
}

// First row whose key is > x.
template <class T>
__device__ __forceinline__ u32 run_upper_bound(const RunView<T> &r,
                                               const RunMeta &m,
                                               typename T::K x) {
  return x == T::kAllOnes ? m.count : run_lower_bound(r, m, x + 1);
}

// Point search; returns the row position or 0xffffffff (row in `row`).
// Keys are unique inside a run, so a binary search that meets the key
// stops there; once the directory window is down to kFindWindow rows, its
// loads are issued together (one memory round trip instead of
// log2(window) dependent ones).

template <class T>
__device__ __forceinline__ u32 run_find(const RunView<T> &r, const RunMeta &m,
                                        typename T::K key,
                                        typename T::Row &row) {
  u32 lo, hi;
  if (!dir_window(r, m, key, lo, hi)) return 0xffffffffu;
  const typename T::K target = key & T::kLow;
  while (hi - lo > kFindWindow) {
    const u32 mid = (lo + hi) >> 1;
    const typename T::Row w = r.rows[mid];
    const typename T::K km = T::kw(w) & T::kLow;
    if (km == target) {
      row = w;
      return mid;
    }
    if (km < target) lo = mid + 1;
    else hi = mid;
  }
  typename T::Row w[kFindWindow];
#pragma unroll
  for (u32 i = 0; i < kFindWindow; ++i)
    if (lo + i < hi) w[i] = r.rows[lo + i];
  u32 at = 0xffffffffu;
#pragma unroll
  for (u32 i = 0; i < kFindWindow; ++i) {
    if (lo + i < hi && (T::kw(w[i]) & T::kLow) == target) {
      at = lo + i;
      row = w[i];
    }
  }
  return at;
}

// ------------------------------------------------------------ batch runs
constexpr int kFinThreads = 512;
constexpr int kFinItems = 8;
constexpr int kFinTile = kFinThreads * kFinItems;  // 4096
constexpr int kFinWaves = kFinThreads / kWave;

enum FinalizeMode : int {
  kModeInsert = 0,
  kModeErase = 1,
  kModeMixed = 2,
};

// Turns a stably sorted batch (keys sk, input positions sp) into a run:
// keep the last occurrence of each key (the newest), gather its value and
// tombstone flag, encode, and compact with a single-pass chained scan.
// Items are striped (item j of a thread is row base + j*512 + tid), so every
// load and store of a row j is one coalesced access of the block; ballots
// give the survivors' ranks.  The run header is written without atomics:
// the tile holding the last key writes the count, the tile where keys
// first reach the top bit (or the last tile) writes the split, and tiles
// with tombstones set the tombstone flag.  The radix directory (dir_bits
// bits) is written here too: sorted position i owns the directory entries
// of the buckets in (bucket(key[i-1]), bucket(key[i])] and writes them the
// number of survivors before position i (the output row of its key).
template <class T>
__global__ __launch_bounds__(kFinThreads) void finalize_batch_kernel(
    const typename T::K *__restrict__ sk, const u32 *__restrict__ sp, u32 n,
    const typename T::V *__restrict__ values,
    const u8 *__restrict__ erase_flags, int mode,
    typename T::Row *__restrict__ out, RunMeta *meta, u32 *__restrict__ dir,
    u32 dir_bits, u64 *status, u32 *ticket) {
  using K = typename T::K;
  using V = typename T::V;
  __shared__ u32 s_tile;
  __shared__ u32 s_off[kFinItems][kFinWaves];
  __shared__ u32 s_total;
  __shared__ u64 s_base;
  __shared__ u32 s_low, s_tomb;
  const u32 lane = lane_id();
  const u32 wave = threadIdx.x / kWave;
  if (threadIdx.x == 0) {
    s_tile = atomicAdd(ticket, 1u);
    s_low = 0;
    s_tomb = 0;
  }
  __syncthreads();
  const u32 tile = s_tile;
  const u64 base = u64(tile) * kFinTile;
  K key[kFinItems];
  u64 mask[kFinItems];
#pragma unroll
  for (int j = 0; j < kFinItems; ++j) {
    const u64 i = base + u64(j) * kFinThreads + threadIdx.x;
    bool keep = false;
    key[j] = 0;
    if (i < n) {
      key[j] = sk[i];
      keep = (i + 1 == n) || sk[i + 1] != key[j];
    }
    mask[j] = ballot(keep);
    if (lane == 0) s_off[j][wave] = popc64(mask[j]);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 run = 0;
    for (int j = 0; j < kFinItems; ++j)
      for (int w = 0; w < kFinWaves; ++w) {
        const u32 c = s_off[j][w];
        s_off[j][w] = run;
        run += c;
      }
    s_total = run;
  }
  __syncthreads();
  const u32 tile_total = s_total;
  if (threadIdx.x < kWave) {
    const u64 ex = wave_lookback(status, tile, tile_total);
    if (threadIdx.x == 0) s_base = ex;
  }
  __syncthreads();
  const u64 out_base = s_base;
  u32 low = 0, tombs = 0;
#pragma unroll
  for (int j = 0; j < kFinItems; ++j) {
    if (!((mask[j] >> lane) & 1ull)) continue;
    const u64 i = base + u64(j) * kFinThreads + threadIdx.x;
    const u64 o = out_base + s_off[j][wave] + popc64(mask[j] & lanemask_lt());
    const u32 p = sp[i];
    bool tomb;
    if (mode == kModeErase) tomb = true;
    else if (mode == kModeMixed) tomb = erase_flags[p] != 0;
    else tomb = false;
    const V v = tomb ? V(0) : values[p];
    out[o] = T::make((key[j] & T::kLow) | (tomb ? T::kTomb : K(0)), v);
    low += key[j] < T::kTomb;
    tombs += tomb;
  }
  // directory entries inside the bucket window
  const u32 shift = T::kBits - dir_bits;
#pragma unroll
  for (int j = 0; j < kFinItems; ++j) {
    const u64 i = base + u64(j) * kFinThreads + threadIdx.x;
    if (i >= n) continue;
    const u32 bk = static_cast<u32>(key[j] >> shift);
    const u32 before = static_cast<u32>(out_base + s_off[j][wave] +
                                        popc64(mask[j] & lanemask_lt()));
    if (i == 0) {
      meta->lo_bucket = bk;
      dir[bk] = 0;
    } else {
      for (u32 b = static_cast<u32>(sk[i - 1] >> shift) + 1; b <= bk; ++b)
        dir[b] = before;
    }
    if (i + 1 == n) {
      meta->hi_bucket = bk;
      meta->dir_bits = dir_bits;
      dir[bk + 1] = static_cast<u32>(out_base + tile_total);
    }
  }
  low = wave_sum_u32(low);
  tombs = wave_sum_u32(tombs);
  if (lane == 0) {
    if (low) atomicAdd(&s_low, low);
    if (tombs) atomicAdd(&s_tomb, tombs);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const u64 in_end = base + kFinTile < n ? base + kFinTile : n;
    const bool last_tile = in_end == n;
    // keys before this tile are all below the top bit and some key of the
    // tile (its last) reaches it, or no key does at all
    const bool transition = (base == 0 || sk[base - 1] < T::kTomb) &&
                            (sk[in_end - 1] >= T::kTomb || last_tile);
    if (last_tile) meta->count = static_cast<u32>(out_base + tile_total);
    if (transition) meta->split = static_cast<u32>(out_base + s_low);
    if (s_tomb) meta->tombs = 1;
  }
}

// ------------------------------------------------ small batches, one launch
// A batch of at most T::kSmallTile keys becomes a run in ONE workgroup: an
// LDS LSD radix sort (the same wave64 ballot multisplit as the onesweep
// passes, with the whole batch as one tile), newest-wins dedupe, value
// gather, encode, and the radix directory.  Replaces nine launches for the
// small arrivals of a serving stream.
constexpr int kSmallThreads = 512;
constexpr int kSmallWaves = kSmallThreads / kWave;

// Where a fused admission reads its batch: entry i's key, value and erase
// flag.  ArrayBatch reads three arrays; the serving engine's step kernels
// read the step's request buffers through its piece table instead.
template <class T>
struct ArrayBatch {
  const typename T::K *keys;
  const typename T::V *values;
  const u8 *erase_flags;
  __device__ __forceinline__ typename T::K key(u32 i) const { return keys[i]; }
  __device__ __forceinline__ typename T::V value(u32 i) const {
    return values[i];
  }
  __device__ __forceinline__ bool erased(u32 i) const {
    return erase_flags[i] != 0;
  }
};

// (the body of small_batch_run_kernel; also run by the serving engine's
// step kernels, which read the arguments from a step descriptor)
template <class T, class Batch>
__device__ __forceinline__ void small_batch_run(
    const Batch &batch, u32 n, int mode, typename T::Row *__restrict__ out,
    RunMeta *meta, u32 *__restrict__ dir, u32 max_bits, u32 rows_log) {
  using K = typename T::K;
  using V = typename T::V;
  constexpr int kTile = T::kSmallTile;
  constexpr int kItems = kTile / kSmallThreads;
  constexpr int kPasses = T::kBits / 8;
  // keys (2 x tile), 16-bit positions (2 x tile), per-wave digit counters:
  // 57 KB for 32-bit keys (4096 keys), 48 KB for 64-bit keys (2048)
  __shared__ K s_k0[kTile];
  __shared__ K s_k1[kTile];
  __shared__ unsigned short s_p0[kTile];
  __shared__ unsigned short s_p1[kTile];
  __shared__ u32 s_hist[kSmallWaves * 256];
  __shared__ u32 s_start[256];
  __shared__ u32 s_scan[32];
  __shared__ u32 s_misc[8];
  const u32 tid = threadIdx.x;
  const u32 lane = lane_id();
  const u32 wave = tid / kWave;

  for (u32 i = tid; i < n; i += kSmallThreads) {
    s_k0[i] = batch.key(i);
    s_p0[i] = static_cast<unsigned short>(i);
  }
  __syncthreads();
  K *src_k = s_k0, *dst_k = s_k1;
  unsigned short *src_p = s_p0, *dst_p = s_p1;
  for (int pass = 0; pass < kPasses; ++pass) {
    const u32 shift = 8u * pass;
    for (u32 i = tid; i < kSmallWaves * 256; i += kSmallThreads) s_hist[i] = 0;
    __syncthreads();
    K key[kItems];
    u32 pos[kItems], rank[kItems];
    u32 *my_hist = s_hist + wave * 256;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const u32 idx = wave * (kWave * kItems) + j * kWave + lane;
      const bool valid = idx < n;
      key[j] = valid ? src_k[idx] : T::kAllOnes;
      pos[j] = valid ? src_p[idx] : 0u;
      const u32 d = static_cast<u32>(key[j] >> shift) & 0xffu;
      u64 match = ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const u64 vote = ballot(bit);
        match &= bit ? vote : ~vote;
      }
      u32 prior = 0;
      if (valid) prior = lds_load(&my_hist[d]);
      rank[j] = prior + popc64(match & lanemask_lt());
      if (valid && (match >> lane) == 1ull)
        lds_store(&my_hist[d], prior + popc64(match));
    }
    __syncthreads();
    // wave prefixes per digit (threads 0..255 own one digit each)
    u32 total = 0;
    if (tid < 256) {
      for (int w = 0; w < kSmallWaves; ++w) {
        const u32 c = s_hist[w * 256 + tid];
        s_hist[w * 256 + tid] = total;
        total += c;
      }
    }
    // exclusive scan of the digit totals over 256 digits
    {
      const u32 v = tid < 256 ? total : 0u;
      const u32 inc = wave_inclusive_scan(v);
      if (lane == kWave - 1) s_scan[wave] = inc;
      __syncthreads();
      if (tid == 0) {
        u32 run = 0;
        for (int w = 0; w < kSmallWaves; ++w) {
          const u32 t = s_scan[w];
          s_scan[w] = run;
          run += t;
        }
      }
      __syncthreads();
      if (tid < 256) s_start[tid] = s_scan[wave] + inc - v;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const u32 idx = wave * (kWave * kItems) + j * kWave + lane;
      if (idx < n) {
        const u32 d = static_cast<u32>(key[j] >> shift) & 0xffu;
        const u32 lp = s_start[d] + s_hist[wave * 256 + d] + rank[j];
        dst_k[lp] = key[j];
        dst_p[lp] = static_cast<unsigned short>(pos[j]);
      }
    }
    __syncthreads();
    K *tk = src_k;
    unsigned short *tp = src_p;
    src_k = dst_k;
    src_p = dst_p;
    dst_k = tk;
    dst_p = tp;
  }
  // sorted (key, position) in src_*; keep the last occurrence of each key.
  // Survivor ranks: ballots per (item row j, wave), scanned by thread 0.
  u32 *s_off = s_hist;  // reuse: kItems * kSmallWaves words
  u64 mask[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const u32 i = j * kSmallThreads + tid;
    const bool keep = i < n && (i + 1 == n || src_k[i + 1] != src_k[i]);
    mask[j] = ballot(keep);
    if (lane == 0) s_off[j * kSmallWaves + wave] = popc64(mask[j]);
  }
  __syncthreads();
  if (tid == 0) {
    u32 run = 0;
    for (int k = 0; k < kItems * kSmallWaves; ++k) {
      const u32 c = s_off[k];
      s_off[k] = run;
      run += c;
    }
    s_misc[0] = run;  // count
    s_misc[1] = 0;    // split
    s_misc[2] = 0;    // tombs
  }
  __syncthreads();
  const u32 count = s_misc[0];
  u32 low = 0, tombs = 0;
  // decoded output keys go to dst_k for the directory pass
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    if (!((mask[j] >> lane) & 1ull)) continue;
    const u32 i = j * kSmallThreads + tid;
    const u32 o = s_off[j * kSmallWaves + wave] +
                  popc64(mask[j] & lanemask_lt());
    const K k = src_k[i];
    const u32 p = src_p[i];
    bool tomb;
    if (mode == kModeErase) tomb = true;
    else if (mode == kModeMixed) tomb = batch.erased(p);
    else tomb = false;
    out[o] = T::make((k & T::kLow) | (tomb ? T::kTomb : K(0)),
                     tomb ? V(0) : batch.value(p));
    dst_k[o] = k;
    low += k < T::kTomb;
    tombs += tomb;
  }
  low = wave_sum_u32(low);
  tombs = wave_sum_u32(tombs);
  if (lane == 0) {
    if (low) atomicAdd(&s_misc[1], low);
    if (tombs) atomicAdd(&s_misc[2], tombs);
  }
  __syncthreads();
  const u32 b = dir_bits_for(count, max_bits, rows_log);
  const u32 shift = T::kBits - b;
  if (tid == 0) {
    // count >= 1 (n >= 1)
    const u32 lo_b = static_cast<u32>(dst_k[0] >> shift);
    const u32 hi_b = static_cast<u32>(dst_k[count - 1] >> shift);
    meta->count = count;
    meta->split = s_misc[1];
    meta->tombs = s_misc[2];
    meta->dir_bits = b;
    meta->lo_bucket = lo_b;
    meta->hi_bucket = hi_b;
    dir[lo_b] = 0;
    dir[hi_b + 1] = count;
  }
  for (u32 t = 1 + tid; t < count; t += kSmallThreads) {
    const u32 bucket = static_cast<u32>(dst_k[t] >> shift);
    for (u32 j = static_cast<u32>(dst_k[t - 1] >> shift) + 1; j <= bucket;
         ++j)
      dir[j] = t;
  }
}

template <class T>
__global__ __launch_bounds__(kSmallThreads) void small_batch_run_kernel(
    const typename T::K *__restrict__ keys, u32 n,
    const typename T::V *__restrict__ values,
    const u8 *__restrict__ erase_flags, int mode,
    typename T::Row *__restrict__ out, RunMeta *meta, u32 *__restrict__ dir,
    u32 max_bits, u32 rows_log) {
  small_batch_run<T>(ArrayBatch<T>{keys, values, erase_flags}, n, mode, out,
                     meta, dir, max_bits, rows_log);
}

// Builds the radix directory of a run from its final count/split (one
// thread per row, grid-stride).
template <class T>
__global__ void build_directory_kernel(const typename T::Row *__restrict__ rows,
                                       RunMeta *meta, u32 *__restrict__ dir,
                                       u32 max_bits, u32 rows_log) {
  const u32 count = meta->count;
  const u32 split = meta->split;
  const u32 b = dir_bits_for(count, max_bits, rows_log);
  const u32 shift = T::kBits - b;
  const u64 gid = u64(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid == 0) {
    meta->dir_bits = b;
    if (count == 0) {
      meta->lo_bucket = 1;
      meta->hi_bucket = 0;
    } else {
      const u32 lo_b = static_cast<u32>(
          decode_key<T>(T::kw(rows[0]), 0, split) >> shift);
      const u32 hi_b = static_cast<u32>(
          decode_key<T>(T::kw(rows[count - 1]), count - 1, split) >> shift);
      meta->lo_bucket = lo_b;
      meta->hi_bucket = hi_b;
      dir[lo_b] = 0;
      dir[hi_b + 1] = count;
    }
  }
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 t = gid + 1; t < count; t += stride) {
    const u32 pos = static_cast<u32>(t);
    const u32 bucket = static_cast<u32>(
        decode_key<T>(T::kw(rows[pos]), pos, split) >> shift);
    for (u32 j = static_cast<u32>(
                     decode_key<T>(T::kw(rows[pos - 1]), pos - 1, split) >>
                     shift) +
                 1;
         j <= bucket; ++j)
      dir[j] = pos;
  }
}

// --------------------------------------------------------------- merging
// merge workgroups: 256 or 512 threads (tile = threads x items per thread)

// LDS slot of element i of an array of E-byte elements, XOR-swizzled: the
// bits of i that select its bank inside a 128-byte line of the 32 banks are
// XORed with the line number, a bijection inside every line.  Lanes whose
// accesses start 4, 8 or 16 elements apart (a thread's merge window, its
// staged output rows) then spread over the banks as consecutive lanes do,
// instead of meeting in a few.  lds_lines rounds an array up to whole lines.
template <int E>
__host__ __device__ constexpr u32 lds_swz(u32 i) {
  constexpr u32 per = 128 / E;
  constexpr u32 gran = E >= 4 ? 1 : 4 / E;
  return i ^ ((((i / per) & 31u) * gran) & (per - 1));
}
template <int E>
__host__ __device__ constexpr u32 lds_lines(u32 n) {
  constexpr u32 per = 128 / E;
  return (n + per - 1) / per * per;
}

template <class T>
struct MergeSide {
  const typename T::Row *rows;
  const RunMeta *meta;
};

template <class T>
__device__ __forceinline__ typename T::K side_key(const MergeSide<T> &s,
                                                  u32 pos, u32 split) {
  return decode_key<T>(T::kw(s.rows[pos]), pos, split);
}

// Number of A elements among the first `diag` merged elements, where A (the
// newer run) precedes B on equal keys.
template <class T>
__device__ __forceinline__ u32 merge_path(const MergeSide<T> &a, u32 ca,
                                          u32 sa, const MergeSide<T> &b,
                                          u32 cb, u32 sb, u64 diag) {
  u64 lo = diag > cb ? diag - cb : 0;
  u64 hi = diag < ca ? diag : ca;
  while (lo < hi) {
    const u64 mid = (lo + hi) >> 1;
    const auto ka = side_key(a, static_cast<u32>(mid), sa);
    const auto kb = side_key(b, static_cast<u32>(diag - mid - 1), sb);
    if (ka <= kb) lo = mid + 1;
    else hi = mid;
  }
  return static_cast<u32>(lo);
}

// Tile split points of the merge, one wave per split point: a 64-ary
// search (every lane tests one candidate, a ballot keeps the interval that
// holds the answer), so a split over 2^26 rows costs 5 rounds of loads
// instead of 26 dependent ones.  Also clears the merge's chained-scan state
// and the output run's header (one launch instead of three).
template <class T>
__global__ void merge_partition_kernel(MergeSide<T> a, MergeSide<T> b,
                                       u32 tiles, u32 tile_rows,
                                       u32 *__restrict__ parts, u64 *status,
                                       u32 *ticket, RunMeta *out_meta,
                                       u32 dir_bits) {
  const u32 lane = lane_id();
  const u64 t = (u64(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  if (t > tiles) return;
  if (lane == 0 && t < tiles) status[t] = 0;
  if (t == 0 && lane == 0) {
    *ticket = 0;
    out_meta->count = 0;
    out_meta->split = 0;
    out_meta->dir_bits = dir_bits;
    out_meta->tombs = 0;
    out_meta->lo_bucket = 1;  // empty until the first and last input
    out_meta->hi_bucket = 0;  // element write their buckets
  }
  const u32 ca = a.meta->count, cb = b.meta->count;
  const u32 sa = a.meta->split, sb = b.meta->split;
  const u64 total = u64(ca) + cb;
  u64 diag = t * tile_rows;
  if (diag > total) diag = total;
  // answer = first m in [lo, hi) with A[m] > B[diag - m - 1], else hi
  u64 lo = diag > cb ? diag - cb : 0;
  u64 hi = diag < ca ? diag : ca;
  while (hi > lo) {
    const u64 span = hi - lo;
    const u64 m = span <= kWave ? lo + lane : lo + ((span * lane) >> 6);
    bool p = false;  // lanes past the interval vote "false"
    if (m < hi)
      p = side_key(a, static_cast<u32>(m), sa) <=
          side_key(b, static_cast<u32>(diag - m - 1), sb);
    const u64 vote = ballot(p);
    const u32 c = popc64(vote);  // the true candidates form a prefix
    if (span <= kWave) {
      lo += c;
      break;
    }
    const u64 m_prev = lo + ((span * (c ? c - 1 : 0)) >> 6);
    const u64 m_next = lo + ((span * c) >> 6);
    if (c == 0) {
      hi = lo;  // A[lo] > B[...]: the answer is lo
    } else {
      lo = m_prev + 1;
      if (c < kWave) hi = m_next;
    }
  }
  if (lane == 0) parts[t] = static_cast<u32>(lo);
}

// Merges run A (newer) with run B (older) into `out`: on equal keys the A
// row survives, and tombstones are dropped when `drop_tombstones` (the
// output becomes the oldest run, GPULSMOpt.cuh:3175-3179).  A tile is
// 256 * kMergeItems output diagonals.  The output's radix directory
// (dir_bits bits) is written here: every input element, kept or dropped,
// owns the entries of the buckets in (bucket(previous input key),
// bucket(its key)] in merged order and writes them the number of survivors
// before it, which is the output row of the first key >= those bounds.
template <class T, int kMergeItems, int kMergeThreads>
__global__ __launch_bounds__(kMergeThreads) void merge_kernel(
    MergeSide<T> a, MergeSide<T> b, const u32 *__restrict__ parts,
    typename T::Row *__restrict__ out, RunMeta *out_meta,
    u32 *__restrict__ dir, u32 dir_bits, int drop_tombstones, u64 *status,
    u32 *ticket) {
  using K = typename T::K;
  using V = typename T::V;
  using Row = typename T::Row;
  constexpr int kMergeTile = kMergeThreads * kMergeItems;
  // survivors are staged as rows over the key array, then the value array
  constexpr u32 kStageHalf = (kMergeTile + 2) / 2;
  static_assert(sizeof(Row) == 2 * sizeof(K) && sizeof(K) == sizeof(V),
                "row = key word + value");
  // LDS slots are swizzled (lds_swz) so that the lanes of a wave, whose
  // merge windows and staged outputs start kMergeItems slots apart, spread
  // over the banks
  constexpr u32 kSlots = lds_lines<sizeof(K)>(kMergeTile + 2);
  constexpr u32 kStageSlots =
      lds_lines<sizeof(Row)>(kStageHalf) * sizeof(Row) / sizeof(K);
  constexpr u32 kKeySlots = kSlots > kStageSlots ? kSlots : kStageSlots;
  auto key_slot = [](u32 i) { return lds_swz<sizeof(K)>(i); };
  auto tomb_slot = [](u32 i) { return lds_swz<1>(i); };
  auto row_slot = [](u32 i) { return lds_swz<sizeof(Row)>(i); };
  __shared__ u32 s_tile;
  // inputs: decoded key, value, tombstone flag
  __shared__ __align__(16) K s_key[kKeySlots];
  __shared__ __align__(16) V s_val[kKeySlots];
  __shared__ u8 s_tomb[lds_lines<1>(kMergeTile + 2)];
  __shared__ u32 s_scan[kMergeThreads / kWave + 1];
  __shared__ u64 s_base;
  __shared__ u32 s_low, s_tombs;
  __shared__ K s_prev_b;  // B[j0 - 1]: with A[i0 - 1], the row before the tile

  if (threadIdx.x == 0) {
    s_tile = atomicAdd(ticket, 1u);
    s_low = 0;
    s_tombs = 0;
  }
  __syncthreads();
  const u32 tile = s_tile;
  const u32 ca = a.meta->count, cb = b.meta->count;
  const u32 sa = a.meta->split, sb = b.meta->split;
  const u64 total = u64(ca) + cb;
  const u64 d0 = u64(tile) * kMergeTile;
  if (d0 >= total) return;  // every later ticket is empty too (an empty
                            // output keeps the empty bucket window)
  const u64 d1 = d0 + kMergeTile < total ? d0 + kMergeTile : total;
  const u32 i0 = parts[tile], i1 = parts[tile + 1];
  const u32 j0 = static_cast<u32>(d0 - i0), j1 = static_cast<u32>(d1 - i1);
  const u32 na = i1 - i0, nb = j1 - j0;

  // LDS layout: [0] = A[i0-1] (previous A row, for the equal-key test),
  // [1, na] = A[i0, i1), [na+1, na+nb] = B[j0, j1).  Slot t of the
  // na + 1 + nb <= kMergeTile + 1 slots is loaded by thread t % 256; every
  // load of a thread is issued before the first one is used.
  constexpr int kLoads = kMergeItems + 1;
  const u32 slots = na + 1 + nb;
  Row rr[kLoads];
#pragma unroll
  for (int r = 0; r < kLoads; ++r) {
    const u32 t = threadIdx.x + r * kMergeThreads;
    if (t < slots) {
      if (t > na) rr[r] = b.rows[j0 + (t - na - 1)];
      else if (t > 0 || i0 > 0) rr[r] = a.rows[i0 + t - 1];
    }
  }
  Row prev_b_row{};
  const bool load_prev_b = threadIdx.x == kMergeThreads - 1 && j0 > 0;
  if (load_prev_b) prev_b_row = b.rows[j0 - 1];
#pragma unroll
  for (int r = 0; r < kLoads; ++r) {
    const u32 t = threadIdx.x + r * kMergeThreads;
    if (t < slots) {
      if (t == 0 && i0 == 0) {
        s_key[0] = 0;
        s_tomb[0] = 2;  // "no previous row"
      } else {
        const bool from_a = t <= na;
        const u32 pos = from_a ? i0 + t - 1 : j0 + (t - na - 1);
        s_key[key_slot(t)] =
            decode_key<T>(T::kw(rr[r]), pos, from_a ? sa : sb);
        s_tomb[tomb_slot(t)] =
            static_cast<u8>(T::kw(rr[r]) >> (T::kBits - 1));
        s_val[key_slot(t)] = T::val(rr[r]);
      }
    }
  }
  if (threadIdx.x == kMergeThreads - 1)
    s_prev_b = load_prev_b ? decode_key<T>(T::kw(prev_b_row), j0 - 1, sb)
                           : K(0);
  __syncthreads();
  // Header fields without atomics: the tile reaching the end writes the
  // count; the first tile whose inputs reach the top key bit (or the last
  // tile) writes the split, as every row before it is below the top bit and
  // every row after it is not; tiles with surviving tombstones set the flag.
  const bool last_tile = d1 == total;
  bool transition = false;
  {
    const bool before_low = (i0 == 0 || s_key[0] < T::kTomb) &&
                            (j0 == 0 || s_prev_b < T::kTomb);
    const bool tile_high = (na && s_key[key_slot(na)] >= T::kTomb) ||
                           (nb && s_key[key_slot(na + nb)] >= T::kTomb);
    transition = before_low && (tile_high || last_tile);
  }

  // Per-thread merge path inside the tile.
  const u32 tile_len = na + nb;
  const u32 dk = threadIdx.x * kMergeItems;
  u32 ia = 0, ib = 0;
  if (dk < tile_len) {
    u32 lo = dk > nb ? dk - nb : 0;
    u32 hi = dk < na ? dk : na;
    while (lo < hi) {
      const u32 mid = (lo + hi) >> 1;
      if (s_key[key_slot(1 + mid)] <=
          s_key[key_slot(na + 1 + (dk - mid - 1))])
        lo = mid + 1;
      else hi = mid;
    }
    ia = lo;
    ib = dk - lo;
  }
  // the input element before this thread's first one, in merged order
  bool has_prev = false;
  K prev = 0;
  if (dk < tile_len) {
    if (ia > 0 || i0 > 0) {
      has_prev = true;
      prev = s_key[key_slot(ia)];  // slot ia: A[ia - 1] (A[i0 - 1] at 0)
    }
    if (ib > 0 || j0 > 0) {
      const K kb = ib > 0 ? s_key[key_slot(na + ib)] : s_prev_b;
      prev = has_prev && prev > kb ? prev : kb;
      has_prev = true;
    }
  }
  K ok[kMergeItems];
  V ov[kMergeItems];
  u8 ot[kMergeItems];
  u32 osrc[kMergeItems];  // LDS slot of each merged element
  bool okeep[kMergeItems];
  u32 kept = 0;
  const u32 steps =
      dk < tile_len ? (tile_len - dk < kMergeItems ? tile_len - dk
                                                   : u32(kMergeItems))
                    : 0u;
  // Serial merge.  Registers carry the next key of each side and the last
  // A key taken (a B row equal to it is shadowed by the newer A row), so
  // the only LDS read on the serial path is the next key of the side that
  // advanced; values and tombstone flags are read afterwards, all at once.
  {
    bool a_left = ia < na, b_left = ib < nb;
    K ka = (steps && a_left) ? s_key[key_slot(1 + ia)] : K(0);
    K kb = (steps && b_left) ? s_key[key_slot(na + 1 + ib)] : K(0);
    // LDS slot ia holds A[ia - 1] (A[i0 - 1] at 0, flag 2 when none)
    bool have_last_a = steps && s_tomb[tomb_slot(ia)] != 2;
    K last_a = steps ? s_key[key_slot(ia)] : K(0);
#pragma unroll
    for (int s = 0; s < kMergeItems; ++s) {
      okeep[s] = false;
      ok[s] = 0;
      osrc[s] = 0;
      if (static_cast<u32>(s) < steps) {
        const bool take_a = a_left && (!b_left || ka <= kb);
        const K k = take_a ? ka : kb;
        ok[s] = k;
        osrc[s] = take_a ? 1 + ia : na + 1 + ib;
        okeep[s] = take_a || !(have_last_a && last_a == k);
        if (take_a) {
          last_a = k;
          have_last_a = true;
        }
        ia += take_a;
        ib += !take_a;
        a_left = ia < na;
        b_left = ib < nb;
        const bool more = take_a ? a_left : b_left;
        const K next =
            more ? s_key[key_slot(take_a ? 1 + ia : na + 1 + ib)] : K(0);
        if (take_a) ka = next;
        else kb = next;
      }
    }
  }
#pragma unroll
  for (int s = 0; s < kMergeItems; ++s) {
    ov[s] = 0;
    ot[s] = 0;
    if (static_cast<u32>(s) < steps) {
      ot[s] = s_tomb[tomb_slot(osrc[s])];
      ov[s] = s_val[key_slot(osrc[s])];
    }
  }
#pragma unroll
  for (int s = 0; s < kMergeItems; ++s) {
    if (okeep[s] && ot[s] && drop_tombstones) okeep[s] = false;
    kept += okeep[s];
  }
  const u32 local = block_exclusive_scan<kMergeThreads>(kept, s_scan);
  const u32 tile_total = s_scan[kMergeThreads / kWave];
  // Wave 0 waits on the look-back while the other waves stage their
  // survivors at tile-local positions (the scan's barriers ordered every
  // LDS input read before these writes).
  if (threadIdx.x < kWave) {
    const u64 ex = wave_lookback(status, tile, tile_total);
    if (threadIdx.x == 0) s_base = ex;
  }
  Row *stage_lo = reinterpret_cast<Row *>(s_key);
  Row *stage_hi = reinterpret_cast<Row *>(s_val);
  u32 low = 0, tombs = 0, at = local;
#pragma unroll
  for (int s = 0; s < kMergeItems; ++s) {
    if (okeep[s]) {
      const Row r =
          T::make((ok[s] & T::kLow) | (ot[s] ? T::kTomb : K(0)), ov[s]);
      if (at < kStageHalf) stage_lo[row_slot(at)] = r;
      else stage_hi[row_slot(at - kStageHalf)] = r;
      ++at;
      low += ok[s] < T::kTomb;
      tombs += ot[s];
    }
  }
  __syncthreads();
  const u64 obase = s_base;
  for (u32 t = threadIdx.x; t < tile_total; t += kMergeThreads)
    out[obase + t] = t < kStageHalf ? stage_lo[row_slot(t)]
                                    : stage_hi[row_slot(t - kStageHalf)];
  // directory entries inside the bucket window (first to last input key)
  {
    const u32 shift = T::kBits - dir_bits;
    u32 pbk = static_cast<u32>(prev >> shift);
    u32 before = static_cast<u32>(obase) + local;
#pragma unroll
    for (int s = 0; s < kMergeItems; ++s) {
      if (static_cast<u32>(s) < steps) {
        const u32 bk = static_cast<u32>(ok[s] >> shift);
        if (s == 0 && !has_prev) {  // the first input element of the merge
          out_meta->lo_bucket = bk;
          dir[bk] = 0;
        } else {
          for (u32 j = pbk + 1; j <= bk; ++j) dir[j] = before;
        }
        pbk = bk;
        before += okeep[s];
      }
    }
    if (last_tile && steps && dk + steps == tile_len) {
      // the last input element closes the window at the row count
      out_meta->hi_bucket = pbk;
      dir[pbk + 1] = static_cast<u32>(obase + tile_total);
    }
  }
  low = wave_sum_u32(low);
  tombs = wave_sum_u32(tombs);
  if (lane_id() == 0) {
    if (low) atomicAdd(&s_low, low);
    if (tombs) atomicAdd(&s_tombs, tombs);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (last_tile) out_meta->count = static_cast<u32>(obase + tile_total);
    if (transition) out_meta->split = static_cast<u32>(obase + s_low);
    if (s_tombs) out_meta->tombs = 1;
  }
}

// --------------------------------------------------------------- queries
// Query kernels read their inputs and write their results through an IO
// policy: PlainLookup / PlainRange over contiguous arrays (the C ABI), or
// the serving engine's piece tables (one launch answers every request of a
// step straight from and into the requests' own buffers).

// A miss writes all-ones (0xFFFFFFFF for 32-bit values) without a found
// array, value 0 + found 0 with one (GPULSMOpt.cuh:4714-4806).
template <class T>
struct PlainLookup {
  const typename T::K *q;
  typename T::V *out;
  u8 *found;
  // IO policies may copy a table they search per query into LDS:
  // stage(lds) returns the policy to use after the block's next barrier
  static constexpr u32 kStageBytes = 0;
  __device__ __forceinline__ PlainLookup stage(void *) const { return *this; }
  struct Slot {
    u64 i;
  };
  __device__ __forceinline__ typename T::K load(u64 i, Slot &s) const {
    s.i = i;
    return q[i];
  }
  __device__ __forceinline__ void put(const Slot &s, bool live,
                                      typename T::V value) const {
    if (found) {
      out[s.i] = live ? value : typename T::V(0);
      found[s.i] = live ? 1 : 0;
    } else {
      out[s.i] = live ? value : T::kAllOnes;
    }
  }
};

// Newest visible version of `key`: live + value, or not live.
template <class T>
__device__ __forceinline__ bool lookup_one(const RunSet<T> &runs,
                                           typename T::K key,
                                           typename T::V &value) {
  for (u32 r = 0; r < runs.count; ++r) {
    const RunView<T> &rv = runs.run[r];
    typename T::Row row;
    const RunMeta m = *rv.meta;
    if (run_find(rv, m, key, row) != 0xffffffffu) {
      value = T::val(row);
      return !(T::kw(row) & T::kTomb);
    }
  }
  return false;
}

template <class T, class IO>
__global__ void lookup_kernel(IO io_arg, u64 n, RunSet<T> runs) {
  __shared__ __align__(16) char s_io[IO::kStageBytes ? IO::kStageBytes : 16];
  const IO io = io_arg.stage(s_io);
  __syncthreads();
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    typename IO::Slot slot;
    const typename T::K key = io.load(i, slot);
    typename T::V value = 0;
    const bool live = lookup_one(runs, key, value);
    io.put(slot, live, value);
  }
}

// Result of a grouped lookup: value and found flag side by side, so the
// gather back to query order reads each result with one access.
template <class T>
struct alignas(2 * sizeof(typename T::V)) GroupedResult {
  typename T::V value;
  typename T::V found;
};

// Lookups over queries partitioned by their top varying key byte (qk =
// grouped keys).  Results are written in the grouped order (res_values[p],
// res_found[p]), so every store is coalesced; lookup_gather_kernel then
// brings them back to query order.  Scattering each result straight to its
// query position would dirty one cache line per query and evict the run
// rows the grouping keeps in L2.  Workgroups are dealt round-robin to the
// 8 XCDs, so block b serves the contiguous chunk (b % 8) * (G / 8) + b / 8:
// every XCD walks its own eighth of the grouped queries, i.e. its own eighth
// of the key space, and the run rows it needs stay in that XCD's L2.
// A thread carries kQ queries through each run in lock step (directory
// loads of all of them, then their row windows), so their memory round
// trips overlap; run headers are read once per workgroup into LDS.
template <class T>
__global__ __launch_bounds__(256) void lookup_grouped_kernel(
    const typename T::K *__restrict__ qk, u64 n,
    GroupedResult<T> *__restrict__ res, RunSet<T> runs) {
  using K = typename T::K;
  using V = typename T::V;
  using Row = typename T::Row;
  constexpr int kQ = 2;
  __shared__ RunMeta s_meta[kMaxRuns];
  if (threadIdx.x < runs.count) s_meta[threadIdx.x] = *runs.run[threadIdx.x].meta;
  __syncthreads();
  const u32 G = gridDim.x;  // a multiple of 8
  const u32 logical = (blockIdx.x % 8u) * (G / 8u) + blockIdx.x / 8u;
  const u64 chunk = (n + G - 1) / G;
  const u64 begin = u64(logical) * chunk;
  const u64 end = begin + chunk < n ? begin + chunk : n;
  for (u64 p0 = begin + threadIdx.x; p0 < end; p0 += u64(kQ) * blockDim.x) {
    K key[kQ];
    V val[kQ];
    bool open[kQ], live[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const u64 p = p0 + u64(q) * blockDim.x;
      open[q] = p < end;
      key[q] = open[q] ? qk[p] : K(0);
      live[q] = false;
      val[q] = 0;
    }
    for (u32 r = 0; r < runs.count; ++r) {
      const RunMeta m = s_meta[r];
      const RunView<T> &rv = runs.run[r];
      u32 lo[kQ], hi[kQ];
#pragma unroll
      for (int q = 0; q < kQ; ++q) {
        lo[q] = hi[q] = 0;
        if (open[q]) (void)dir_window(rv, m, key[q], lo[q], hi[q]);
      }
      bool hit[kQ];
      Row row[kQ];
#pragma unroll
      for (int q = 0; q < kQ; ++q) {
        hit[q] = false;
        // wide directory buckets (skewed keys): bisect down to the window
        const K target = key[q] & T::kLow;
        while (hi[q] - lo[q] > kFindWindow) {
          const u32 mid = (lo[q] + hi[q]) >> 1;
          const Row w = rv.rows[mid];
          const K km = T::kw(w) & T::kLow;
          if (km == target) {
            hit[q] = true;
            row[q] = w;
            lo[q] = hi[q] = mid;
          } else if (km < target) {
            lo[q] = mid + 1;
          } else {
            hi[q] = mid;
          }
        }
      }
      // the row windows: 8-byte rows are loaded in aligned pairs (one
      // 16-byte access covers two rows; row arrays carry one spare row, so
      // the pair of a window's last row stays inside the allocation)
      constexpr u32 kPair = sizeof(Row) == 8 ? 2 : 1;
      constexpr u32 kSlots = kPair == 2 ? kFindWindow + 2 : kFindWindow;
      Row w[kQ][kSlots];
      u32 first[kQ];
#pragma unroll
      for (int q = 0; q < kQ; ++q) {
        first[q] = kPair == 2 ? (lo[q] & ~1u) : lo[q];
#pragma unroll
        for (u32 i = 0; i < kSlots; i += kPair) {
          const u32 at = first[q] + i;
          if (at < hi[q]) {
            if constexpr (kPair == 2) {
              const uint4 v = *reinterpret_cast<const uint4 *>(rv.rows + at);
              w[q][i] = T::make(v.x, v.y);
              w[q][i + 1] = T::make(v.z, v.w);
            } else {
              w[q][i] = rv.rows[at];
            }
          }
        }
      }
#pragma unroll
      for (int q = 0; q < kQ; ++q) {
        const K target = key[q] & T::kLow;
#pragma unroll
        for (u32 i = 0; i < kSlots; ++i) {
          const u32 at = first[q] + i;
          if (at >= lo[q] && at < hi[q] &&
              (T::kw(w[q][i]) & T::kLow) == target) {
            hit[q] = true;
            row[q] = w[q][i];
          }
        }
        if (open[q] && hit[q]) {
          open[q] = false;
          live[q] = !(T::kw(row[q]) & T::kTomb);
          val[q] = T::val(row[q]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const u64 p = p0 + u64(q) * blockDim.x;
      if (p < end) {
        GroupedResult<T> g;
        g.value = live[q] ? val[q] : V(0);
        g.found = live[q] ? 1 : 0;
        res[p] = g;
      }
    }
  }
}

// Result i = grouped result dest[i], written through the IO policy: dest
// (the partition pass's destination of query i) is read in query order and
// the grouped results come from 256 streams (one per key group) that each
// advance through their own segment, so reads and stores both stream.
// Blocks are dealt to the XCDs like lookup_grouped_kernel's: XCD x answers
// the x-th eighth of the queries, whose results are the x-th eighth of
// every group's segment (the partition is stable), so no result line is
// fetched into more than one XCD's L2.
template <class T, class IO>
__global__ __launch_bounds__(256) void lookup_gather_kernel(
    IO io_arg, const u32 *__restrict__ dest, u64 n,
    const GroupedResult<T> *__restrict__ res) {
  __shared__ __align__(16) char s_io[IO::kStageBytes ? IO::kStageBytes : 16];
  const IO io = io_arg.stage(s_io);
  __syncthreads();
  constexpr int kQ = 4;
  const u32 G = gridDim.x;  // a multiple of 8
  const u32 logical = (blockIdx.x % 8u) * (G / 8u) + blockIdx.x / 8u;
  const u64 chunk = (n + G - 1) / G;
  const u64 begin = u64(logical) * chunk;
  const u64 end = begin + chunk < n ? begin + chunk : n;
  for (u64 p0 = begin + threadIdx.x; p0 < end; p0 += u64(kQ) * blockDim.x) {
    u32 d[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const u64 i = p0 + u64(q) * blockDim.x;
      d[q] = i < end ? dest[i] : 0u;
    }
    GroupedResult<T> g[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const u64 i = p0 + u64(q) * blockDim.x;
      if (i < end) g[q] = res[d[q]];
    }
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const u64 i = p0 + u64(q) * blockDim.x;
      if (i < end) {
        typename IO::Slot slot;
        (void)io.load(i, slot);
        io.put(slot, g[q].found != 0, g[q].value);
      }
    }
  }
}

// Workgroups of a grouped lookup over n queries: a multiple of 8 (one
// share per XCD), about four queries per thread.
inline u32 grouped_lookup_blocks(u64 n) {
  const u64 g = (n + 1023) / 1024;
  const u64 c = g < 1 ? 1 : (g > (1u << 16) ? (1u << 16) : g);
  return static_cast<u32>((c + 7) & ~u64(7));
}

// Wave per query.  Rows of run r count when they are live and no newer run
// holds the same key (newer versions and tombstones shadow them).  Sums
// accumulate in 64 bits and are truncated to the value width.
constexpr int kRangeThreads = 256;
constexpr int kRangeWaves = kRangeThreads / kWave;

template <class T>
struct PlainRange {
  const typename T::K *lo, *hi;
  typename T::V *out;
  int count_only;
  static constexpr u32 kStageBytes = 0;
  __device__ __forceinline__ PlainRange stage(void *) const { return *this; }
  struct Slot {
    u64 q;
  };
  // bounds of query q; `count` selects count (1 per live row) or sum
  __device__ __forceinline__ void load(u64 q, Slot &s, typename T::K &a,
                                       typename T::K &z, bool &count) const {
    s.q = q;
    a = lo[q];
    z = hi[q];
    count = count_only != 0;
  }
  __device__ __forceinline__ void put(const Slot &s, typename T::V v) const {
    out[s.q] = v;
  }
};

// Queries whose bounds a wave resolves together: one lane per (query,
// run, side), at least one query.
__host__ __device__ __forceinline__ u32 range_queries_per_wave(u32 runs) {
  const u32 per = 2 * (runs ? runs : 1);
  return per >= static_cast<u32>(kWave) ? 1u : static_cast<u32>(kWave) / per;
}

// One wave per round of qw queries.  The lanes first resolve the lower and
// upper bounds of the round's queries in every run at once (one lane per
// query, run and side).  A row counts when it is live and no newer run
// holds its key.  Then, for the whole round at once:
//   * the rows of every run newer than a query's oldest run with rows are
//     staged in LDS (decoded keys, tombstone bits) and set bits of a round
//     filter;
//   * a count over an oldest run without tombstones takes that run's row
//     count, minus one for every key whose newest version lies in a newer
//     run and that the oldest run also holds (one point search per newer
//     key), plus the newer keys' own newest versions that are live; only
//     the newer rows are visited;
//   * otherwise every row of the query's segments is visited, and a row of
//     a run other than the newest searches the staged keys of the newer
//     runs when its filter bit is set;
// with the work items of all queries of the round flattened over the
// lanes, so a round costs a few memory round trips however many queries
// it holds.  The count shortcut replaces the visit of the oldest run's
// rows (one coalesced load per 64 rows) by a point search per newer row
// (about three dependent round trips), so a query takes it only when its
// oldest run holds at least kRangeFastRatio times as many rows in the
// range as the newer runs.  When the staged keys of a round exceed the
// stage (kRangeStageBytes per wave), its queries are answered one at a
// time by the whole wave instead.  Sums accumulate in 64 bits and are
// truncated to the value width.
constexpr u32 kRangeStageBytes = 8192;   // staged keys per wave, in bytes
constexpr u32 kRangeFastRatio = 2;
constexpr u32 kRangeFilterLog = 13;      // filter bits per wave (log2)

template <class K>
__device__ __forceinline__ u32 range_filter_bit(K key) {
  u32 x = static_cast<u32>(key);
  if constexpr (sizeof(K) == 8) x ^= static_cast<u32>(key >> 32);
  return (x * 0x9E3779B1u) >> (32 - kRangeFilterLog);
}
// The range kernel's work on queries [q_begin, q_end) by the calling
// workgroup, whose waves are wave_id .. of a grid of wave_stride waves
// (the range kernel's grid, or the serving engine's step kernel taking
// chunks of queries).  `io` is already staged; every thread of the
// workgroup calls it.
template <class T, class IO>
__device__ __forceinline__ void range_block(const IO &io, u64 q_begin,
                                            u64 q_end, const RunSet<T> &runs,
                                            u64 wave_id, u64 wave_stride) {
  using K = typename T::K;
  constexpr u32 kRangeCache = kRangeStageBytes / sizeof(K);
  constexpr u32 kSlots =
      2 * kMaxRuns > kWave ? 2 * kMaxRuns : static_cast<u32>(kWave);
  // s_b[wave][2 * (qi * runs + r) + side]: lower (side 0) and upper (side
  // 1) bound of query qi of the round in run r
  __shared__ u32 s_b[kRangeWaves][kSlots];
  __shared__ RunMeta s_meta[kMaxRuns];
  __shared__ K s_cache[kRangeWaves][kRangeCache];
  __shared__ u32 s_filter[kRangeWaves][(1u << kRangeFilterLog) / 32];
  __shared__ u32 s_tomb[kRangeWaves][kRangeCache / 32];
  constexpr u32 kQMax = kWave / 2;  // queries per round (two bounds each)
  __shared__ typename IO::Slot s_slot[kRangeWaves][kQMax];
  // per query of the round: flags (exists, a <= z, count), plan (first |
  // oldest << 8 | fast << 16), work and staged-key offsets, accumulator
  __shared__ u32 s_flags[kRangeWaves][kQMax];
  __shared__ u32 s_plan[kRangeWaves][kQMax];
  __shared__ u32 s_woff[kRangeWaves][kQMax];
  __shared__ u32 s_coff[kRangeWaves][kQMax];
  __shared__ u32 s_stg[kRangeWaves][kQMax];  // newer rows (staged)
  __shared__ u32 s_wrk[kRangeWaves][kQMax];  // work items
  __shared__ unsigned long long s_acc[kRangeWaves][kQMax];
  const u32 nr = runs.count;
  __syncthreads();  // a previous call's readers of s_meta are done
  if (threadIdx.x < nr) s_meta[threadIdx.x] = *runs.run[threadIdx.x].meta;
  __syncthreads();
  if (nr == 0) {  // no runs: every range is empty
    const u64 stride = wave_stride * kWave;
    for (u64 q = q_begin + (wave_id - wave_id % kRangeWaves) * kWave +
                 threadIdx.x;
         q < q_end; q += stride) {
      typename IO::Slot slot;
      K a, z;
      bool count_only;
      io.load(q, slot, a, z, count_only);
      io.put(slot, typename T::V(0));
    }
    return;
  }
  const u32 lane = lane_id();
  const u32 wave = threadIdx.x / kWave;
  const u32 per = 2 * nr;  // >= 2: the bounds pass writes every query's
                           // flags and slot
  const u32 qw = range_queries_per_wave(nr);
  const u64 n = q_end;
  for (u64 q0 = q_begin + wave_id * qw; q0 < n; q0 += wave_stride * qw) {
    // every bound of the round at once, one lane each
    for (u32 t = lane; t < qw * per; t += kWave) {
      const u64 q = q0 + t / per;
      const u32 r = (t % per) >> 1;
      u32 b = 0;
      u32 flags = 0;
      typename IO::Slot slot{};
      if (q < n) {
        K a, z;
        bool count_only;
        io.load(q, slot, a, z, count_only);
        flags = 1u | (a <= z ? 2u : 0u) | (count_only ? 4u : 0u);
        if (a <= z) {
          const RunView<T> &rv = runs.run[r];
          const RunMeta m = s_meta[r];
          b = (t & 1) ? run_upper_bound(rv, m, z) : run_lower_bound(rv, m, a);
        }
      }
      s_b[wave][t] = b;
      if (t % per == 0) {
        s_slot[wave][t / per] = slot;
        s_flags[wave][t / per] = flags;
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

    // ---- plan of query q0 + lane (lanes below qw)
    u32 work = 0, staged = 0;
    if (lane < qw) {
      const u32 flags = s_flags[wave][lane];
      u32 plan = 0;
      u64 base = 0;
      if ((flags & 3u) == 3u) {
        const u32 *bnd = &s_b[wave][lane * per];
        u32 first = nr, oldest = 0;
        for (u32 r = 0; r < nr; ++r) {
          if (bnd[2 * r] < bnd[2 * r + 1]) {
            if (first == nr) first = r;
            oldest = r;
          }
        }
        if (first < nr) {
          u32 newer = 0;
          for (u32 r = first; r < oldest; ++r) newer += bnd[2 * r + 1] - bnd[2 * r];
          const u32 old_rows = bnd[2 * oldest + 1] - bnd[2 * oldest];
          const bool fast = (flags & 4u) && s_meta[oldest].tombs == 0 &&
                            old_rows >= kRangeFastRatio * newer;
          staged = newer;
          work = fast ? newer : newer + old_rows;
          if (fast) base = old_rows;
          plan = first | (oldest << 8) | (fast ? 1u << 16 : 0u);
        }
      }
      s_plan[wave][lane] = plan;
      s_acc[wave][lane] = base;
    }
    const u32 work_inc = wave_inclusive_scan(work);
    const u32 staged_inc = wave_inclusive_scan(staged);
    if (lane < qw) {
      s_woff[wave][lane] = work_inc - work;
      s_coff[wave][lane] = staged_inc - staged;
      s_stg[wave][lane] = staged;
      s_wrk[wave][lane] = work;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

    // one query by the whole wave: every row of its segments, the newer
    // runs searched in place (its newer rows exceed the stage)
    auto alone = [&](u32 qi) {
      const u64 q = q0 + qi;
      if (q >= n) return;
      typename IO::Slot slot;
      K a, z;
      bool count_only;
      io.load(q, slot, a, z, count_only);
      if (a > z) {
        if (lane == 0) io.put(slot, typename T::V(0));
        return;
      }
      const u32 *bnd = &s_b[wave][qi * per];  // [2r] lower, [2r + 1] upper
      // the newest run with rows in the range
      u32 first = nr;
      for (u32 r = 0; r < nr; ++r) {
        if (bnd[2 * r] < bnd[2 * r + 1]) {
          first = r;
          break;
        }
      }
      u64 acc = 0;
      u32 skip = nr;  // a run counted without a scan
      if (first < nr && count_only && s_meta[first].tombs == 0) {
        if (lane == 0) acc = bnd[2 * first + 1] - bnd[2 * first];
        skip = first;
      }
      u32 total = 0, oldest = first;
      for (u32 r = first; r < nr; ++r) {
        const u32 cnt = bnd[2 * r + 1] - bnd[2 * r];
        if (r != skip) total += cnt;
        if (cnt) oldest = r;
      }
      // stage the keys of runs first .. oldest - 1, run after run
      u32 cached = 0;
      for (u32 r = first; r < oldest; ++r) cached += bnd[2 * r + 1] - bnd[2 * r];
      const bool use_cache = cached <= kRangeCache;
      u32 *filter = s_filter[wave];
      if (cached) {
        // (the previous query's reads of s_cache and s_filter are done)
        __builtin_amdgcn_wave_barrier();
        for (u32 i = lane; i < (1u << kRangeFilterLog) / 32; i += kWave)
          filter[i] = 0;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (u32 f = lane; f < cached; f += kWave) {
          u32 r = first, off = f;
          while (off >= bnd[2 * r + 1] - bnd[2 * r]) {
            off -= bnd[2 * r + 1] - bnd[2 * r];
            ++r;
          }
          const u32 idx = bnd[2 * r] + off;
          const K key = decode_key<T>(T::kw(runs.run[r].rows[idx]), idx,
                                      s_meta[r].split);
          if (use_cache) s_cache[wave][f] = key;
          const u32 bit = range_filter_bit(key);
          atomicOr(&filter[bit >> 5], 1u << (bit & 31));
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
      for (u32 f = lane; f < total; f += kWave) {
        // run r and row idx of flattened row f
        u32 r = first, off = f;
        while (true) {
          if (r != skip) {
            const u32 cnt = bnd[2 * r + 1] - bnd[2 * r];
            if (off < cnt) break;
            off -= cnt;
          }
          ++r;
        }
        const u32 idx = bnd[2 * r] + off;
        const RunView<T> &rv = runs.run[r];
        const typename T::Row w = rv.rows[idx];
        if (T::kw(w) & T::kTomb) continue;
        bool shadowed = false;
        const K key = decode_key<T>(T::kw(w), idx, s_meta[r].split);
        const u32 bit = range_filter_bit(key);
        if (r > first && ((filter[bit >> 5] >> (bit & 31)) & 1u)) {
          u32 base = 0;  // staged keys of runs first .. s - 1
          for (u32 s = first; s < r && !shadowed; ++s) {
            u32 l2 = bnd[2 * s], h2 = bnd[2 * s + 1];
            const u32 cnt = h2 - l2;
            if (!cnt) continue;
            if (use_cache) {
              const K *c = &s_cache[wave][base];
              u32 lo = 0, hi = cnt;
              while (lo < hi) {
                const u32 mid = (lo + hi) >> 1;
                if (c[mid] < key) lo = mid + 1;
                else hi = mid;
              }
              shadowed = lo < cnt && c[lo] == key;
              base += cnt;
              continue;
            }
            const RunView<T> &nv = runs.run[s];
            const u32 ns = s_meta[s].split;
            const u32 end = h2;
            while (l2 < h2) {
              const u32 mid = (l2 + h2) >> 1;
              if (decode_key<T>(T::kw(nv.rows[mid]), mid, ns) < key)
                l2 = mid + 1;
              else
                h2 = mid;
            }
            if (l2 < end && decode_key<T>(T::kw(nv.rows[l2]), l2, ns) == key)
              shadowed = true;
          }
        }
        if (!shadowed) acc += count_only ? u64(1) : u64(T::val(w));
      }
      acc = wave_sum_u64(acc);
      if (lane == 0) io.put(slot, static_cast<typename T::V>(acc));
        };

    // query of a flattened index f among queries [g0, g1): the last query
    // whose offset is <= f
    auto query_in = [&](const u32 *off, u32 g0, u32 g1, u32 f) {
      u32 lo = g0, hi = g1;
      while (hi - lo > 1) {
        const u32 mid = (lo + hi) >> 1;
        if (off[mid] <= f) lo = mid;
        else hi = mid;
      }
      return lo;
    };
    u32 *filter = s_filter[wave];
    for (u32 g0 = 0; g0 < qw;) {
      // the group [g0, g1): consecutive queries whose newer rows fit the
      // stage together
      const u32 cbase = s_coff[wave][g0];
      u32 g1 = g0;
      while (g1 < qw &&
             s_coff[wave][g1] + s_stg[wave][g1] - cbase <= kRangeCache)
        ++g1;
      if (g1 == g0) {
        alone(g0);
        ++g0;
        __builtin_amdgcn_wave_barrier();
        continue;
      }
      const u32 wbase = s_woff[wave][g0];
      const u32 wend = s_woff[wave][g1 - 1] + s_wrk[wave][g1 - 1];
      const u32 ncache = s_coff[wave][g1 - 1] + s_stg[wave][g1 - 1] - cbase;
      if (ncache) {
        // (the previous group's reads of the stage and filter are done)
        __builtin_amdgcn_wave_barrier();
        for (u32 i = lane; i < (1u << kRangeFilterLog) / 32; i += kWave)
          filter[i] = 0;
        for (u32 i = lane; i < kRangeCache / 32; i += kWave)
          s_tomb[wave][i] = 0;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // stage the newer rows: query after query, run after run
        for (u32 f = cbase + lane; f < cbase + ncache; f += kWave) {
          const u32 qi = query_in(s_coff[wave], g0, g1, f);
          const u32 *bnd = &s_b[wave][qi * per];
          u32 r = s_plan[wave][qi] & 0xffu, off = f - s_coff[wave][qi];
          while (off >= bnd[2 * r + 1] - bnd[2 * r]) {
            off -= bnd[2 * r + 1] - bnd[2 * r];
            ++r;
          }
          const u32 idx = bnd[2 * r] + off;
          const typename T::K w = T::kw(runs.run[r].rows[idx]);
          const K key = decode_key<T>(w, idx, s_meta[r].split);
          const u32 ci = f - cbase;
          s_cache[wave][ci] = key;
          if (w & T::kTomb) atomicOr(&s_tomb[wave][ci >> 5], 1u << (ci & 31));
          const u32 bit = range_filter_bit(key);
          atomicOr(&filter[bit >> 5], 1u << (bit & 31));
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
      // whether a run newer than r (from `first` on) of query qi holds key
      auto newer_holds = [&](u32 qi, u32 first, u32 r, K key) {
        const u32 *bnd = &s_b[wave][qi * per];
        u32 base = s_coff[wave][qi] - cbase;
        for (u32 s2 = first; s2 < r; ++s2) {
          const u32 cnt = bnd[2 * s2 + 1] - bnd[2 * s2];
          if (!cnt) continue;
          const K *c = &s_cache[wave][base];
          u32 lo = 0, hi = cnt;
          while (lo < hi) {
            const u32 mid = (lo + hi) >> 1;
            if (c[mid] < key) lo = mid + 1;
            else hi = mid;
          }
          if (lo < cnt && c[lo] == key) return true;
          base += cnt;
        }
        return false;
      };
      for (u32 f = wbase + lane; f < wend; f += kWave) {
        const u32 qi = query_in(s_woff[wave], g0, g1, f);
        const u32 plan = s_plan[wave][qi];
        const u32 first = plan & 0xffu, oldest = (plan >> 8) & 0xffu;
        const bool fast = (plan >> 16) & 1u;
        const bool count = (s_flags[wave][qi] & 4u) != 0;
        const u32 *bnd = &s_b[wave][qi * per];
        u32 r = first, off = f - s_woff[wave][qi];
        while (off >= bnd[2 * r + 1] - bnd[2 * r]) {
          off -= bnd[2 * r + 1] - bnd[2 * r];
          ++r;
        }
        const u32 idx = bnd[2 * r] + off;
        u64 contrib = 0;
        if (fast) {
          // a newer row: staged at the same position within its query
          const u32 ci = s_coff[wave][qi] - cbase + (f - s_woff[wave][qi]);
          const K key = s_cache[wave][ci];
          if (!newer_holds(qi, first, r, key)) {
            const bool tomb = (s_tomb[wave][ci >> 5] >> (ci & 31)) & 1u;
            typename T::Row row;
            const bool in_oldest =
                run_find(runs.run[oldest], s_meta[oldest], key, row) !=
                0xffffffffu;
            contrib = u64(tomb ? 0 : 1) - u64(in_oldest ? 1 : 0);
          }
        } else {
          const typename T::Row w = runs.run[r].rows[idx];
          if (!(T::kw(w) & T::kTomb)) {
            const K key = decode_key<T>(T::kw(w), idx, s_meta[r].split);
            const u32 bit = range_filter_bit(key);
            const bool maybe =
                r > first && ((filter[bit >> 5] >> (bit & 31)) & 1u);
            if (!(maybe && newer_holds(qi, first, r, key)))
              contrib = count ? u64(1) : u64(T::val(w));
          }
        }
        if (contrib)
          atomicAdd(&s_acc[wave][qi], static_cast<unsigned long long>(contrib));
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (lane >= g0 && lane < g1 && (s_flags[wave][lane] & 1u))
        io.put(s_slot[wave][lane],
               static_cast<typename T::V>(s_acc[wave][lane]));
      __builtin_amdgcn_wave_barrier();
      g0 = g1;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

template <class T, class IO>
__global__ __launch_bounds__(kRangeThreads) void range_kernel(IO io_arg, u64 n,
                                                              RunSet<T> runs) {
  __shared__ __align__(16) char s_io[IO::kStageBytes ? IO::kStageBytes : 16];
  const IO io = io_arg.stage(s_io);  // range_block's first barrier orders it
  range_block<T, IO>(io, 0, n, runs,
                     u64(blockIdx.x) * kRangeWaves + threadIdx.x / kWave,
                     u64(gridDim.x) * kRangeWaves);
}

// Smallest live key >= query, all-ones when there is none.
template <class T>
__global__ void successor_kernel(const typename T::K *__restrict__ queries,
                                 u64 n, typename T::K *__restrict__ out,
                                 RunSet<T> runs) {
  using K = typename T::K;
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    K lower = queries[i];
    K result = T::kAllOnes;
    while (true) {
      bool any = false;
      K best = 0;
      for (u32 r = 0; r < runs.count; ++r) {
        const RunView<T> &rv = runs.run[r];
        const RunMeta m = *rv.meta;
        const u32 p = run_lower_bound(rv, m, lower);
        if (p < m.count) {
          const K k = decode_key<T>(T::kw(rv.rows[p]), p, m.split);
          if (!any || k < best) best = k;
          any = true;
        }
      }
      if (!any) break;
      // the newest run that holds `best` decides its visibility
      bool live = false;
      for (u32 r = 0; r < runs.count; ++r) {
        const RunView<T> &rv = runs.run[r];
        typename T::Row row;
        if (run_find(rv, *rv.meta, best, row) != 0xffffffffu) {
          live = !(T::kw(row) & T::kTomb);
          break;
        }
      }
      if (live) {
        result = best;
        break;
      }
      if (best == T::kAllOnes) break;
      lower = best + 1;
    }
    out[i] = result;
  }
}

// Writes a tombstone-free run as plain (key, value) arrays.
template <class T>
__global__ void decode_run_kernel(const typename T::Row *__restrict__ rows,
                                  const RunMeta *meta,
                                  typename T::K *__restrict__ ok,
                                  typename T::V *__restrict__ ov) {
  const u32 count = meta->count, split = meta->split;
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += stride) {
    const u32 pos = static_cast<u32>(i);
    const typename T::Row r = rows[pos];
    ok[i] = decode_key<T>(T::kw(r), pos, split);
    ov[i] = T::val(r);
  }
}

}  // namespace lsm
}  // namespace gdict
