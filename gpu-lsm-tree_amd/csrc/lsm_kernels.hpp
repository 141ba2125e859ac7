// Device side of the LSM dictionary (uint32 keys, uint32 values).
//
// Data model in HBM.  The dictionary is a stack of immutable sorted runs,
// newest first.  A run of c rows is three allocations' worth of streams:
//   keys[c]  u32: low 31 bits of the key | tombstone flag in bit 31
//   vals[c]  u32: value (0 for a tombstone)
//   dir      RunMeta header (count, split, directory bits, tombstones)
//            followed by a radix directory of 2^b + 1 u32 row offsets:
//            dir[j] = first row whose key >= j << (32 - b).
// Bit 31 of a key is implied by the row position (rows >= split have it
// set), which frees it for the tombstone flag, so a row is exactly 8 bytes
// like the reference's Row (GPULSMOpt.cuh:343-377).  Every search enters
// through the directory (b >= 1), and inside one directory bucket all keys
// share bit 31, so comparisons use the low 31 bits only.
#pragma once

#include "device_utils.hpp"

namespace gdict {
namespace lsm {

constexpr u32 kTombBit = 0x80000000u;
constexpr u32 kLowMask = 0x7fffffffu;
constexpr u32 kInvalid = 0xffffffffu;
constexpr int kMaxRuns = 48;
constexpr u32 kMaxDirBits = 24;

struct RunMeta {
  u32 count;
  u32 split;     // rows with key < 2^31
  u32 dir_bits;  // b, 1 <= b <= kMaxDirBits
  u32 tombs;     // tombstone rows
};
constexpr u64 kMetaBytes = 256;  // header slot in front of the directory

struct RunView {
  const u32 *keys;
  const u32 *vals;
  const u32 *dir;
  const RunMeta *meta;
};

struct RunSet {
  u32 count;
  u32 pad;
  RunView run[kMaxRuns];
};

__device__ __forceinline__ u32 decode_key(u32 w, u32 pos, u32 split) {
  return (w & kLowMask) | (pos >= split ? kTombBit : 0u);
}

__host__ __device__ __forceinline__ u32 dir_bits_for(u64 count,
                                                     u32 max_bits) {
  // about 8 rows per bucket
  u32 lg = 0;
  while (lg < 63 && (u64(1) << (lg + 1)) <= count) ++lg;
  u32 b = lg > 3 ? lg - 3 : 1;
  if (b < 1) b = 1;
  if (b > max_bits) b = max_bits;
  return b;
}

// lower_bound(x) over the whole run, x in [0, 2^32].
__device__ __forceinline__ u32 run_lower_bound(const RunView &r, u32 b,
                                               u32 count, u64 x) {
  if (x > 0xffffffffull) return count;
  const u32 key = static_cast<u32>(x);
  const u32 bucket = key >> (32u - b);
  u32 lo = r.dir[bucket];
  u32 hi = r.dir[bucket + 1];
  const u32 target = key & kLowMask;
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    if ((r.keys[mid] & kLowMask) < target) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Point search; returns the row position or kInvalid.
__device__ __forceinline__ u32 run_find(const RunView &r, u32 b, u32 key,
                                        u32 &word) {
  const u32 bucket = key >> (32u - b);
  u32 lo = r.dir[bucket];
  const u32 end = r.dir[bucket + 1];
  u32 hi = end;
  const u32 target = key & kLowMask;
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    if ((r.keys[mid] & kLowMask) < target) lo = mid + 1;
    else hi = mid;
  }
  if (lo < end) {
    const u32 w = r.keys[lo];
    if ((w & kLowMask) == target) {
      word = w;
      return lo;
    }
  }
  return kInvalid;
}

// ------------------------------------------------------------ batch runs
constexpr int kFinThreads = 256;
constexpr int kFinItems = 8;
constexpr int kFinTile = kFinThreads * kFinItems;

enum FinalizeMode : int {
  kModeInsert = 0,
  kModeErase = 1,
  kModeMixed = 2,
};

// Turns a stably sorted batch (keys sk, input positions sp) into a run:
// keep the last occurrence of each key (the newest), gather its value and
// tombstone flag, encode, and compact with a single-pass chained scan.
static __global__ __launch_bounds__(kFinThreads) void finalize_batch_kernel(
    const u32 *__restrict__ sk, const u32 *__restrict__ sp, u32 n,
    const u32 *__restrict__ values, const u8 *__restrict__ erase_flags,
    int mode, u32 *__restrict__ out_keys, u32 *__restrict__ out_vals,
    RunMeta *meta, u64 *status, u32 *ticket) {
  __shared__ u32 s_tile;
  __shared__ u32 s_scan[kFinThreads / kWave + 1];
  __shared__ u64 s_base;
  __shared__ u32 s_low, s_tomb;
  if (threadIdx.x == 0) {
    s_tile = atomicAdd(ticket, 1u);
    s_low = 0;
    s_tomb = 0;
  }
  __syncthreads();
  const u32 tile = s_tile;
  const u64 first = u64(tile) * kFinTile + u64(threadIdx.x) * kFinItems;

  u32 key[kFinItems];
  u32 keep_mask = 0;
  u32 kept = 0;
#pragma unroll
  for (int j = 0; j < kFinItems; ++j) {
    const u64 i = first + j;
    key[j] = 0;
    if (i < n) {
      key[j] = sk[i];
      const bool last = (i + 1 == n) || sk[i + 1] != key[j];
      if (last) {
        keep_mask |= 1u << j;
        ++kept;
      }
    }
  }
  const u32 local = block_exclusive_scan<kFinThreads>(kept, s_scan);
  const u32 tile_total = s_scan[kFinThreads / kWave];
  if (threadIdx.x < kWave) {
    const u64 ex = wave_lookback(status, tile, tile_total);
    if (threadIdx.x == 0) s_base = ex;
  }
  __syncthreads();
  u64 o = s_base + local;
  u32 low = 0, tombs = 0;
#pragma unroll
  for (int j = 0; j < kFinItems; ++j) {
    if (!(keep_mask >> j & 1u)) continue;
    const u32 p = sp[first + j];
    bool tomb;
    if (mode == kModeErase) tomb = true;
    else if (mode == kModeMixed) tomb = erase_flags[p] != 0;
    else tomb = false;
    const u32 v = tomb ? 0u : values[p];
    out_keys[o] = (key[j] & kLowMask) | (tomb ? kTombBit : 0u);
    out_vals[o] = v;
    low += key[j] < kTombBit;
    tombs += tomb;
    ++o;
  }
  // order-independent totals
  low = wave_sum_u32(low);
  tombs = wave_sum_u32(tombs);
  if (lane_id() == 0) {
    if (low) atomicAdd(&s_low, low);
    if (tombs) atomicAdd(&s_tomb, tombs);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (tile_total) atomicAdd(&meta->count, tile_total);
    if (s_low) atomicAdd(&meta->split, s_low);
    if (s_tomb) atomicAdd(&meta->tombs, s_tomb);
  }
}

// Builds the radix directory of a run from its final count/split.  Launch
// over max(count_ub, 2^max_bits + 1) threads (grid-stride).
static __global__ void build_directory_kernel(const u32 *__restrict__ keys,
                                       RunMeta *meta, u32 *__restrict__ dir,
                                       u32 max_bits) {
  const u32 count = meta->count;
  const u32 split = meta->split;
  const u32 b = dir_bits_for(count, max_bits);
  const u32 shift = 32u - b;
  const u32 nb = 1u << b;
  if (blockIdx.x == 0 && threadIdx.x == 0) meta->dir_bits = b;
  const long long last_bucket =
      count ? static_cast<long long>(
                  decode_key(keys[count - 1], count - 1, split) >> shift)
            : -1;
  const u64 items = count > u64(nb) + 1 ? u64(count) : u64(nb) + 1;
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 t = u64(blockIdx.x) * blockDim.x + threadIdx.x; t < items;
       t += stride) {
    if (t < count) {
      const u32 pos = static_cast<u32>(t);
      const u32 bucket = decode_key(keys[pos], pos, split) >> shift;
      const long long prev =
          pos ? static_cast<long long>(
                    decode_key(keys[pos - 1], pos - 1, split) >> shift)
              : -1;
      for (long long j = prev + 1; j <= static_cast<long long>(bucket); ++j)
        dir[j] = pos;
    }
    if (t <= nb && static_cast<long long>(t) > last_bucket) dir[t] = count;
  }
}

// --------------------------------------------------------------- merging
constexpr int kMergeThreads = 256;
constexpr int kMergeItems = 8;
constexpr int kMergeTile = kMergeThreads * kMergeItems;  // 2048

struct MergeSide {
  const u32 *keys;
  const u32 *vals;
  const RunMeta *meta;
};

__device__ __forceinline__ u32 side_key(const MergeSide &s, u32 pos,
                                        u32 split) {
  return decode_key(s.keys[pos], pos, split);
}

// Number of A elements among the first `diag` merged elements, where A (the
// newer run) precedes B on equal keys.
__device__ __forceinline__ u32 merge_path(const MergeSide &a, u32 ca, u32 sa,
                                          const MergeSide &b, u32 cb, u32 sb,
                                          u64 diag) {
  u64 lo = diag > cb ? diag - cb : 0;
  u64 hi = diag < ca ? diag : ca;
  while (lo < hi) {
    const u64 mid = (lo + hi) >> 1;
    const u32 ka = side_key(a, static_cast<u32>(mid), sa);
    const u32 kb = side_key(b, static_cast<u32>(diag - mid - 1), sb);
    if (ka <= kb) lo = mid + 1;
    else hi = mid;
  }
  return static_cast<u32>(lo);
}

static __global__ void merge_partition_kernel(MergeSide a, MergeSide b, u32 tiles,
                                       u32 *__restrict__ parts) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > tiles) return;
  const u32 ca = a.meta->count, cb = b.meta->count;
  const u64 total = u64(ca) + cb;
  u64 diag = u64(t) * kMergeTile;
  if (diag > total) diag = total;
  parts[t] = merge_path(a, ca, a.meta->split, b, cb, b.meta->split, diag);
}

// Merges run A (newer) with run B (older) into `out`: on equal keys the A
// row survives, and tombstones are dropped when `drop_tombstones` (the
// output becomes the oldest run, GPULSMOpt.cuh:3175-3179).
static __global__ __launch_bounds__(kMergeThreads) void merge_kernel(
    MergeSide a, MergeSide b, const u32 *__restrict__ parts,
    u32 *__restrict__ out_keys, u32 *__restrict__ out_vals, RunMeta *out_meta,
    int drop_tombstones, u64 *status, u32 *ticket) {
  __shared__ u32 s_tile;
  __shared__ u32 s_key[kMergeTile + 2];
  __shared__ u32 s_val[kMergeTile + 2];
  __shared__ u8 s_tomb[kMergeTile + 2];
  __shared__ u32 s_scan[kMergeThreads / kWave + 1];
  __shared__ u64 s_base;
  __shared__ u32 s_low, s_tombs;

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
  if (d0 >= total) return;  // every later ticket is empty too
  const u64 d1 = d0 + kMergeTile < total ? d0 + kMergeTile : total;
  const u32 i0 = parts[tile], i1 = parts[tile + 1];
  const u32 j0 = static_cast<u32>(d0 - i0), j1 = static_cast<u32>(d1 - i1);
  const u32 na = i1 - i0, nb = j1 - j0;

  // LDS layout: [0] = A[i0-1] (previous A row, for the equal-key test),
  // [1, na] = A[i0, i1), [na+1, na+nb] = B[j0, j1).
  for (u32 t = threadIdx.x; t < na + 1; t += kMergeThreads) {
    if (t == 0 && i0 == 0) {
      s_key[0] = 0;
      s_tomb[0] = 2;  // "no previous row"
      continue;
    }
    const u32 pos = i0 + t - 1;
    const u32 w = a.keys[pos];
    s_key[t] = decode_key(w, pos, sa);
    s_tomb[t] = static_cast<u8>(w >> 31);
    if (t) s_val[t] = a.vals[pos];
  }
  for (u32 t = threadIdx.x; t < nb; t += kMergeThreads) {
    const u32 pos = j0 + t;
    const u32 w = b.keys[pos];
    s_key[na + 1 + t] = decode_key(w, pos, sb);
    s_tomb[na + 1 + t] = static_cast<u8>(w >> 31);
    s_val[na + 1 + t] = b.vals[pos];
  }
  __syncthreads();

  // Per-thread merge path inside the tile.
  const u32 tile_len = na + nb;
  const u32 dk = threadIdx.x * kMergeItems;
  u32 ia = 0, ib = 0;
  if (dk < tile_len) {
    u32 lo = dk > nb ? dk - nb : 0;
    u32 hi = dk < na ? dk : na;
    while (lo < hi) {
      const u32 mid = (lo + hi) >> 1;
      if (s_key[1 + mid] <= s_key[na + 1 + (dk - mid - 1)]) lo = mid + 1;
      else hi = mid;
    }
    ia = lo;
    ib = dk - lo;
  }
  u32 ok[kMergeItems], ov[kMergeItems];
  u8 ot[kMergeItems];
  bool okeep[kMergeItems];
  u32 kept = 0;
  const u32 steps =
      dk < tile_len ? (tile_len - dk < kMergeItems ? tile_len - dk
                                                   : u32(kMergeItems))
                    : 0u;
#pragma unroll
  for (int s = 0; s < kMergeItems; ++s) {
    okeep[s] = false;
    ok[s] = 0;
    ov[s] = 0;
    ot[s] = 0;
    if (static_cast<u32>(s) < steps) {
      bool take_a;
      if (ia >= na) take_a = false;
      else if (ib >= nb) take_a = true;
      else take_a = s_key[1 + ia] <= s_key[na + 1 + ib];
      bool keep = true;
      if (take_a) {
        ok[s] = s_key[1 + ia];
        ov[s] = s_val[1 + ia];
        ot[s] = s_tomb[1 + ia];
        ++ia;
      } else {
        ok[s] = s_key[na + 1 + ib];
        ov[s] = s_val[na + 1 + ib];
        ot[s] = s_tomb[na + 1 + ib];
        ++ib;
        // shadowed by the newer A row immediately before it?
        if (s_tomb[ia] != 2 && s_key[ia] == ok[s]) keep = false;
      }
      if (keep && ot[s] && drop_tombstones) keep = false;
      okeep[s] = keep;
      kept += keep;
    }
  }
  const u32 local = block_exclusive_scan<kMergeThreads>(kept, s_scan);
  const u32 tile_total = s_scan[kMergeThreads / kWave];
  if (threadIdx.x < kWave) {
    const u64 ex = wave_lookback(status, tile, tile_total);
    if (threadIdx.x == 0) s_base = ex;
  }
  __syncthreads();  // also: every LDS input read is complete
  // stage survivors in LDS, then write them out contiguously
  u32 low = 0, tombs = 0, at = local;
#pragma unroll
  for (int s = 0; s < kMergeItems; ++s) {
    if (okeep[s]) {
      s_key[at] = (ok[s] & kLowMask) | (ot[s] ? kTombBit : 0u);
      s_val[at] = ov[s];
      ++at;
      low += ok[s] < kTombBit;
      tombs += ot[s];
    }
  }
  __syncthreads();
  const u64 base = s_base;
  for (u32 t = threadIdx.x; t < tile_total; t += kMergeThreads) {
    out_keys[base + t] = s_key[t];
    out_vals[base + t] = s_val[t];
  }
  low = wave_sum_u32(low);
  tombs = wave_sum_u32(tombs);
  if (lane_id() == 0) {
    if (low) atomicAdd(&s_low, low);
    if (tombs) atomicAdd(&s_tombs, tombs);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (tile_total) atomicAdd(&out_meta->count, tile_total);
    if (s_low) atomicAdd(&out_meta->split, s_low);
    if (s_tombs) atomicAdd(&out_meta->tombs, s_tombs);
  }
}

// --------------------------------------------------------------- queries
static __global__ void lookup_kernel(const u32 *__restrict__ queries, u64 n,
                              u32 *__restrict__ out_values,
                              u8 *__restrict__ out_found, RunSet runs) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const u32 key = queries[i];
    bool found = false, live = false;
    u32 value = 0;
    for (u32 r = 0; r < runs.count; ++r) {
      const RunView &rv = runs.run[r];
      const u32 b = rv.meta->dir_bits;
      u32 w;
      const u32 pos = run_find(rv, b, key, w);
      if (pos != kInvalid) {
        found = true;
        live = !(w & kTombBit);
        if (live) value = rv.vals[pos];
        break;
      }
    }
    (void)found;
    if (out_found) {
      out_values[i] = live ? value : 0u;
      out_found[i] = live ? 1 : 0;
    } else {
      out_values[i] = live ? value : kInvalid;
    }
  }
}

// Wave per query.  Rows of run r count when they are live and no newer run
// holds the same key (newer versions and tombstones shadow them).
constexpr int kRangeThreads = 256;
constexpr int kRangeWaves = kRangeThreads / kWave;

static __global__ __launch_bounds__(kRangeThreads) void range_kernel(
    const u32 *__restrict__ lo, const u32 *__restrict__ hi, u64 n,
    u32 *__restrict__ out, int count_only, RunSet runs) {
  __shared__ u32 s_lo[kRangeWaves][kMaxRuns];
  __shared__ u32 s_hi[kRangeWaves][kMaxRuns];
  const u32 lane = lane_id();
  const u32 wave = threadIdx.x / kWave;
  const u64 wave_stride = u64(gridDim.x) * kRangeWaves;
  const u32 nr = runs.count;
  for (u64 q = u64(blockIdx.x) * kRangeWaves + wave; q < n; q += wave_stride) {
    const u32 a = lo[q], z = hi[q];
    if (a > z) {
      if (lane == 0) out[q] = 0;
      continue;
    }
    // segment bounds of every run, one lane per run
    for (u32 r = lane; r < nr; r += kWave) {
      const RunView &rv = runs.run[r];
      const u32 b = rv.meta->dir_bits, c = rv.meta->count;
      s_lo[wave][r] = run_lower_bound(rv, b, c, a);
      s_hi[wave][r] = run_lower_bound(rv, b, c, u64(z) + 1);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    u64 acc = 0;
    bool newer_rows = false;
    for (u32 r = 0; r < nr; ++r) {
      const RunView &rv = runs.run[r];
      const u32 L = s_lo[wave][r], H = s_hi[wave][r];
      if (L >= H) continue;
      const u32 tombs = rv.meta->tombs;
      if (count_only && !newer_rows && tombs == 0) {
        if (lane == 0) acc += H - L;
        newer_rows = true;
        continue;
      }
      const u32 split = rv.meta->split;
      for (u32 idx = L + lane; idx < H; idx += kWave) {
        const u32 w = rv.keys[idx];
        if (w & kTombBit) continue;
        bool shadowed = false;
        if (newer_rows) {
          const u32 key = decode_key(w, idx, split);
          for (u32 s = 0; s < r && !shadowed; ++s) {
            u32 l2 = s_lo[wave][s], h2 = s_hi[wave][s];
            if (l2 >= h2) continue;
            const RunView &nv = runs.run[s];
            const u32 ns = nv.meta->split;
            while (l2 < h2) {
              const u32 mid = (l2 + h2) >> 1;
              if (decode_key(nv.keys[mid], mid, ns) < key) l2 = mid + 1;
              else h2 = mid;
            }
            if (l2 < s_hi[wave][s] &&
                decode_key(nv.keys[l2], l2, ns) == key)
              shadowed = true;
          }
        }
        if (!shadowed) acc += count_only ? 1u : rv.vals[idx];
      }
      newer_rows = true;
    }
    acc = wave_sum_u64(acc);
    if (lane == 0) out[q] = static_cast<u32>(acc);
    __builtin_amdgcn_wave_barrier();
  }
}

static __global__ void successor_kernel(const u32 *__restrict__ queries, u64 n,
                                 u32 *__restrict__ out, RunSet runs) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    u64 lower = queries[i];
    u32 result = kInvalid;
    while (lower <= 0xffffffffull) {
      u64 best = 0x100000000ull;
      for (u32 r = 0; r < runs.count; ++r) {
        const RunView &rv = runs.run[r];
        const u32 b = rv.meta->dir_bits, c = rv.meta->count;
        const u32 p = run_lower_bound(rv, b, c, lower);
        if (p < c) {
          const u64 k = decode_key(rv.keys[p], p, rv.meta->split);
          if (k < best) best = k;
        }
      }
      if (best > 0xffffffffull) break;
      // the newest run that holds `best` decides its visibility
      bool live = false;
      for (u32 r = 0; r < runs.count; ++r) {
        const RunView &rv = runs.run[r];
        u32 w;
        if (run_find(rv, rv.meta->dir_bits, static_cast<u32>(best), w) !=
            kInvalid) {
          live = !(w & kTombBit);
          break;
        }
      }
      if (live) {
        result = static_cast<u32>(best);
        break;
      }
      lower = best + 1;
    }
    out[i] = result;
  }
}

// Writes a tombstone-free run as plain (key, value) arrays.
static __global__ void decode_run_kernel(const u32 *__restrict__ keys,
                                  const u32 *__restrict__ vals,
                                  const RunMeta *meta, u32 *__restrict__ ok,
                                  u32 *__restrict__ ov) {
  const u32 count = meta->count, split = meta->split;
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += stride) {
    const u32 pos = static_cast<u32>(i);
    ok[i] = decode_key(keys[pos], pos, split);
    ov[i] = vals[pos];
  }
}

static __global__ void zero_words_kernel(u32 *p, u64 n) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = 0;
}

}  // namespace lsm
}  // namespace gdict
