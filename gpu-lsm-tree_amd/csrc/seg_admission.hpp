// Segmented admission of a 32-bit batch: the batch becomes a run in two
// sorting launches instead of four radix passes and a finalize launch.
//
//   seg_tile_hist_kernel  per tile (4096 keys): the OR of (key ^ keys[0])
//                     (the bits that vary inside the tile) and counts of
//                     the digit at the tile's own shift (the same launch
//                     clears the sort's control state and the run header)
//   seg_tile_scan_kernel  s0 = (highest varying bit of the batch) - 7;
//                     every tile's counts of the digit (key >> s0) & 0xff,
//                     derived from its own, so the 256 segments split the
//                     batch's actual key spread, not the key space; per-tile
//                     digit prefixes and segment sizes
//   onesweep pass     stable partition by that digit, carrying the values
//                     (SegPayload), so the sort needs no random gathers;
//                     its tiles read their digit offsets instead of
//                     chaining them through a look-back
//   seg_sort_kernel   one workgroup per segment, in key order (ticket =
//                     segment): an LDS LSD sort of the bits below s0, the
//                     newest-wins dedupe, tombstone decoding, row
//                     encoding, the chained scan for the output offset, the
//                     radix-directory entries of the segment's key range
//                     and the run header.
//
// A segment holds up to kSegTile keys in LDS; a larger one (heavy
// duplication: hot keys) is first deduplicated through an LDS hash table
// and then sorted in LDS; one with more distinct keys than the table takes
// is sorted by its workgroup in global memory (stable LSD, chunk by
// chunk), which is slower but exact.
#pragma once

#include "lsm_kernels.hpp"
#include "radix_sort.hpp"

namespace gdict {
namespace lsm {

constexpr int kSegThreads = 1024;
constexpr int kSegItems = 8;
constexpr int kSegTile = kSegThreads * kSegItems;  // 8192
constexpr int kSegWaves = kSegThreads / kWave;
static_assert(kSegWaves >= 2, "the look-back and the offset scan use two waves");
// Top set bit of x, 0 when x == 0.
template <class K>
__device__ __forceinline__ u32 top_bit(K x) {
  if (!x) return 0;
  if constexpr (sizeof(K) == 8) return 63u - static_cast<u32>(__clzll(x));
  else return 31u - static_cast<u32>(__clz(x));
}

// Per-tile digit counts of a spread partition, digit-major, each at the
// tile's own shift: block t (one onesweep tile, keys [t * 4096,
// (t + 1) * 4096)) writes tile_span[t] = OR over its keys of
// (key ^ keys[0]) and counts the digit (key >> st) & 0xff into
// tile_hist[d * tiles + t], st = max(h_t, 7) - 7 for the tile's top
// varying bit h_t.  Every key of the tile carries keys[0]'s bits above
// max(h_t, 7), which lets seg_tile_scan_kernel derive the counts at the
// batch's shift.  The launch also clears a[0..na), b[0..nb) and c[0..nc),
// every block its share.
template <class K>
static __global__ __launch_bounds__(rs::kThreads) void seg_tile_hist_kernel(
    const K *__restrict__ keys, u32 n, u32 *__restrict__ tile_hist,
    K *__restrict__ tile_span, u32 *a, u64 na, u32 *b, u64 nb, u32 *c,
    u64 nc) {
  static_assert(rs::kThreads == rs::kRadix &&
                    rs::kThreads * rs::kHistItems == 4096,
                "a thread per digit, one onesweep tile per block");
  __shared__ u32 s_hist[rs::kRadix];
  __shared__ K s_or[rs::kThreads / kWave];
  const u64 total = na + nb + nc;
  for (u64 i = u64(blockIdx.x) * rs::kThreads + threadIdx.x; i < total;
       i += u64(gridDim.x) * rs::kThreads) {
    if (i < na) a[i] = 0;
    else if (i < na + nb) b[i - na] = 0;
    else c[i - na - nb] = 0;
  }
  s_hist[threadIdx.x] = 0;
  const K k0 = keys[0];
  const u64 base = u64(blockIdx.x) * 4096;
  K k[rs::kHistItems];
  K acc = 0;
#pragma unroll
  for (int j = 0; j < rs::kHistItems; ++j) {
    const u64 i = base + u64(j) * rs::kThreads + threadIdx.x;
    k[j] = i < n ? keys[i] : k0;
  }
#pragma unroll
  for (int j = 0; j < rs::kHistItems; ++j) acc |= k[j] ^ k0;
#pragma unroll
  for (int offset = kWave / 2; offset > 0; offset >>= 1)
    acc |= __shfl_xor(acc, offset, kWave);
  if (lane_id() == 0) s_or[threadIdx.x / kWave] = acc;
  __syncthreads();  // also orders the counter clearing before the counts
  K o = 0;
#pragma unroll
  for (int w = 0; w < rs::kThreads / kWave; ++w) o |= s_or[w];
  const u32 h = top_bit(o);
  const u32 st = h >= 7 ? h - 7 : 0u;
#pragma unroll
  for (int j = 0; j < rs::kHistItems; ++j)
    if (base + u64(j) * rs::kThreads + threadIdx.x < n)
      atomicAdd(&s_hist[static_cast<u32>(k[j] >> st) & 0xffu], 1u);
  __syncthreads();
  tile_hist[u64(threadIdx.x) * gridDim.x + blockIdx.x] = s_hist[threadIdx.x];
  if (threadIdx.x == 0) tile_span[blockIdx.x] = o;
}

// One block per digit d at the batch's shift s0 = hg - 7, hg = max(h, 7)
// for the top bit h of the OR of the tile spreads.  Tile t's count of d
// comes from its counts at its own shift (seg_tile_hist_kernel): with
// ht = max(h_t, 7) and e = hg - ht, its keys carry keys[0]'s bits ht + 1
// .. hg, so d counts nothing unless its top e bits equal those, and then
// gathers the tile's digits whose top 8 - e bits are d's low bits (e >= 8:
// every key of the tile has keys[0]'s digit).  Writes prefix[d * tiles +
// t] = the exclusive prefix of d's counts over the tiles and totals[d];
// block 0 also stores the shift.  The onesweep pass then places tile t's
// keys of digit d at (the digit's start, a scan of the totals) +
// prefix[d * tiles + t], without a look-back.  Chunks of 4 consecutive
// tiles per thread, 1024 tiles per round.
template <class K>
static __global__ __launch_bounds__(rs::kThreads) void seg_tile_scan_kernel(
    const K *__restrict__ keys, u32 n, const u32 *__restrict__ tile_hist,
    const K *__restrict__ tile_span, u32 tiles, u32 *__restrict__ prefix,
    u32 *__restrict__ totals, u32 *__restrict__ shift_out) {
  constexpr u32 kPer = 4;
  constexpr u32 kRound = kPer * rs::kThreads;
  __shared__ u32 s_scan[rs::kThreads / kWave + 1];
  __shared__ K s_or[rs::kThreads / kWave];
  const u32 d = blockIdx.x;
  // keys[0] and the first round's spreads and counts load with the spread
  // reduction (one memory round trip when tiles <= 1024)
  const K k0 = keys[0];
  K sp0[kPer];
  u32 raw0[kPer];
  K o = 0;
#pragma unroll
  for (u32 q = 0; q < kPer; ++q) {
    const u32 t = threadIdx.x * kPer + q;
    sp0[q] = t < tiles ? tile_span[t] : K(0);
    raw0[q] = t < tiles ? tile_hist[u64(d) * tiles + t] : 0u;
    o |= sp0[q];
  }
  // the rest (tiles > 1024, large query sets): eight loads in flight
#pragma unroll 8
  for (u32 t = kRound + threadIdx.x; t < tiles; t += rs::kThreads)
    o |= tile_span[t];
#pragma unroll
  for (int offset = kWave / 2; offset > 0; offset >>= 1)
    o |= __shfl_xor(o, offset, kWave);
  if (lane_id() == 0) s_or[threadIdx.x / kWave] = o;
  __syncthreads();
  o = 0;
#pragma unroll
  for (int w = 0; w < rs::kThreads / kWave; ++w) o |= s_or[w];
  const u32 hg = top_bit(o) > 7 ? top_bit(o) : 7u;
  const u32 s0 = hg - 7;
  if (d == 0 && threadIdx.x == 0) *shift_out = s0;
  const u32 own = static_cast<u32>(k0 >> s0) & 0xffu;  // keys[0]'s digit
  // tile t's count of d from its spread and its count of its own digit d
  auto count = [&](u32 t, K span_t, u32 raw_t) -> u32 {
    const u32 h_t = top_bit(span_t);
    const u32 ht = h_t > 7 ? h_t : 7u;
    const u32 e = hg - ht;
    if (e == 0) return raw_t;
    if (e >= 8) {
      const u64 first = u64(t) * 4096;
      return d == own ? static_cast<u32>(n - first < 4096 ? n - first : 4096)
                      : 0u;
    }
    const u32 top = static_cast<u32>(k0 >> (ht + 1)) & ((1u << e) - 1u);
    if ((d >> (8 - e)) != top) return 0;
    const u32 j0 = (d & ((1u << (8 - e)) - 1u)) << e;
    u32 sum = 0;
    for (u32 j = j0; j < j0 + (1u << e); ++j)
      sum += tile_hist[u64(j) * tiles + t];
    return sum;
  };
  u32 *col = prefix + u64(d) * tiles;
  u32 carry = 0;
  for (u32 r0 = 0; r0 < tiles; r0 += kRound) {
    const u32 first = r0 + threadIdx.x * kPer;
    u32 cnt[kPer], sum = 0;
#pragma unroll
    for (u32 q = 0; q < kPer; ++q) {
      const u32 t = first + q;
      cnt[q] = 0;
      if (t < tiles) {
        const K sp = r0 == 0 ? sp0[q] : tile_span[t];
        const u32 raw = r0 == 0 ? raw0[q] : tile_hist[u64(d) * tiles + t];
        cnt[q] = count(t, sp, raw);
      }
      sum += cnt[q];
    }
    u32 run = carry + block_exclusive_scan<rs::kThreads>(sum, s_scan);
#pragma unroll
    for (u32 q = 0; q < kPer; ++q) {
      if (first + q < tiles) {
        col[first + q] = run;
        run += cnt[q];
      }
    }
    carry += s_scan[rs::kThreads / kWave];
    __syncthreads();  // s_scan is reused by the next round
  }
  if (threadIdx.x == 0) totals[d] = carry;
}

// Control words of a spread partition of n keys: the partition's own, the
// per-tile spans and the device-side shift.
template <class K>
struct SpreadLayout {
  u64 span_at, shift_at, words;
  explicit SpreadLayout(u64 n) {
    const u64 part = rs::control_words(n, 1);
    span_at = part + (part & 1);  // 8-byte aligned for 64-bit spans
    shift_at = span_at + u64(rs::tiles_for(n)) * (sizeof(K) / 4);
    words = shift_at + 2;
  }
};

// The counting launches of partition_by_spread: per-tile digit counts and
// prefixes at the keys' spread shift (in the room of the look-back state,
// which the pass does not use) and the shift itself, at control +
// SpreadLayout<K>(n).shift_at.  Nothing of the partition's control needs
// clearing, only the shift and the caller's extra region (the spans are
// written whole).
template <class K>
inline void spread_counts(const K *keys, u32 n, u32 *control, hipStream_t s,
                          u32 *extra = nullptr, u64 extra_words = 0) {
  const SpreadLayout<K> L(n);
  const rs::Control pc = rs::control_layout(control, 1);
  K *span = reinterpret_cast<K *>(control + L.span_at);
  const u32 tiles = rs::tiles_for(n);
  u32 *tile_hist = reinterpret_cast<u32 *>(pc.status);
  u32 *tile_prefix = tile_hist + u64(tiles) * rs::kRadix;
  seg_tile_hist_kernel<K><<<tiles, rs::kThreads, 0, s>>>(
      keys, n, tile_hist, span, control, 0, control + L.shift_at,
      L.words - L.shift_at, extra, extra_words);
  seg_tile_scan_kernel<K><<<rs::kRadix, rs::kThreads, 0, s>>>(
      keys, n, tile_hist, span, tiles, tile_prefix, pc.offsets,
      control + L.shift_at);
}

// Stable partition of n keys by their top 8 varying bits (the 256 groups
// split the keys' actual spread, not the key space): keys_out receives the
// grouped keys, pos_out (nullable) each one's input position, dest
// (nullable) each input key's grouped position, in input order; *shift_out
// (inside `control`) the digit shift.  `control` holds
// SpreadLayout<K>(n).words words; payload selects what pos_out carries.
// The per-tile digit prefixes give every tile its offsets, so the pass
// needs no look-back.
template <class K, class Payload>
inline void partition_by_spread(const K *keys, u32 n, K *keys_out,
                                u32 *pos_out, u32 *dest, Payload payload,
                                u32 *control, hipStream_t s, int block,
                                u32 *extra = nullptr, u64 extra_words = 0) {
  spread_counts<K>(keys, n, control, s, extra, extra_words);
  const SpreadLayout<K> L(n);
  const rs::Control pc = rs::control_layout(control, 1);
  const u32 tiles = rs::tiles_for(n);
  const u32 *tile_prefix =
      reinterpret_cast<u32 *>(pc.status) + u64(tiles) * rs::kRadix;
  rs::launch_pass<K>(block ? block : rs::block_for(n), tiles, keys, payload,
                     keys_out, pos_out, n, 0, pc.offsets, pc.status,
                     pc.tickets, s, dest, control + L.shift_at,
                     tile_prefix);
}

// Payload of the partition pass.  When s0 <= 23 the top varying key bit
// is at most bit 30, so every key of the batch shares bit 31 and the pass
// may carry the erase flag there:
//   insert            the update's value
//   erase             nothing (every row is a tombstone)
//   mixed, s0 <= 23   the value; the erase flag rides in key bit 31
//   mixed, s0 = 24    the input position (value and flag are gathered
//                     after the sort)
constexpr u32 kStashBit = 0x80000000u;

__device__ __forceinline__ bool seg_stashed(int mode, u32 s0) {
  return mode == kModeMixed && s0 <= 23;
}

struct SegPayload {
  const u32 *values;
  const u8 *flags;
  int mode;
  template <class K>
  __device__ __forceinline__ u32 operator()(u64 idx, bool valid, K &key,
                                            u32 shift) const {
    if (!valid || mode == kModeErase) return 0u;
    if (mode == kModeMixed) {
      if (shift > 23) return static_cast<u32>(idx);
      key = (key & ~kStashBit) | (flags[idx] ? kStashBit : 0u);
    }
    return values[idx];
  }
};

// Output offset of segment `seg` of the 256: the sum of the survivor counts
// the segments before it publish in status[] (each publishes once, as soon
// as it has counted its survivors, and waits for nothing later).  One wave
// reads all of them in one round, four per lane, instead of walking back a
// window at a time.
__device__ __forceinline__ u64 segments_before(u64 *status, u32 seg) {
  const u32 lane = lane_id();
  u64 sum = 0;
#pragma unroll
  for (u32 k = 0; k < 4; ++k) {
    const u32 t = lane * 4 + k;
    if (t < seg) {
      u64 word;
      do {
        word = load_relaxed(&status[t]);
      } while ((word >> kFlagShift) == 0);
      sum += word & kValueMask;
    }
  }
  return wave_sum_u64(sum);
}

// Shared state of seg_sort_kernel.
struct SegShared {
  u32 k[2][kSegTile];
  u32 p[2][kSegTile];
  u32 wave_hist[kSegWaves][256];
  u32 digit_start[256];  // chunk-local start of each digit
  u32 run[256];          // running global offsets (global-memory sort)
  u32 scan[kSegWaves + 1];
  u32 off[kSegItems][kSegWaves];
  u32 misc[8];
  u64 base;
};

// Stable ranking of one chunk of `count` elements, ipt per lane: wave w
// holds chunk elements [w*64*ipt, (w+1)*64*ipt), item j of lane l is
// w*64*ipt + j*64 + l, so all waves share the work of a short chunk.
// hist takes the per-wave digit counts (cleared on entry: every reader of
// the previous round has passed a barrier since).  On return dst_local[j]
// is each item's position in the chunk sorted by its digit and
// sh.digit_start the chunk-local start of each digit.
__device__ __forceinline__ void seg_rank(SegShared &sh, u32 (*hist)[256],
                                         const u32 (&key)[kSegItems],
                                         u32 count, u32 ipt, u32 shift,
                                         u32 (&dst_local)[kSegItems]) {
  const u32 tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const u32 wave_base = wave * kWave * ipt;
  for (u32 i = tid; i < kSegWaves * 256; i += kSegThreads)
    (&hist[0][0])[i] = 0;
  __syncthreads();
  u32 *my_hist = hist[wave];
  u32 rank[kSegItems];
#pragma unroll
  for (int j = 0; j < kSegItems; ++j) {
    if (static_cast<u32>(j) < ipt) {
      const bool valid = wave_base + j * kWave + lane < count;
      const u32 d = (key[j] >> shift) & 0xffu;
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
  }
  __syncthreads();
  // per digit (threads 0..255): counts of the waves, digit start from a
  // scan over digits, then hist[w][d] = start of digit d in wave w
  u32 c[kSegWaves];
  u32 total = 0;
  if (tid < 256) {
#pragma unroll
    for (int w = 0; w < kSegWaves; ++w) {
      c[w] = hist[w][tid];
      total += c[w];
    }
  }
  const u32 inclusive = wave_inclusive_scan(total);
  if (tid < 256 && lane == kWave - 1) sh.scan[wave] = inclusive;
  __syncthreads();
  if (tid < 256) {
    u32 start = inclusive - total;
    for (u32 w = 0; w < wave; ++w) start += sh.scan[w];
    sh.digit_start[tid] = start;
#pragma unroll
    for (int w = 0; w < kSegWaves; ++w) {
      hist[w][tid] = start;
      start += c[w];
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSegItems; ++j) {
    if (static_cast<u32>(j) < ipt) {
      const u32 d = (key[j] >> shift) & 0xffu;
      dst_local[j] = hist[wave][d] + rank[j];
    }
  }
}

// Distinct keys of a segment beyond the LDS tile (hot keys: many updates
// of a few keys), through an LDS hash table of key -> last index in the
// segment; the stable partition keeps the segment in input order, so the
// last index is the newest update.  On success the distinct keys, each
// with the key word and payload of its last occurrence, are in sh.k[0] /
// sh.p[0] (in table order) and `count` holds their number.  Returns false
// when the segment has more than kHashMax distinct keys (it is then sorted
// in global memory).  The lanes of a wave holding its first lane's key
// fold into one insert, so a hot key costs one table update per wave.
constexpr u32 kHashSlots = kSegTile;  // 2^13 (u64 slots over sh.k)
constexpr u32 kHashMax = kHashSlots / 4 * 3;

__device__ __forceinline__ bool seg_dedupe(SegShared &sh,
                                           const u32 *__restrict__ keys,
                                           const u32 *__restrict__ pays, u32 m,
                                           u32 key_mask, u32 &count) {
  static_assert(sizeof(sh.k) == kHashSlots * sizeof(u64), "table over sh.k");
  const u32 tid = threadIdx.x, lane = lane_id();
  u64 *slot = reinterpret_cast<u64 *>(&sh.k[0][0]);
  u32 *last = sh.p[0];
  for (u32 i = tid; i < kHashSlots; i += kSegThreads) {
    slot[i] = 0;
    last[i] = 0;
  }
  if (tid == 0) {
    sh.misc[4] = 0;  // distinct keys
    sh.misc[5] = 0;  // overflow
  }
  __syncthreads();
  for (u32 c0 = 0; c0 < m; c0 += kSegTile) {
    u32 k[kSegItems];
#pragma unroll
    for (int j = 0; j < kSegItems; ++j) {
      const u32 i = c0 + j * kSegThreads + tid;
      k[j] = i < m ? keys[i] & key_mask : 0u;
    }
#pragma unroll
    for (int j = 0; j < kSegItems; ++j) {
      const u32 i = c0 + j * kSegThreads + tid;
      const bool valid = i < m;  // a prefix of the wave's lanes
      const u32 lead = __builtin_amdgcn_readfirstlane(k[j]);
      const bool same = valid && k[j] == lead;
      const u64 sm = ballot(same);
      u32 idx = i;
      bool skip = false;
      if (popc64(sm) > 1) {
        u32 v = same ? i : 0u;
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
          const u32 o = __shfl_xor(v, off, kWave);
          v = o > v ? o : v;
        }
        if (same) idx = v;
        skip = same && lane != ffs64(sm) - 1;
      }
      if (!valid || skip || sh.misc[5]) continue;
      const u64 tag = (u64(1) << 32) | k[j];
      u32 h = (k[j] * 0x9E3779B1u) >> 19;  // 13 bits
      bool placed = false;
      for (u32 probe = 0; probe < kHashSlots && !placed; ++probe) {
        u64 cur = __hip_atomic_load(&slot[h], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == 0) {
          cur = atomicCAS(reinterpret_cast<unsigned long long *>(&slot[h]),
                          0ull, static_cast<unsigned long long>(tag));
          if (cur == 0) {
            if (atomicAdd(&sh.misc[4], 1u) >= kHashMax) sh.misc[5] = 1;
            cur = tag;
          }
        }
        if (cur == tag) {
          atomicMax(&last[h], idx);
          placed = true;
        } else {
          h = (h + 1) & (kHashSlots - 1);
        }
      }
      if (!placed) sh.misc[5] = 1;
    }
    __syncthreads();
    if (sh.misc[5]) return false;  // uniform: read after the barrier
  }
  // compact the occupied slots: raw key word and payload of each key's
  // last occurrence
  bool occ[kSegItems];
  u32 at[kSegItems];
  u32 mine = 0;
#pragma unroll
  for (int j = 0; j < kSegItems; ++j) {
    const u32 s = j * kSegThreads + tid;
    occ[j] = slot[s] != 0;
    at[j] = last[s];
    mine += occ[j];
  }
  u32 pos = block_exclusive_scan<kSegThreads>(mine, sh.scan);
  count = sh.scan[kSegWaves];
  u32 rk[kSegItems], rp[kSegItems];
#pragma unroll
  for (int j = 0; j < kSegItems; ++j) {
    if (occ[j]) {
      rk[j] = keys[at[j]];
      rp[j] = pays[at[j]];
    }
  }
#pragma unroll
  for (int j = 0; j < kSegItems; ++j) {
    if (occ[j]) {
      sh.k[0][pos] = rk[j];
      sh.p[0][pos] = rp[j];
      ++pos;
    }
  }
  __syncthreads();
  return true;
}

// One workgroup per segment (ticket order = key order).  keys_in: the
// batch (only its first key is read, for the shared high bits); pk/pp: the
// partitioned keys and SegPayload payloads; tk/tp: scratch of n entries
// for segments too large for LDS.  offsets: segment starts (exclusive scan
// of the segment histogram).  Writes the run (rows, header, directory).
// Sorts a segment of m <= kSegTile keys by their bits below s0 into
// sh.k[0] / sh.p[0], stably (equal keys keep their input order), without
// the LSD passes: one counting pass over kBuckets buckets (the top
// min(s0, 12) of those bits; about one key per bucket for a 4096-key
// segment), every key claims a slot of its bucket, then takes the place
// its (key, input index) rank inside the bucket gives it.  Returns false, having changed
// nothing the LSD passes read, when a bucket holds more than kBucketMax
// keys (dense or repeated keys).
constexpr u32 kBuckets = kSegWaves * 256;  // counters over sh.wave_hist
constexpr u32 kBucketMax = 16;

__device__ __forceinline__ bool seg_bucket_sort(SegShared &sh,
                                                const u32 *__restrict__ keys,
                                                const u32 *__restrict__ pays,
                                                u32 m, u32 s0) {
  static_assert(kBuckets == 4096 && kBuckets % kSegThreads == 0,
                "4096 buckets, an equal share per thread");
  constexpr u32 kPer = kBuckets / kSegThreads;
  const u32 tid = threadIdx.x, lane = lane_id();
  u32 *cnt = &sh.wave_hist[0][0];
  const u32 hb = s0 < 12 ? s0 : 12;
  const u32 lsh = s0 - hb;
  const u32 low = (1u << s0) - 1u;  // the bits that order a segment
  for (u32 i = tid; i < kBuckets; i += kSegThreads) cnt[i] = 0;
  if (tid == 0) sh.misc[6] = 0;  // largest bucket
  __syncthreads();
  u32 k[kSegItems], p[kSegItems], b[kSegItems];
#pragma unroll
  for (int j = 0; j < kSegItems; ++j) {
    const u32 i = tid + j * kSegThreads;
    if (i < m) {
      k[j] = keys[i];
      p[j] = pays[i];
    }
  }
#pragma unroll
  for (int j = 0; j < kSegItems; ++j) {
    const u32 i = tid + j * kSegThreads;
    b[j] = 0;
    if (i < m) {
      b[j] = (k[j] & low) >> lsh;
      atomicAdd(&cnt[b[j]], 1u);
    }
  }
  __syncthreads();
  // bucket starts: thread t owns buckets [t * kPer, (t + 1) * kPer)
  u32 c[kPer], sum = 0, most = 0;
#pragma unroll
  for (u32 q = 0; q < kPer; ++q) {
    c[q] = cnt[tid * kPer + q];
    sum += c[q];
    most = c[q] > most ? c[q] : most;
  }
  const u32 ex = block_exclusive_scan<kSegThreads>(sum, sh.scan);
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const u32 o = __shfl_xor(most, off, kWave);
    most = o > most ? o : most;
  }
  if (lane == 0 && most) atomicMax(&sh.misc[6], most);
  u32 run = ex;
#pragma unroll
  for (u32 q = 0; q < kPer; ++q) {
    cnt[tid * kPer + q] = run;
    run += c[q];
  }
  __syncthreads();
  if (sh.misc[6] > kBucketMax) return false;  // uniform
  // every key claims a slot of its bucket in sh.k[1] / sh.p[1] (key, input
  // index); cnt[b] then ends at the bucket's end
  u32 slot[kSegItems];
#pragma unroll
  for (int j = 0; j < kSegItems; ++j) {
    const u32 i = tid + j * kSegThreads;
    slot[j] = 0;
    if (i < m) slot[j] = atomicAdd(&cnt[b[j]], 1u);
  }
#pragma unroll
  for (int j = 0; j < kSegItems; ++j) {
    const u32 i = tid + j * kSegThreads;
    if (i < m) {
      sh.k[1][slot[j]] = k[j];
      sh.p[1][slot[j]] = i;
    }
  }
  __syncthreads();
  // each key's place: its bucket's start plus the keys of the bucket before
  // it in (key, input index) order (no dependent steps, unlike insertion)
#pragma unroll
  for (int j = 0; j < kSegItems; ++j) {
    const u32 i = tid + j * kSegThreads;
    if (i < m) {
      const u32 end = cnt[b[j]];
      const u32 beg = b[j] ? cnt[b[j] - 1] : 0u;
      const u32 kx = k[j] & low;
      u32 rank = 0;
      for (u32 y = beg; y < end; ++y) {
        const u32 ky = sh.k[1][y] & low;
        rank += (ky < kx || (ky == kx && sh.p[1][y] < i)) ? 1u : 0u;
      }
      sh.k[0][beg + rank] = k[j];
      sh.p[0][beg + rank] = p[j];
    }
  }
  __syncthreads();
  return true;
}

template <class T>
__global__ __launch_bounds__(kSegThreads) void seg_sort_kernel(
    const u32 *__restrict__ keys_in, u32 *pk, u32 *pp, u32 *tk, u32 *tp,
    u32 n, const u32 *__restrict__ totals,
    const u32 *shift_in, const u32 *__restrict__ values,
    const u8 *__restrict__ erase_flags, int mode,
    typename T::Row *__restrict__ out, RunMeta *meta, u32 *__restrict__ dir,
    u32 dir_bits, u64 *status, u32 *ticket) {
  static_assert(sizeof(typename T::K) == 4, "32-bit keys");
  __shared__ SegShared sh;
  __shared__ u32 s_ticket;
  const u32 tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  if (tid == 0) s_ticket = atomicAdd(ticket, 1u);
  __syncthreads();
  // grid = 256: one ticket per segment.  Values derived from it are
  // wave-uniform; saying so keeps them (and the per-item predicates on
  // them) in scalar registers instead of per-lane masks
  const u32 seg = __builtin_amdgcn_readfirstlane(s_ticket);
  const u32 s0 = *shift_in;
  // the segment's start: the keys of the segments before it (a scan of
  // the per-segment totals)
  {
    const u32 mine = tid < 256 ? totals[tid] : 0u;
    const u32 ex = block_exclusive_scan<kSegThreads>(mine, sh.scan);
    if (tid == seg) sh.misc[7] = ex;
    __syncthreads();
  }
  const u32 start = __builtin_amdgcn_readfirstlane(sh.misc[7]);
  const u32 end = start + totals[seg];
  // elements sorted below: the segment, or its distinct keys when a
  // segment beyond the LDS tile is deduplicated first
  u32 m = end - start;
  const u32 passes = (s0 + 7) / 8;
  // key bit 31 of the sorted keys may carry the erase flag; restore()
  // gives the key itself
  const bool stashed = seg_stashed(mode, s0);
  const bool gathered = mode == kModeMixed && !stashed;
  const u32 key_mask = stashed ? ~kStashBit : ~0u;
  const u32 high_bit = keys_in[0] & kStashBit;
  auto restore = [&](u32 x) { return stashed ? (x & ~kStashBit) | high_bit : x; };

  // ---------------------------------------------------------------- sort
  // sorted keys / payloads of the segment: in LDS (cur) or in global
  // memory (gk, gp)
  int cur = 0;
  const u32 *gk = nullptr, *gp = nullptr;
  bool in_lds = m <= static_cast<u32>(kSegTile);
  bool staged = false;  // the elements are in sh.k[0] / sh.p[0] already
  if (!in_lds) {
    u32 distinct = 0;
    if (seg_dedupe(sh, pk + start, pp + start, m, key_mask, distinct)) {
      m = __builtin_amdgcn_readfirstlane(distinct);
      in_lds = staged = true;
    }
  }
  if (in_lds) {
    const u32 ipt = __builtin_amdgcn_readfirstlane(
        m ? (m + kSegThreads - 1) / kSegThreads : 1u);
    const u32 wave_base = wave * kWave * ipt;
    // the bucket sort handles every segment without dense keys; the LSD
    // passes below take the rest
    const bool bucketed =
        passes >= 2 && !staged && m && seg_bucket_sort(sh, pk + start,
                                                        pp + start, m, s0);
    if (passes == 0 && !staged) {
      // nothing to sort (each segment holds one key): copy it to LDS, all
      // loads of a thread issued before the first store
      u32 lk[kSegItems], lp[kSegItems];
#pragma unroll
      for (int j = 0; j < kSegItems; ++j) {
        const u32 i = tid + j * kSegThreads;
        if (i < m) {
          lk[j] = pk[start + i];
          lp[j] = pp[start + i];
        }
      }
#pragma unroll
      for (int j = 0; j < kSegItems; ++j) {
        const u32 i = tid + j * kSegThreads;
        if (i < m) {
          sh.k[0][i] = lk[j];
          sh.p[0][i] = lp[j];
        }
      }
      __syncthreads();
    }
    for (u32 pass = 0; pass < (bucketed ? 0u : passes); ++pass) {
      // the first pass reads the partitioned segment from global memory,
      // the later ones the previous pass's LDS arrays
      u32 key[kSegItems], pay[kSegItems], dst[kSegItems];
      if (pass == 0 && !staged) {
#pragma unroll
        for (int j = 0; j < kSegItems; ++j) {
          const u32 idx = wave_base + j * kWave + lane;
          const bool valid = static_cast<u32>(j) < ipt && idx < m;
          key[j] = valid ? pk[start + idx] : 0xffffffffu;
          pay[j] = valid ? pp[start + idx] : 0u;
        }
      } else {
#pragma unroll
        for (int j = 0; j < kSegItems; ++j) {
          const u32 idx = wave_base + j * kWave + lane;
          const bool valid = static_cast<u32>(j) < ipt && idx < m;
          key[j] = valid ? sh.k[cur][idx] : 0xffffffffu;
          pay[j] = valid ? sh.p[cur][idx] : 0u;
        }
      }
      seg_rank(sh, sh.wave_hist, key, m, ipt, 8 * pass, dst);
#pragma unroll
      for (int j = 0; j < kSegItems; ++j) {
        const u32 idx = wave_base + j * kWave + lane;
        if (static_cast<u32>(j) < ipt && idx < m) {
          sh.k[cur ^ 1][dst[j]] = key[j];
          sh.p[cur ^ 1][dst[j]] = pay[j];
        }
      }
      __syncthreads();
      cur ^= 1;
    }
  } else {
    // stable LSD in global memory, chunk by chunk, ping-pong pk <-> tk
    u32 *src_k = pk + start, *src_p = pp + start;
    u32 *dst_k = tk + start, *dst_p = tp + start;
    for (u32 pass = 0; pass < passes; ++pass) {
      const u32 shift = 8 * pass;
      for (u32 i = tid; i < 256; i += kSegThreads) sh.run[i] = 0;
      __syncthreads();
      for (u32 i = tid; i < m; i += kSegThreads)
        atomicAdd(&sh.run[(src_k[i] >> shift) & 0xffu], 1u);
      __syncthreads();
      {
        const u32 v = tid < 256 ? sh.run[tid] : 0u;
        const u32 ex = block_exclusive_scan<kSegThreads>(v, sh.scan);
        if (tid < 256) sh.run[tid] = ex;
        __syncthreads();
      }
      for (u32 c0 = 0; c0 < m; c0 += kSegTile) {
        const u32 cnt = m - c0 < u32(kSegTile) ? m - c0 : u32(kSegTile);
        const u32 ipt = __builtin_amdgcn_readfirstlane(
            (cnt + kSegThreads - 1) / kSegThreads);
        const u32 wave_base = wave * kWave * ipt;
        u32 key[kSegItems], pay[kSegItems], dst[kSegItems];
#pragma unroll
        for (int j = 0; j < kSegItems; ++j) {
          const u32 idx = wave_base + j * kWave + lane;
          const bool valid = static_cast<u32>(j) < ipt && idx < cnt;
          key[j] = valid ? src_k[c0 + idx] : 0xffffffffu;
          pay[j] = valid ? src_p[c0 + idx] : 0u;
        }
        seg_rank(sh, sh.wave_hist, key, cnt, ipt, shift, dst);
#pragma unroll
        for (int j = 0; j < kSegItems; ++j) {
          const u32 idx = wave_base + j * kWave + lane;
          if (static_cast<u32>(j) < ipt && idx < cnt) {
            const u32 d = (key[j] >> shift) & 0xffu;
            const u32 o = sh.run[d] + dst[j] - sh.digit_start[d];
            dst_k[o] = key[j];
            dst_p[o] = pay[j];
          }
        }
        __syncthreads();
        // advance the running offsets by this chunk's digit counts
        if (tid < 256) {
          const u32 next = tid < 255 ? sh.digit_start[tid + 1] : cnt;
          sh.run[tid] += next - sh.digit_start[tid];
        }
        __syncthreads();
      }
      u32 *t0 = src_k, *t1 = src_p;
      src_k = dst_k;
      src_p = dst_p;
      dst_k = t0;
      dst_p = t1;
      __threadfence_block();
      __syncthreads();
    }
    gk = src_k;
    gp = src_p;
  }

  // The rest runs once per location of the sorted segment, so that each
  // copy reads it with LDS or global loads rather than generic ones.
  auto finish = [&](const u32 *sk, const u32 *sp) {
    // ------------------------------------------- survivors: count, offset
    // keep the last occurrence of each key (stable sort: the newest update)
    auto keep_at = [&](u32 i) {
      return i < m && (i + 1 == m || ((sk[i + 1] ^ sk[i]) & key_mask) != 0);
    };
    u32 total = 0;
    for (u32 c0 = 0; c0 < m; c0 += kSegTile) {
      u32 cnt = 0;
#pragma unroll
      for (int j = 0; j < kSegItems; ++j)
        cnt += keep_at(c0 + j * kSegThreads + tid);
      total += wave_sum_u32(cnt);
    }
    if (lane == 0) sh.scan[wave] = total;
    __syncthreads();
    if (tid == 0) {
      u32 t = 0;
      for (int w = 0; w < kSegWaves; ++w) t += sh.scan[w];
      sh.misc[0] = t;
    }
    __syncthreads();
    const u32 seg_total = sh.misc[0];
    if (tid == 0) {
      sh.misc[1] = 0;  // survivors below the top key bit
      sh.misc[2] = 0;  // tombstones
      store_relaxed(&status[seg], kFlagAggregate | seg_total);
    }
    // the offset over the segments before (wave 0) is gathered beside the
    // first chunk's survivor offsets (wave 1); an empty segment gathers it
    // on its own
    auto lookback = [&]() {
      if (wave == 0) {
        const u64 ex = segments_before(status, seg);
        if (tid == 0) sh.base = ex;
      }
    };
    u64 base = 0;
    if (m == 0) {
      lookback();
      __syncthreads();
      base = sh.base;
    }

    // ------------------------------------------- rows and directory entries
    const u32 shift = 32u - dir_bits;
    u32 low = 0, tombs = 0;
    u32 running = 0;  // survivors of earlier chunks
    for (u32 c0 = 0; c0 < m; c0 += kSegTile) {
      // per-lane keep bits (bit j: item j), turned into wave masks where
      // needed: sixteen live 64-bit ballots would crowd the scalar registers
      u32 keep = 0;
#pragma unroll
      for (int j = 0; j < kSegItems; ++j) {
        const bool kj = keep_at(c0 + j * kSegThreads + tid);
        keep |= kj ? (1u << j) : 0u;
        const u64 mj = ballot(kj);
        if (lane == 0) sh.off[j][wave] = popc64(mj);
      }
      __syncthreads();
      if (c0 == 0) lookback();
      // exclusive scan of the (item, wave) survivor counts in row order, by
      // wave 1: lane l owns entries 2l and 2l + 1
      if (wave == 1) {
        u32 *flat = &sh.off[0][0];
        const u32 a = flat[2 * lane], b = flat[2 * lane + 1];
        const u32 inc = wave_inclusive_scan(a + b);
        const u32 ex = running + inc - (a + b);
        flat[2 * lane] = ex;
        flat[2 * lane + 1] = ex + a;
        if (lane == kWave - 1) sh.misc[3] = running + inc;
      }
      __syncthreads();
      base = sh.base;
      u32 k[kSegItems], prev[kSegItems], pay[kSegItems];
      bool tomb[kSegItems];
#pragma unroll
      for (int j = 0; j < kSegItems; ++j) {
        const u32 i = c0 + j * kSegThreads + tid;
        const u32 raw = i < m ? sk[i] : 0u;
        k[j] = restore(raw);
        prev[j] = (i < m && i > 0) ? restore(sk[i - 1]) : k[j];
        pay[j] = (i < m && ((keep >> j) & 1u)) ? sp[i] : 0u;
        tomb[j] = mode == kModeErase || (stashed && (raw & kStashBit));
      }
      // a position payload: gather value and flag (all items issued before
      // any row is written, so their latencies overlap)
      u32 val[kSegItems];
#pragma unroll
      for (int j = 0; j < kSegItems; ++j) {
        const u32 i = c0 + j * kSegThreads + tid;
        const bool kept = i < m && ((keep >> j) & 1u);
        if (gathered) {
          tomb[j] = kept && erase_flags[pay[j]] != 0;
          val[j] = kept ? values[pay[j]] : 0u;
        } else {
          val[j] = pay[j];
        }
      }
#pragma unroll
      for (int j = 0; j < kSegItems; ++j) {
        const u32 i = c0 + j * kSegThreads + tid;
        if (i >= m) continue;
        // survivors before position i: the output row of key k
        const u32 before = static_cast<u32>(base) + sh.off[j][wave] +
                           popc64(ballot((keep >> j) & 1u) & lanemask_lt());
        if (prev[j] != k[j]) {
          for (u32 b = (prev[j] >> shift) + 1; b <= (k[j] >> shift); ++b)
            dir[b] = before;
        }
        if ((keep >> j) & 1u) {
          out[before] = T::make((k[j] & T::kLow) | (tomb[j] ? T::kTomb : 0u),
                                tomb[j] ? 0u : val[j]);
          low += k[j] < T::kTomb;
          tombs += tomb[j];
        }
      }
      running = sh.misc[3];
      __syncthreads();
    }
    low = wave_sum_u32(low);
    tombs = wave_sum_u32(tombs);
    if (lane == 0) {
      if (low) atomicAdd(&sh.misc[1], low);
      if (tombs) atomicAdd(&sh.misc[2], tombs);
    }
    __syncthreads();

    // the segment's key range [lo_key, hi_key) and its directory bounds:
    // bounds below its first key and above its last key (all of them when
    // it is empty) take the rows before / after the segment
    const u64 unit = u64(1) << s0;
    // every key of the batch shares the bits at and above s0 + 8
    const u64 common =
        s0 + 8 >= 32 ? 0 : (u64(keys_in[0]) >> (s0 + 8)) << (s0 + 8);
    const u64 lo_key = common + u64(seg) * unit;
    const u64 hi_key = lo_key + unit;  // exclusive, may be 2^32
    const u64 ub = u64(1) << shift;    // bucket width
    const u64 first_b = (lo_key + ub - 1) >> shift;  // first bound >= lo_key
    const u64 end_b = (hi_key + ub - 1) >> shift;    // first bound >= hi_key
    const u32 seg_end_row = static_cast<u32>(base) + seg_total;
    const u32 first_key = m ? restore(sk[0]) : 0u;
    const u32 last_key = m ? restore(sk[m - 1]) : 0u;
    if (m == 0) {
      for (u64 b = first_b + tid; b < end_b; b += kSegThreads)
        dir[b] = static_cast<u32>(base);
    } else {
      const u64 k0b = u64(first_key) >> shift;  // bounds in [lo_key, first]
      for (u64 b = first_b + tid; b <= k0b; b += kSegThreads)
        dir[b] = static_cast<u32>(base);
      const u64 klb = (u64(last_key) >> shift) + 1;  // bounds after the last
      for (u64 b = klb + tid; b < end_b; b += kSegThreads) dir[b] = seg_end_row;
    }

    if (tid == 0) {
      if (seg == 255) meta->count = seg_end_row;
      const u64 tb = T::kTomb;
      if ((lo_key <= tb && tb < hi_key) || (seg == 0 && tb < lo_key) ||
          (seg == 255 && hi_key <= tb))
        meta->split = static_cast<u32>(base) + sh.misc[1];
      if (sh.misc[2]) meta->tombs = 1;
      if (seg == 0) meta->dir_bits = dir_bits;
      if (m && start == 0) {  // the first keys of the run
        meta->lo_bucket = first_key >> shift;
        dir[first_key >> shift] = 0;
      }
      if (m && end == n) {  // the last keys of the run
        const u32 hb = last_key >> shift;
        meta->hi_bucket = hb;
        dir[hb + 1] = seg_end_row;
      }
    }
  };
  if (in_lds) finish(sh.k[cur], sh.p[cur]);
  else finish(gk, gp);
}

}  // namespace lsm
}  // namespace gdict
