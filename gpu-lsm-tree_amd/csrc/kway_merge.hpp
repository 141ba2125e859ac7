// Merging up to kMaxWay runs in one pass.
//
// The binary-counter run policy merges a cascade of runs whenever a batch
// arrives on top of runs of equal size (runs of 1, 1, 2, 4, ... batches):
// done pairwise, the rows of the newest runs are rewritten once per level
// of the cascade; merged at once, every row is read and written once.  Over
// 64 equal batches that is 224 instead of 384 batch-sized rewrites.
//
// Tiles are key ranges, not rank ranges.  Tile t holds, from every input
// run, the rows with keys in [x_t, x_{t+1}), so every version of a key
// lands in one tile (the newest-wins rule needs no look across tiles) and a
// tile's split points are plain lower bounds.  The partition picks x_t so
// that the rows below it number about t * target (kway_partition_kernel),
// which bounds every tile by the LDS capacity; each tile then merges its
// inputs in LDS with log2(ways) rounds of pairwise merge paths and writes
// the survivors like the two-way merge (chained scan over tiles).
#pragma once

#include "lsm_kernels.hpp"

namespace gdict {
namespace lsm {

constexpr int kMaxWay = 8;

template <class T>
struct KWayIn {
  RunView<T> run[kMaxWay];  // newest first; entries >= ways are unused
  u32 ways;
};

// Merge workgroup shape, rows per tile (LDS slots: 64 short of threads x
// items, which keeps the 32-bit tile's LDS within 40 KB, four workgroups
// per CU), the partition's target rows per tile and the largest directory
// bucket a split may fall into unrefined: a tile holds fewer than
// target + max(bucket, ways) rows.
template <class T>
struct KWayShape {
  static constexpr u32 kThreads = sizeof(typename T::K) == 4 ? 512 : 256;
  static constexpr u32 kItems = 8;
  static constexpr u32 kSlots = kThreads * kItems - 64;
  static constexpr u32 kLightMax = 128;
  static constexpr u32 kTarget = kSlots - kLightMax;
};

// Lower bound of key c << (B - g), c in [0, 2^g]: one directory entry
// when the run's directory has at least g bits, else a search inside the
// directory bucket that holds the key.
template <class T>
__device__ __forceinline__ u32 boundary_lb(const RunView<T> &r,
                                           const RunMeta &m, u32 c, u32 g) {
  using K = typename T::K;
  if (m.count == 0) return 0;
  if (m.dir_bits >= g) {
    const u64 j = u64(c) << (m.dir_bits - g);
    if (j <= m.lo_bucket) return 0;       // dir[lo_bucket] = 0
    if (j > m.hi_bucket) return m.count;  // dir[hi_bucket + 1] = count
    return r.dir[j];
  }
  if ((u64(c) >> g) != 0) return m.count;  // c = 2^g: past every key
  return run_lower_bound<T>(r, m, K(c) << (T::kBits - g));
}

// Bounds on the rows below key c << (B - g), c < 2^g, summed over the
// runs, from directory entries only (every load issued before any use):
// exact for a run whose directory resolves the key; for a coarser run whose
// bucket holds the key strictly inside, its rows below the bucket (lo) and
// below the bucket's end (hi).
template <class T>
__device__ __forceinline__ void bounds_at(const KWayIn<T> &in,
                                          const RunMeta (&m)[kMaxWay],
                                          u32 ways, u32 c, u32 g, u64 &lo,
                                          u64 &hi) {
  u32 v[kMaxWay], e[kMaxWay];
#pragma unroll
  for (int i = 0; i < kMaxWay; ++i) {
    v[i] = e[i] = 0;
    if (static_cast<u32>(i) < ways && m[i].count) {
      const u32 db = m[i].dir_bits;
      const u64 j = db >= g ? u64(c) << (db - g) : u64(c) >> (g - db);
      const bool inner = db < g && (c & ((1u << (g - db)) - 1u)) != 0;
      if (j < m[i].lo_bucket || (!inner && j == m[i].lo_bucket)) {
        v[i] = e[i] = 0;
      } else if (j > m[i].hi_bucket) {
        v[i] = e[i] = m[i].count;
      } else {
        v[i] = in.run[i].dir[j];
        e[i] = inner ? in.run[i].dir[j + 1] : v[i];
      }
    }
  }
  lo = hi = 0;
#pragma unroll
  for (int i = 0; i < kMaxWay; ++i) {
    lo += v[i];
    hi += e[i];
  }
}

// Tile boundaries by buckets, one thread per key bucket c at the coarsest
// directory resolution g of the inputs (every run's directory resolves it
// with one entry): with P(c) the rows below the bucket's start summed over
// the runs, the bucket holds boundary t when P(c) <= t * target <
// P(c + 1), and records c for it.  Every boundary 0 < t < tiles with
// t * target below the row total lands in exactly one bucket.  Reads each
// needed directory entry once instead of searching for every boundary.
template <class T>
__global__ void kway_bucket_kernel(KWayIn<T> in, u32 tiles, u32 target,
                                   u32 *__restrict__ bucket_of) {
  const u32 ways = in.ways;
  RunMeta m[kMaxWay];
  u32 g = kMaxDirBits;
#pragma unroll
  for (int i = 0; i < kMaxWay; ++i) {
    m[i] = RunMeta{0, 0, kMaxDirBits, 0, 1, 0};
    if (static_cast<u32>(i) < ways) {
      m[i] = *in.run[i].meta;
      g = m[i].dir_bits < g ? m[i].dir_bits : g;
    }
  }
  const u64 c = u64(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >> g) return;
  u32 v0[kMaxWay], v1[kMaxWay];
#pragma unroll
  for (int i = 0; i < kMaxWay; ++i) {
    v0[i] = v1[i] = 0;
    if (static_cast<u32>(i) < ways) {
      v0[i] = boundary_lb(in.run[i], m[i], static_cast<u32>(c), g);
      v1[i] = boundary_lb(in.run[i], m[i], static_cast<u32>(c) + 1, g);
    }
  }
  u64 p0 = 0, p1 = 0;
#pragma unroll
  for (int i = 0; i < kMaxWay; ++i) {
    p0 += v0[i];
    p1 += v1[i];
  }
  u64 t = p0 == 0 ? 1 : (p0 + target - 1) / target;
  for (; t < tiles && t * target < p1; ++t)
    bucket_of[t] = static_cast<u32>(c);
}

// Split points of the k-way merge, one wave per tile boundary t in
// [0, tiles]: parts[t * kMaxWay + i] = rows of run i with keys below x_t.
// Starting from the boundary's coarse bucket (kway_bucket_kernel), a 64-ary
// search over its sub-buckets at the finest directory resolution of the
// inputs finds the bucket c whose start has at most R = t * target rows
// below it (summed over the runs) and whose end has more.  When that bucket
// holds at most light_max rows, x_t is its start.  Otherwise (dense keys)
// phase 2 searches the keys inside the bucket for the largest x with at
// most R rows below it.  Also clears the merge's chained-scan state and the
// output run's header.
template <class T>
__global__ void kway_partition_kernel(KWayIn<T> in, u32 tiles, u32 target,
                                      u32 light_max,
                                      const u32 *__restrict__ bucket_of,
                                      u32 *__restrict__ parts, u64 *status,
                                      u32 *ticket, RunMeta *out_meta,
                                      u32 dir_bits) {
  using K = typename T::K;
  const u32 lane = lane_id();
  const u64 t = (u64(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  if (t > tiles) return;
  // the boundary's coarse bucket, loaded beside the run headers (unused,
  // and possibly never written, for t = 0 and t * target >= rows)
  const u32 c0 = bucket_of[t];
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
  const u32 ways = in.ways;
  RunMeta m[kMaxWay];
  u64 total = 0;
  u32 g_lo = kMaxDirBits, g = 1;
#pragma unroll
  for (int i = 0; i < kMaxWay; ++i) {
    m[i] = RunMeta{0, 0, kMaxDirBits, 0, 1, 0};
    if (static_cast<u32>(i) < ways) {
      m[i] = *in.run[i].meta;
      total += m[i].count;
      g_lo = m[i].dir_bits < g_lo ? m[i].dir_bits : g_lo;
      g = m[i].dir_bits > g ? m[i].dir_bits : g;
    }
  }
  u32 *out = parts + t * kMaxWay;
  const u64 R = t * u64(target);
  if (t == 0 || R >= total) {  // below every row / above every row
#pragma unroll
    for (int i = 0; i < kMaxWay; ++i)
      if (lane == static_cast<u32>(i) && static_cast<u32>(i) < ways)
        out[i] = t == 0 ? 0u : m[i].count;
    return;
  }
  // this lane's run (lanes >= ways answer nothing)
  RunView<T> rv = in.run[0];
  RunMeta rm = m[0];
#pragma unroll
  for (int i = 1; i < kMaxWay; ++i)
    if (lane == static_cast<u32>(i)) {
      rv = in.run[i];
      rm = m[i];
    }
  // phase 1: the boundary's coarse bucket c0 (exact), then the last of its
  // sub-buckets at the finest resolution whose start has at most R rows
  // below it by the upper bound; a start becomes x_t when its lower bound
  // leaves fewer than light_max rows short of R, which bounds the tile,
  // and every run's split is then its lower bound of x_t.  The coarse
  // start is preferred when it qualifies.
  u64 below_lo = 0;  // rows below the coarse start, summed over the runs
  u32 a_mine = 0, b_mine = 0;  // this lane's run's rows below c0, c0 + 1
  K x = K(c0) << (T::kBits - g_lo);
  bool light = false;
  if (g > g_lo) {
    // 64-ary search for the last c in [lo, hi) with hi_bound(c) <= R,
    // given that of lo <= R < that of hi.  Lane 0 also evaluates lo in
    // the first round (exact there: every run's directory resolves the
    // coarse start), so below_lo comes from the same loads, and each
    // round keeps the chosen start's lower bound, so none is reloaded.
    const u32 d = g - g_lo;
    u32 lo = c0 << d, hi = (c0 + 1) << d;
    u64 plo_lo = 0;
    bool first = true;
    while (hi - lo > 1) {
      const u32 span = hi - lo;
      const bool fine = span <= static_cast<u32>(kWave);
      const u32 c =
          fine ? lo + lane : lo + static_cast<u32>((u64(span) * lane) >> 6);
      bool p = lane == 0;  // hi_bound(lo) <= R is known
      u64 plo = 0, phi = 0;
      if ((lane == 0 && first) || (lane != 0 && c < hi)) {
        bounds_at<T>(in, m, ways, c, g, plo, phi);
        if (lane != 0) p = phi <= R;
      }
      if (first) {
        below_lo = static_cast<u64>(
            __shfl(static_cast<unsigned long long>(plo), 0, kWave));
        plo_lo = below_lo;
        first = false;
      }
      const u32 L = popc64(ballot(p)) - 1;  // the true lanes form a prefix
      const u64 plo_l = static_cast<u64>(__shfl(
          static_cast<unsigned long long>(plo), static_cast<int>(L), kWave));
      if (L != 0) plo_lo = plo_l;
      const u32 cl =
          fine ? lo + L : lo + static_cast<u32>((u64(span) * L) >> 6);
      const u32 cn =
          fine ? lo + L + 1
               : (L + 1 < static_cast<u32>(kWave)
                      ? lo + static_cast<u32>((u64(span) * (L + 1)) >> 6)
                      : hi);
      lo = cl;
      hi = cn;
    }
    light = R - below_lo < light_max;
    if (!light && R - plo_lo < light_max) {
      light = true;
      x = K(lo) << (T::kBits - g);
    }
  } else {
    // every run's directory resolves g: the coarse bounds are entries
#pragma unroll
    for (int i = 0; i < kMaxWay; ++i) {
      if (static_cast<u32>(i) < ways) {
        const u32 ai = boundary_lb(in.run[i], m[i], c0, g_lo);
        const u32 bi = boundary_lb(in.run[i], m[i], c0 + 1, g_lo);
        below_lo += ai;
        if (lane == static_cast<u32>(i)) {
          a_mine = ai;
          b_mine = bi;
        }
      }
    }
    light = R - below_lo < light_max;
  }
  u32 mine = 0;
  if (light) {
    if (lane < ways) mine = run_lower_bound<T>(rv, rm, x);
  } else {
    // phase 2 (dense keys): the largest x in the coarse bucket with at
    // most R rows below it, by a 64-ary search over its keys; inside one
    // bucket every key shares the top bit, so rows compare on the low bits
    if (g > g_lo && lane < ways) {
      a_mine = boundary_lb(rv, rm, c0, g_lo);
      b_mine = boundary_lb(rv, rm, c0 + 1, g_lo);
    }
    u32 a[kMaxWay], b[kMaxWay];
#pragma unroll
    for (int i = 0; i < kMaxWay; ++i) {
      a[i] = __shfl(a_mine, i, kWave);
      b[i] = __shfl(b_mine, i, kWave);
    }
    const u64 need = R - below_lo;
    u64 span = u64(1) << (T::kBits - g_lo);  // 1 <= g_lo <= 24
    auto rank_in = [&](K key) -> u64 {
      u64 r = 0;
#pragma unroll
      for (int i = 0; i < kMaxWay; ++i)
        if (static_cast<u32>(i) < ways)
          r += window_lower_bound<T>(in.run[i].rows, a[i], b[i],
                                     key & T::kLow) -
               a[i];
      return r;
    };
    while (span > 1) {
      const bool fine = span <= static_cast<u64>(kWave);
      const u64 step = fine ? 1 : span >> 6;
      const u64 off = u64(lane) * step;
      bool p = lane == 0;
      if (lane != 0 && off < span) p = rank_in(x + K(off)) <= need;
      const u32 L = popc64(ballot(p)) - 1;
      x += K(u64(L) * step);
      span = fine ? 1 : (L + 1 < static_cast<u32>(kWave) ? step
                                                         : span - u64(L) * step);
    }
    if (lane < ways)
      mine = window_lower_bound<T>(rv.rows, a_mine, b_mine, x & T::kLow);
  }
  if (lane < ways) out[lane] = mine;
}

// Merges the tile's rows of every input run (newest first) into `out`: on
// equal keys the newest row survives, tombstones drop when
// `drop_tombstones`; header and directory as in merge_kernel.  Rows are
// staged in LDS in run order; round r merges adjacent groups of 2^(r-1)
// runs (newer group first on equal keys) with one merge path per thread,
// carrying (key, slot) pairs, so after ceil(log2(ways)) rounds the tile is
// in (key, age) order and each row's value is read once from its slot.  A
// slot holds the row's LDS position and, in bit 15, its tombstone flag.
template <class T, int kItems, int kThreads>
__global__ __launch_bounds__(kThreads) void kway_merge_kernel(
    KWayIn<T> in, const u32 *__restrict__ parts,
    typename T::Row *__restrict__ out, u64 out_rows, RunMeta *out_meta,
    u32 *__restrict__ dir, u32 dir_bits, int drop_tombstones, u64 *status,
    u32 *ticket) {
  using K = typename T::K;
  using V = typename T::V;
  using Row = typename T::Row;
  constexpr u32 kTile = KWayShape<T>::kSlots;  // LDS slots
  static_assert(kThreads * kItems >= kTile, "every slot has a thread");
  constexpr u32 kWaves = kThreads / kWave;
  constexpr u32 kStageHalf = kTile / 2;  // rows over the key array
  constexpr u32 kTombSlot = 0x8000u;
  static_assert(sizeof(Row) == 2 * sizeof(K) && sizeof(K) == sizeof(V),
                "row = key word + value");
  static_assert(kTile <= kTombSlot, "slots are 15-bit");
  // LDS slots are swizzled (lds_swz): a thread's merge window and its
  // staged rows start kItems slots after the previous lane's
  constexpr u32 kKeySlots0 = lds_lines<sizeof(K)>(kTile);
  constexpr u32 kStageSlots =
      lds_lines<sizeof(Row)>(kStageHalf) * sizeof(Row) / sizeof(K);
  constexpr u32 kKeySlots = kKeySlots0 > kStageSlots ? kKeySlots0
                                                     : kStageSlots;
  auto key_slot = [](u32 i) { return lds_swz<sizeof(K)>(i); };
  auto slot_slot = [](u32 i) { return lds_swz<2>(i); };
  auto row_slot = [](u32 i) { return lds_swz<sizeof(Row)>(i); };
  __shared__ u32 s_tile;
  __shared__ __align__(16) K s_key[kKeySlots];
  __shared__ __align__(16) V s_val[kKeySlots];
  __shared__ unsigned short s_slot[lds_lines<2>(kTile)];
  __shared__ u32 s_scan[kWaves + 1];
  __shared__ K s_wave_last[kWaves];
  __shared__ u64 s_base;
  __shared__ u32 s_low, s_tombs, s_tile_high;
  __shared__ K s_prev;  // largest key before the tile
  __shared__ u32 s_has_prev, s_before_low;
  // per segment: source rows, row of slot s = s + delta, top-bit split
  __shared__ const Row *s_seg_rows[kMaxWay];
  __shared__ u32 s_seg_delta[kMaxWay], s_seg_split[kMaxWay];

  if (threadIdx.x == 0) {
    s_tile = atomicAdd(ticket, 1u);
    s_low = 0;
    s_tombs = 0;
    s_tile_high = 0;
  }
  __syncthreads();
  const u32 tile = __builtin_amdgcn_readfirstlane(s_tile);
  const u32 ways = in.ways;
  const u32 lane = lane_id();
  const u32 wave = threadIdx.x / kWave;
  RunMeta m[kMaxWay];
  u32 beg[kMaxWay], off[kMaxWay + 1];
  u64 total = 0, ends = 0;
  bool trailing = true;  // every row lies before this tile
  off[0] = 0;
#pragma unroll
  for (int i = 0; i < kMaxWay; ++i) {
    m[i] = RunMeta{0, 0, 0, 0, 1, 0};
    beg[i] = 0;
    u32 len = 0;
    if (static_cast<u32>(i) < ways) {
      m[i] = *in.run[i].meta;
      // split points are nondecreasing lower bounds; the clamps only keep
      // every access inside the run and the tile whatever they hold
      const u32 cnt = m[i].count;
      const u32 b0 = parts[u64(tile) * kMaxWay + i];
      const u32 e0 = parts[(u64(tile) + 1) * kMaxWay + i];
      beg[i] = b0 < cnt ? b0 : cnt;
      const u32 end = e0 < beg[i] ? beg[i] : (e0 < cnt ? e0 : cnt);
      len = end - beg[i];
      if (len > kTile - off[i]) len = kTile - off[i];
      total += cnt;
      ends += end;
      trailing = trailing && beg[i] == cnt;
    }
    off[i + 1] = off[i] + len;
  }
  if (tile > 0 && trailing) {
    // nothing here or after; an empty aggregate keeps every look-back
    // finite whatever the split points hold
    if (threadIdx.x == 0) store_relaxed(&status[tile], kFlagAggregate);
    return;
  }
  const u32 n = off[kMaxWay];
  const bool ends_all = ends == total;
#pragma unroll
  for (int i = 0; i < kMaxWay; ++i) {
    if (threadIdx.x == static_cast<u32>(i)) {
      s_seg_rows[i] = in.run[i].rows;
      s_seg_delta[i] = beg[i] - off[i];
      s_seg_split[i] = m[i].split;
    }
  }
  __syncthreads();

  // Load: slot s of [0, n) holds row beg[i] + (s - off[i]) of run i, where
  // segment i is the last one starting at or before s; every load of a
  // thread is issued before the first one is used.
  Row rr[kItems];
  u32 segs = 0;  // 3 bits per item
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const u32 s = threadIdx.x + r * kThreads;
    if (s < n) {
      u32 i = 0;
#pragma unroll
      for (int j = 1; j < kMaxWay; ++j) i += s >= off[j];
      segs |= i << (3 * r);
      rr[r] = s_seg_rows[i][s + s_seg_delta[i]];
    }
  }
  // the rows just before the tile (the largest key below it and whether
  // every row below it has the top key bit clear)
  if (wave == 0) {
    K pk = 0;
    bool has = false;
#pragma unroll
    for (int i = 0; i < kMaxWay; ++i) {
      if (lane == static_cast<u32>(i) && static_cast<u32>(i) < ways &&
          beg[i] > 0) {
        pk = decode_key<T>(T::kw(in.run[i].rows[beg[i] - 1]), beg[i] - 1,
                           m[i].split);
        has = true;
      }
    }
    const u64 any = ballot(has);
    const u64 high = ballot(has && pk >= T::kTomb);
    K best = 0;
#pragma unroll
    for (int i = 0; i < kMaxWay; ++i) {
      const K v = __shfl(pk, i, kWave);
      if (((any >> i) & 1ull) && v > best) best = v;
    }
    if (lane == 0) {
      s_prev = best;
      s_has_prev = any != 0;
      s_before_low = high == 0;
    }
  }
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const u32 s = threadIdx.x + r * kThreads;
    if (s < n) {
      const u32 i = (segs >> (3 * r)) & 7u;
      const u32 split = s_seg_split[i], pos = s + s_seg_delta[i];
      const K w = T::kw(rr[r]);
      s_key[key_slot(s)] = decode_key<T>(w, pos, split);
      s_val[key_slot(s)] = T::val(rr[r]);
      s_slot[slot_slot(s)] = static_cast<unsigned short>(
          s | ((w >> (T::kBits - 1)) ? kTombSlot : 0u));
    }
  }
  __syncthreads();

  // Merge rounds.  Thread t produces merged positions [t * kItems, +kItems)
  // of the round; a group is [off[u * w], off[(u + 1) * w]) with w = 2^r
  // runs, its newer half ending at off[u * w + w / 2].  The next key of
  // each side rides in a register: a step reads one key and one slot.
  const u32 p0 = threadIdx.x * kItems;
  const u32 levels = ways <= 2 ? 1u : (ways <= 4 ? 2u : 3u);
  K ok[kItems];
  u32 os[kItems];
#pragma unroll
  for (int s = 0; s < kItems; ++s) {
    ok[s] = 0;
    os[s] = 0;
  }
#pragma unroll
  for (int r = 1; r <= 3; ++r) {
    if (static_cast<u32>(r) > levels) break;
    const int w = 1 << r;
    // group of position q: bounds [glo, gmid) newer, [gmid, ghi) older
    auto group_of = [&](u32 q, u32 &glo, u32 &gmid, u32 &ghi) {
      glo = gmid = ghi = 0;
#pragma unroll
      for (int u = 0; u < kMaxWay; u += 2) {
        if (u % w == 0 && q >= off[u] && q < off[u + w]) {
          glo = off[u];
          gmid = off[u + w / 2];
          ghi = off[u + w];
        }
      }
    };
    if (p0 < n) {
      u32 glo, gmid, ghi;
      group_of(p0, glo, gmid, ghi);
      const u32 na = gmid - glo, nb = ghi - gmid;
      const u32 diag = p0 - glo;
      u32 lo = diag > nb ? diag - nb : 0;
      u32 hi = diag < na ? diag : na;
      while (lo < hi) {
        const u32 mid = (lo + hi) >> 1;
        if (s_key[key_slot(glo + mid)] <=
            s_key[key_slot(gmid + (diag - mid - 1))])
          lo = mid + 1;
        else hi = mid;
      }
      u32 ia = glo + lo, ib = gmid + (diag - lo);  // absolute positions
      K ka = ia < gmid ? s_key[key_slot(ia)] : K(0);
      K kb = ib < ghi ? s_key[key_slot(ib)] : K(0);
#pragma unroll
      for (int s = 0; s < kItems; ++s) {
        const u32 q = p0 + s;
        if (q < n) {
          if (q == ghi) {  // the next group starts here (diagonal 0)
            group_of(q, glo, gmid, ghi);
            ia = glo;
            ib = gmid;
            ka = ia < gmid ? s_key[key_slot(ia)] : K(0);
            kb = ib < ghi ? s_key[key_slot(ib)] : K(0);
          }
          const bool a_left = ia < gmid;
          const bool b_left = ib < ghi;
          const bool take_a = a_left && (!b_left || ka <= kb);
          const u32 from = take_a ? ia : ib;
          ok[s] = take_a ? ka : kb;
          os[s] = s_slot[slot_slot(from)];
          const u32 nxt = from + 1;
          const bool more = nxt < (take_a ? gmid : ghi);
          const K kn = more ? s_key[key_slot(nxt)] : K(0);
          if (take_a) {
            ia = nxt;
            ka = kn;
          } else {
            ib = nxt;
            kb = kn;
          }
        }
      }
    }
    __syncthreads();  // every read of this round precedes the writes
    if (static_cast<u32>(r) < levels) {
#pragma unroll
      for (int s = 0; s < kItems; ++s)
        if (p0 + s < n) {
          s_key[key_slot(p0 + s)] = ok[s];
          s_slot[slot_slot(p0 + s)] = static_cast<unsigned short>(os[s]);
        }
      __syncthreads();
    }
  }

  // ok / os: this thread's merged positions.  The key before a thread's
  // first position is the last key of the thread before (every thread
  // before the last busy one has kItems positions), or for thread 0 the
  // largest key below the tile.
  const u32 steps = p0 < n ? (n - p0 < kItems ? n - p0 : kItems) : 0u;
  const K before_me = __shfl_up(ok[kItems - 1], 1, kWave);
  if (lane == kWave - 1) s_wave_last[wave] = ok[kItems - 1];
  if (steps && p0 + steps == n) s_tile_high = ok[steps - 1] >= T::kTomb;
  __syncthreads();
  bool has_prev = false;
  K prev = 0;
  if (threadIdx.x == 0) {
    has_prev = s_has_prev;
    prev = s_prev;
  } else if (steps) {
    has_prev = true;
    prev = lane ? before_me : s_wave_last[wave - 1];
  }
  const bool tile_high = s_tile_high != 0;
  V ov[kItems];
  bool okeep[kItems];
  u32 kept = 0;
#pragma unroll
  for (int s = 0; s < kItems; ++s) {
    ov[s] = 0;
    okeep[s] = false;
    if (static_cast<u32>(s) < steps) {
      ov[s] = s_val[key_slot(os[s] & (kTombSlot - 1))];
      // every version of a key is in this tile: the first (newest) stays
      const bool dup = s > 0 ? ok[s - 1] == ok[s]
                             : (threadIdx.x > 0 && prev == ok[s]);
      okeep[s] = !dup && !((os[s] & kTombSlot) && drop_tombstones);
      kept += okeep[s];
    }
  }
  const u32 local = block_exclusive_scan<kThreads>(kept, s_scan);
  const u32 tile_total = s_scan[kWaves];
  // Wave 0 waits on the look-back while the other waves stage their
  // survivors at tile-local positions (the scan's barriers ordered every
  // LDS input read before these writes).
  if (wave == 0) {
    const u64 ex = wave_lookback(status, tile, tile_total);
    if (threadIdx.x == 0) s_base = ex;
  }
  Row *stage_lo = reinterpret_cast<Row *>(s_key);
  Row *stage_hi = reinterpret_cast<Row *>(s_val);
  u32 low = 0, tombs = 0, at = local;
#pragma unroll
  for (int s = 0; s < kItems; ++s) {
    if (okeep[s]) {
      const bool tomb = (os[s] & kTombSlot) != 0;
      const Row row =
          T::make((ok[s] & T::kLow) | (tomb ? T::kTomb : K(0)), ov[s]);
      if (at < kStageHalf) stage_lo[row_slot(at)] = row;
      else stage_hi[row_slot(at - kStageHalf)] = row;
      ++at;
      low += ok[s] < T::kTomb;
      tombs += tomb;
    }
  }
  __syncthreads();
  const u64 obase = s_base;
  for (u32 t = threadIdx.x; t < tile_total; t += kThreads)
    if (obase + t < out_rows)
      out[obase + t] = t < kStageHalf ? stage_lo[row_slot(t)]
                                      : stage_hi[row_slot(t - kStageHalf)];
  // directory entries inside the bucket window (first to last input key):
  // every input element owns the buckets in (bucket(previous input key),
  // bucket(its key)] and writes them the survivors before it
  {
    const u32 shift = T::kBits - dir_bits;
    u32 pbk = static_cast<u32>(prev >> shift);
    u32 before = static_cast<u32>(obase) + local;
#pragma unroll
    for (int s = 0; s < kItems; ++s) {
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
    if (ends_all && steps && p0 + steps == n) {
      // the last input element closes the window at the row count
      out_meta->hi_bucket = pbk;
      dir[pbk + 1] = static_cast<u32>(obase + tile_total);
    }
  }
  low = wave_sum_u32(low);
  tombs = wave_sum_u32(tombs);
  if (lane == 0) {
    if (low) atomicAdd(&s_low, low);
    if (tombs) atomicAdd(&s_tombs, tombs);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // the first tile whose rows reach the top key bit (or the last one)
    // writes the split, the one holding the last row the count
    if (ends_all) out_meta->count = static_cast<u32>(obase + tile_total);
    if (s_before_low && (tile_high || ends_all))
      out_meta->split = static_cast<u32>(obase + s_low);
    if (s_tombs) out_meta->tombs = 1;
  }
}

}  // namespace lsm
}  // namespace gdict
