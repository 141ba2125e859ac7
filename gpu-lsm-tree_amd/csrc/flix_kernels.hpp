// Device side of the FliX coarse-granular updatable sorted index.
//
// Data model in HBM (one pool of fixed-size nodes):
//   node_keys[node * S + slot], node_vals[node * S + slot]   key/value slots
//   node_count[node]  live slots (packed at the front, ascending keys)
//   node_next[node]   1-based index of the next node of the chain, 0 = end
//   maxv[b]           upper bound of bucket b's key range; bucket b holds the
//                     keys in (maxv[b-1], maxv[b]], the last bucket ends at
//                     the largest key (FliX/impl_cg_rtx_index_updates.cuh:
//                     208-240, find_maxvalues_kernel)
//   bsize[b], blen[b] live keys and nodes of bucket b's chain, kept by every
//                     update so an insert can size its new nodes up front
//   bdir[j]           bucket directory: the first bucket whose max is >= the
//                     j-th of 2^d equal key slices (j << (key bits - d)),
//                     2^d about the bucket count, so finding a key's
//                     bucket is a short search in [bdir[j], bdir[j + 1]]
// Nodes [0, P) are the bucket heads; nodes [P, total) form the allocation
// region, handed out through a ring of free node ids (bump order at build,
// released nodes of deletes appended).  Inside a chain every key of node i
// is smaller than every key of node i+1; only a head may be empty.  A batch
// key belongs to the chain node that precedes the first node whose first
// key exceeds it, which needs no per-node maximum.
//
// S (slots per node) is at most 64: one wave64 lane per slot, so a node's
// keys arrive in one coalesced load and ballots give match/compaction masks.
#pragma once

#include "device_utils.hpp"
#include "radix_sort.hpp"

namespace gdict {
namespace flix {

constexpr int kMaxSlots = 64;

template <class K, class V>
struct Pool {
  K *keys;
  V *vals;
  u32 *count;
  u32 *next;
  const K *maxv;
  u32 *bsize;
  u32 *blen;
  u32 S;
  u32 buckets;
  u64 total_nodes;
  const u32 *bdir;  // 2^d + 1 entries
  u32 bdir_shift;   // key bits - d
};

// Ring of free node ids.  alloc_cursor / free_count grow monotonically;
// live ids sit at [alloc_cursor, free_count) modulo ring_cap.
struct Ring {
  u32 *ids;
  u64 cap;
  unsigned long long *alloc_cursor;
  unsigned long long *free_count;
};

// First bucket whose max >= key: the directory slice of the key bounds it
// (the last bucket's max is all-ones, so every key has one).
template <class K, class V>
__device__ __forceinline__ u32 bucket_of(const Pool<K, V> &p, K key) {
  const u64 j = static_cast<u64>(key) >> p.bdir_shift;
  u32 lo = p.bdir[j], hi = p.bdir[j + 1];
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    if (p.maxv[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// bdir[j] = first bucket whose max >= j << shift (j < 2^d), bdir[2^d] = the
// last bucket.  Bucket b is that bucket for the slices j with
// maxv[b-1] >> shift < j <= maxv[b] >> shift; bucket_dir_kernel writes them
// for a thread per bucket when they are at most kDirSpan, and
// bucket_dir_fill_kernel finds the rest (slices inside wide gaps between
// keys, left at ~0 by the clear) by a lower bound over maxv.
constexpr u32 kDirSpan = 64;

static __global__ void bucket_dir_clear_kernel(u32 *bdir, u64 n) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    bdir[i] = ~0u;
}

template <class K>
static __global__ void bucket_dir_kernel(const K *maxv, u32 buckets, u32 *bdir,
                                         u32 d, u32 shift) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 bb = u64(blockIdx.x) * blockDim.x + threadIdx.x; bb < buckets;
       bb += stride) {
    const u32 b = static_cast<u32>(bb);
    const u64 j0 = b == 0 ? 0 : (static_cast<u64>(maxv[b - 1]) >> shift) + 1;
    const u64 j1 = static_cast<u64>(maxv[b]) >> shift;  // < 2^d
    if (j0 <= j1 && j1 - j0 < kDirSpan)
      for (u64 j = j0; j <= j1; ++j) bdir[j] = b;
    if (b + 1 == buckets) bdir[u64(1) << d] = b;
  }
}

template <class K>
static __global__ void bucket_dir_fill_kernel(const K *maxv, u32 buckets,
                                              u32 *bdir, u32 d, u32 shift) {
  const u64 slices = u64(1) << d;
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 j = u64(blockIdx.x) * blockDim.x + threadIdx.x; j < slices;
       j += stride) {
    if (bdir[j] != ~0u) continue;
    const K v = static_cast<K>(j << shift);
    u32 lo = 0, hi = buckets - 1;
    while (lo < hi) {
      const u32 mid = (lo + hi) >> 1;
      if (maxv[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    bdir[j] = lo;
  }
}

// lower_bound over a sorted global array segment [lo, hi).
template <class K>
__device__ __forceinline__ u64 lower_bound_g(const K *a, u64 lo, u64 hi,
                                             K key) {
  while (lo < hi) {
    const u64 mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// First slot of a node's c ascending keys whose key >= key, in two rounds
// of independent loads instead of log2(c) dependent ones: the last key of
// every 8-slot block (at most 8 of them, S <= 64) picks the block, then the
// block's keys give the slot.
template <class K>
__device__ __forceinline__ u32 node_lower_bound(const K *nk, u32 c, K key) {
  u32 blk = 0;
  if (c > 8) {
    K last[8];
#pragma unroll
    for (u32 j = 0; j < 8; ++j) {
      const u32 at = 8 * j + 7;
      last[j] = at < c ? nk[at] : K(0);
    }
#pragma unroll
    for (u32 j = 0; j < 8; ++j) blk += (8 * j + 7 < c && last[j] < key);
  }
  const u32 base = 8 * blk;
  K k[8];
#pragma unroll
  for (u32 j = 0; j < 8; ++j) k[j] = base + j < c ? nk[base + j] : K(0);
  u32 below = 0;
#pragma unroll
  for (u32 j = 0; j < 8; ++j) below += (base + j < c && k[j] < key);
  return base + below;
}

// node_lower_bound with fewer loads and more rounds: a binary search over
// the last keys of the node's full 8-slot blocks (one load per round, at
// most four rounds), then the chosen block's keys.  Queries in random order
// miss the caches, and a thread per query is bound by its loads more than
// by its rounds there; queries in key order find the node's lines in cache
// and take node_lower_bound's single round of block lasts.
template <class K>
__device__ __forceinline__ u32 node_lower_bound_bisect(const K *nk, u32 c,
                                                       K key) {
  u32 blk = 0, hi = c >> 3;
  while (blk < hi) {
    const u32 mid = (blk + hi) >> 1;
    if (nk[8 * mid + 7] < key) blk = mid + 1;
    else hi = mid;
  }
  const u32 base = 8 * blk;
  K k[8];
#pragma unroll
  for (u32 j = 0; j < 8; ++j) k[j] = base + j < c ? nk[base + j] : K(0);
  u32 below = 0;
#pragma unroll
  for (u32 j = 0; j < 8; ++j) below += (base + j < c && k[j] < key);
  return base + below;
}

// Slots of a node's c ascending keys that hold a key <= key (the first slot
// above it), by the same two rounds as node_lower_bound.
template <class K>
__device__ __forceinline__ u32 node_upper_bound(const K *nk, u32 c, K key) {
  u32 blk = 0;
  if (c > 8) {
    K last[8];
#pragma unroll
    for (u32 j = 0; j < 8; ++j) {
      const u32 at = 8 * j + 7;
      last[j] = at < c ? nk[at] : K(0);
    }
#pragma unroll
    for (u32 j = 0; j < 8; ++j) blk += (8 * j + 7 < c && last[j] <= key);
  }
  const u32 base = 8 * blk;
  K k[8];
#pragma unroll
  for (u32 j = 0; j < 8; ++j) k[j] = base + j < c ? nk[base + j] : K(0);
  u32 upto = 0;
#pragma unroll
  for (u32 j = 0; j < 8; ++j) upto += (base + j < c && k[j] <= key);
  return base + upto;
}

// Node of bucket b's chain that owns `key`.
template <class K, class V>
__device__ __forceinline__ u32 owner_node(const Pool<K, V> &p, u32 b, K key) {
  u32 node = b;
  while (true) {
    const u32 nxt = p.next[node];
    if (!nxt) return node;
    if (key < p.keys[u64(nxt - 1) * p.S]) return node;
    node = nxt - 1;
  }
}

// ----------------------------------------------------------------- build
// Sorted unique keys -> heads.  Bucket b takes rows [b*fill, (b+1)*fill).
template <class K, class V>
static __global__ void layout_kernel(Pool<K, V> p, K *maxv_out, const K *keys,
                              const V *vals, u64 n, u32 fill) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  const u64 slots = u64(p.buckets) * p.S;
  for (u64 t = u64(blockIdx.x) * blockDim.x + threadIdx.x; t < slots;
       t += stride) {
    const u32 b = static_cast<u32>(t / p.S);
    const u32 s = static_cast<u32>(t % p.S);
    const u64 row = u64(b) * fill + s;
    const u64 end = u64(b + 1) * fill < n ? u64(b + 1) * fill : n;
    if (s < fill && row < end) {
      p.keys[t] = keys[row];
      p.vals[t] = vals[row];
    }
    if (s == 0) {
      const u64 first = u64(b) * fill;
      const u32 c = first < n ? static_cast<u32>(end - first) : 0u;
      p.count[b] = c;
      p.next[b] = 0;
      p.bsize[b] = c;
      p.blen[b] = 1;
      maxv_out[b] = (b + 1 == p.buckets) ? ~K(0) : keys[end - 1];
    }
  }
}

// Rebuild without a flattened copy: a wave per bucket chain of the old
// layer p0 writes the chain's keys and values straight to their rows of the
// new layer p (offsets[b] = the live keys of the chains before chain b, so
// row r of the sorted contents goes to bucket r / fill, slot r % fill);
// layout_heads_kernel then writes the new heads.  S <= 64.
template <class K, class V>
static __global__ __launch_bounds__(256) void relayout_kernel(
    Pool<K, V> p0, const u64 *offsets, Pool<K, V> p, u32 fill) {
  const u32 lane = lane_id();
  const u32 b = blockIdx.x * 4 + threadIdx.x / kWave;
  if (b >= p0.buckets) return;
  u64 at = offsets[b];
  u32 node = b;
  while (true) {
    const u32 c = p0.count[node];
    const u32 nxt = p0.next[node];
    if (lane < c) {
      // row at + lane: the node's first row by one division, the lane's
      // step from it (fewer than fill + 64 slots) by a 32-bit one
      const u64 q0 = at / fill;
      const u32 t = static_cast<u32>(at - q0 * fill) + lane;
      const u64 nb = q0 + t / fill;
      const u64 to = nb * p.S + t % fill;
      p.keys[to] = p0.keys[u64(node) * p0.S + lane];
      p.vals[to] = p0.vals[u64(node) * p0.S + lane];
    }
    at += c;
    if (!nxt) break;
    node = nxt - 1;
  }
}

// Heads of a layer whose n rows relayout_kernel placed: counts, empty
// links, chain sizes and bucket bounds (a bucket's last key, all-ones for
// the last bucket), as layout_kernel writes them.
template <class K, class V>
static __global__ void layout_heads_kernel(Pool<K, V> p, K *maxv_out, u64 n,
                                           u32 fill) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 bb = u64(blockIdx.x) * blockDim.x + threadIdx.x; bb < p.buckets;
       bb += stride) {
    const u32 b = static_cast<u32>(bb);
    const u64 first = u64(b) * fill;
    const u64 end = first + fill < n ? first + fill : n;
    const u32 c = first < n ? static_cast<u32>(end - first) : 0u;
    p.count[b] = c;
    p.next[b] = 0;
    p.bsize[b] = c;
    p.blen[b] = 1;
    maxv_out[b] = (b + 1 == p.buckets) ? ~K(0)
                  : c                  ? p.keys[u64(b) * p.S + c - 1]
                                       : K(0);
  }
}

static __global__ void ring_init_kernel(u32 *ids, u64 first_node, u64 n) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    ids[i] = static_cast<u32>(first_node + i);
}

// Batch slice [begin[b], begin[b+1]) of every bucket for a sorted batch.
template <class K>
static __global__ void bucket_slices_kernel(const K *maxv, u32 buckets,
                                     const K *batch, u64 n_arg,
                                     const u64 *n_dev, u64 *begin) {
  const u32 b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > buckets) return;
  // the batch length: n_arg, or the count a previous kernel left at n_dev
  const u64 n = n_dev ? *n_dev : n_arg;
  if (b == 0) {
    begin[0] = 0;
    return;
  }
  if (b == buckets) {
    begin[buckets] = n;
    return;
  }
  // first batch key > maxv[b-1]
  const K m = maxv[b - 1];
  u64 lo = 0, hi = n;
  while (lo < hi) {
    const u64 mid = (lo + hi) >> 1;
    if (batch[mid] <= m) lo = mid + 1;
    else hi = mid;
  }
  begin[b] = lo;
}

// Marks batch keys already stored (insert-if-absent) and duplicates inside
// the batch (the first occurrence of the sorted, stable batch wins).
template <class K, class V>
static __global__ void mark_new_kernel(Pool<K, V> p, const K *batch, u64 n,
                                u32 *keep) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const K key = batch[i];
    bool fresh = i == 0 || batch[i - 1] != key;
    if (fresh) {
      const u32 b = bucket_of(p, key);
      const u32 node = owner_node(p, b, key);
      const u32 c = p.count[node];
      const K *nk = p.keys + u64(node) * p.S;
      const u32 lo = node_lower_bound(nk, c, key);
      if (lo < c && nk[lo] == key) fresh = false;
    }
    keep[i] = fresh ? 1u : 0u;
  }
}

// ------------------------------------------------- planned chain inserts
// An insert sizes every touched bucket's chain before it writes anything:
// from the bucket's live keys t0 (bsize) and nodes L (blen) and its slice
// of m new keys, the chain ends with chain_nodes_after(t0 + m, L) nodes.
// One scan over the touched buckets gives each its run of new nodes in the
// ring, so the apply passes take nodes without atomics, and a batch the
// ring cannot hold is refused before any bucket changes.  Every touched
// chain is rewritten whole, its t rows spread evenly over its old nodes
// (in chain order) followed by its new ones.
//
// Buckets whose chain and slice fit registers and LDS (at most kFastNodes
// nodes, kFastKeys keys, kFastOut nodes after) take insert_fast_kernel, a
// wave each; the others take insert_big_kernel, a workgroup each, which
// stages the chain's rows in global scratch (sized by the same plan).
constexpr u32 kFastNodes = 16;
constexpr u32 kFastKeys = 512;
constexpr u32 kFastOut = 64;

// Nodes of a chain of L nodes after it holds t keys: unchanged while they
// fit, else half-full nodes (the reference splits a full node into two
// halves, FliX/tile_utils.cuh:306-391).
__host__ __device__ __forceinline__ u32 chain_nodes_after(u64 t, u32 L,
                                                          u32 S) {
  if (t <= u64(L) * S) return L;
  const u64 h = S / 2;
  const u64 n = (t + h - 1) / h;
  return n > L ? static_cast<u32>(n) : L;
}

// Device counters of one insert's plan (cleared before it).
struct InsertPlan {
  u32 n_fast;          // buckets rewritten by insert_fast_kernel
  u32 n_big;           // buckets rewritten by insert_big_kernel
  u32 new_nodes;       // ring nodes the insert takes
  u32 max_nodes;       // longest chain after the insert
  u32 big_nodes;       // chain nodes of the big buckets (scratch ids)
  u32 pad;
  unsigned long long big_keys;  // chain rows of the big buckets (scratch)
};

// A big bucket's work: its first new node's offset from the ring cursor
// and its scratch (node ids and row ranks at id_off, rows at key_off).
struct BigItem {
  u32 bucket;
  u32 node_off;
  u32 id_off;
  u32 pad;
  u64 key_off;
};

// Touched buckets -> fast items (bucket, first new node's offset from the
// ring cursor) and big items.  A block covers kPlanBuckets consecutive
// buckets (16 per thread) and reserves its share of every list, of the new
// nodes and of the scratch with one atomic per counter.
constexpr u32 kPlanBuckets = 4096;
static __global__ __launch_bounds__(256) void insert_plan_kernel(
    const u64 *begin, const u32 *bsize, const u32 *blen, u32 buckets, u32 S,
    uint2 *fast_items, BigItem *big_items, InsertPlan *plan) {
  constexpr u32 kPer = kPlanBuckets / 256;
  constexpr u32 kW = 256 / kWave;
  __shared__ u32 s_f[kW + 1], s_n[kW + 1], s_b[kW + 1], s_i[kW + 1];
  __shared__ u64 s_k[kW + 1];
  __shared__ u32 s_base[4];
  __shared__ u64 s_kbase;
  const u64 first = (u64(blockIdx.x) * 256 + threadIdx.x) * kPer;
  u32 fast_mask = 0, big_mask = 0, nf = 0, nb = 0, nodes = 0, ids = 0;
  u32 longest = 0;
  u64 keys = 0;
#pragma unroll
  for (u32 j = 0; j < kPer; ++j) {
    const u64 b = first + j;
    if (b >= buckets) break;
    const u64 m = begin[b + 1] - begin[b];
    if (!m) continue;
    const u32 L = blen[b];
    const u32 t0 = bsize[b];
    const u32 Lp = chain_nodes_after(t0 + m, L, S);
    nodes += Lp - L;
    longest = Lp > longest ? Lp : longest;
    if (L <= kFastNodes && t0 + m <= kFastKeys && Lp <= kFastOut) {
      fast_mask |= 1u << j;
      ++nf;
    } else {
      big_mask |= 1u << j;
      ++nb;
      ids += L;
      keys += t0;
    }
  }
  const u32 ex_f = block_exclusive_scan<256>(nf, s_f);
  const u32 ex_n = block_exclusive_scan<256>(nodes, s_n);
  const u32 ex_b = block_exclusive_scan<256>(nb, s_b);
  const u32 ex_i = block_exclusive_scan<256>(ids, s_i);
  const u64 ex_k = block_exclusive_scan_u64<256>(keys, s_k);
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const u32 o = __shfl_xor(longest, off, kWave);
    longest = o > longest ? o : longest;
  }
  if (lane_id() == 0 && longest) atomicMax(&plan->max_nodes, longest);
  if (threadIdx.x == 0) {
    s_base[0] = s_f[kW] ? atomicAdd(&plan->n_fast, s_f[kW]) : 0;
    s_base[1] = s_n[kW] ? atomicAdd(&plan->new_nodes, s_n[kW]) : 0;
    s_base[2] = s_b[kW] ? atomicAdd(&plan->n_big, s_b[kW]) : 0;
    s_base[3] = s_i[kW] ? atomicAdd(&plan->big_nodes, s_i[kW]) : 0;
    s_kbase = s_k[kW] ? atomicAdd(&plan->big_keys,
                                  (unsigned long long)s_k[kW])
                      : 0;
  }
  __syncthreads();
  u32 of = s_base[0] + ex_f, on = s_base[1] + ex_n, ob = s_base[2] + ex_b;
  u32 oi = s_base[3] + ex_i;
  u64 ok = s_kbase + ex_k;
#pragma unroll
  for (u32 j = 0; j < kPer; ++j) {
    const u64 b = first + j;
    const bool fast = (fast_mask >> j) & 1u, big = (big_mask >> j) & 1u;
    if (!fast && !big) continue;
    const u64 m = begin[b + 1] - begin[b];
    const u32 L = blen[b];
    const u32 t0 = bsize[b];
    const u32 add = chain_nodes_after(t0 + m, L, S) - L;
    if (fast) {
      fast_items[of++] = make_uint2(static_cast<u32>(b), on);
    } else {
      big_items[ob++] = BigItem{static_cast<u32>(b), on, oi, 0u, ok};
      oi += L;
      ok += t0;
    }
    on += add;
  }
}

// position of the rank-th of t rows spread over n nodes (the first t % n
// nodes hold one row more)
__device__ __forceinline__ void spread_slot(u32 rank, u32 q, u32 r, u32 &j,
                                            u32 &slot) {
  const u32 big = r * (q + 1);
  if (rank < big) {
    j = rank / (q + 1);
    slot = rank - j * (q + 1);
  } else {
    const u32 k = rank - big;
    j = r + k / q;
    slot = k - (j - r) * q;
  }
}

// One wave per fast item (a half-wave when nodes have at most 32 slots, so
// a wave rewrites two chains): rewrites the bucket's chain with its slice
// merged in (see the plan above).  The batch holds only keys absent from the
// index, sorted and unique, so every rank is distinct.
template <class K, class V, u32 kLanes>
static __global__ __launch_bounds__(256) void insert_fast_kernel(
    Pool<K, V> p, const u32 *ring_ids, u64 ring_cap,
    const unsigned long long *alloc_cursor, const K *bk, const V *bv,
    const u64 *begin, const uint2 *items, u32 n_items) {
  static_assert(kLanes == 32 || kLanes == 64, "a node per wave or half-wave");
  constexpr u32 kParts = kWave / kLanes;  // items per wave
  constexpr u32 kItemsPerBlock = 4 * kParts;
  __shared__ K s_k[kItemsPerBlock][kFastKeys];
  __shared__ u32 s_seq[kItemsPerBlock][kFastOut];
  const u32 lane = lane_id() % kLanes;  // the slot this lane holds
  const u32 w = threadIdx.x / kLanes;   // the item of the block
  const u32 item = blockIdx.x * kItemsPerBlock + w;
  if (item >= n_items) return;
  const uint2 it = items[item];
  const u32 b = it.x;
  const u64 pos = begin[b];
  const u32 m = static_cast<u32>(begin[b + 1] - pos);
  const u32 L = p.blen[b];
  const u32 t0 = p.bsize[b];
  const u32 t = t0 + m;
  const u32 S = p.S;
  const u32 Lp = chain_nodes_after(t, L, S);
  K *sk = s_k[w];
  u32 *seq = s_seq[w];
  // the slice's keys behind the chain's keys: sk = [chain keys | slice]
  for (u32 j = lane; j < m; j += kLanes) sk[t0 + j] = bk[pos + j];
  // the chain, a node per round trip: its slots stay in registers
  K ak[kFastNodes];
  V av[kFastNodes];
  u32 at[kFastNodes];  // rank of the node's first key among the chain keys
  u32 cn[kFastNodes];
  u32 node = b, a_at = 0;
#pragma unroll
  for (u32 i = 0; i < kFastNodes; ++i) {
    cn[i] = 0;
    at[i] = a_at;
    if (i < L && node < p.total_nodes) {
      const u32 c = p.count[node];
      const u32 nxt = p.next[node];
      const u64 base = u64(node) * S;
      if (lane < c) {
        ak[i] = p.keys[base + lane];
        av[i] = p.vals[base + lane];
        if (a_at + lane < t0) sk[a_at + lane] = ak[i];
      }
      if (lane == 0) seq[i] = node;
      cn[i] = c;
      a_at += c;
      node = nxt - 1;
    }
  }
  // new nodes reserved by the plan
  const unsigned long long first_new = *alloc_cursor + it.y;
  for (u32 j = lane; j + L < Lp; j += kLanes)
    seq[L + j] = ring_ids[(first_new + j) % ring_cap];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const u32 q = t / Lp, r = t % Lp;
  // chain rows: rank = own rank + slice keys below
#pragma unroll
  for (u32 i = 0; i < kFastNodes; ++i) {
    if (lane < cn[i]) {
      const K key = ak[i];
      u32 lo = 0, hi = m;
      while (lo < hi) {
        const u32 mid = (lo + hi) >> 1;
        if (sk[t0 + mid] < key) lo = mid + 1;
        else hi = mid;
      }
      const u32 rank = at[i] + lane + lo;
      if (rank < t) {  // always, while bsize matches the chain
        u32 j, slot;
        spread_slot(rank, q, r, j, slot);
        const u64 dst = u64(seq[j]) * S + slot;
        p.keys[dst] = key;
        p.vals[dst] = av[i];
      }
    }
  }
  // slice rows: rank = own index + chain keys below
  for (u32 x = lane; x < m; x += kLanes) {
    const K key = sk[t0 + x];
    u32 lo = 0, hi = t0;
    while (lo < hi) {
      const u32 mid = (lo + hi) >> 1;
      if (sk[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    u32 j, slot;
    spread_slot(x + lo, q, r, j, slot);
    const u64 dst = u64(seq[j]) * S + slot;
    p.keys[dst] = key;
    p.vals[dst] = bv[pos + x];
  }
  // counts and links: chain nodes in order, then the new ones
  for (u32 j = lane; j < Lp; j += kLanes) {
    const u32 id = seq[j];
    p.count[id] = q + (j < r ? 1u : 0u);
    p.next[id] = j + 1 < Lp ? seq[j + 1] + 1 : 0u;
  }
  if (lane == 0) {
    p.bsize[b] = t;
    p.blen[b] = Lp;
  }
}

// One workgroup per big item (a chain or slice too large for a wave's
// registers and LDS): thread 0 walks the chain, recording its node ids and
// each node's first row rank in scratch; the waves copy the chain's rows
// to scratch; then every row takes its merged rank (a binary search in the
// other sequence) and its place in the rewritten chain, as in
// insert_fast_kernel.  The pool is only written after every read of it.
template <class K, class V>
static __global__ __launch_bounds__(256) void insert_big_kernel(
    Pool<K, V> p, const u32 *ring_ids, u64 ring_cap,
    const unsigned long long *alloc_cursor, const K *bk, const V *bv,
    const u64 *begin, const BigItem *items, u32 *sc_ids, u32 *sc_at,
    K *sc_keys, V *sc_vals) {
  constexpr u32 kWaves = 256 / kWave;
  const BigItem it = items[blockIdx.x];
  const u32 b = it.bucket;
  const u64 pos = begin[b];
  const u32 m = static_cast<u32>(begin[b + 1] - pos);
  const u32 L = p.blen[b];
  const u32 t0 = p.bsize[b];
  const u32 t = t0 + m;
  const u32 S = p.S;
  const u32 Lp = chain_nodes_after(t, L, S);
  u32 *ids = sc_ids + it.id_off;
  u32 *at = sc_at + it.id_off;
  K *ak = sc_keys + it.key_off;
  V *av = sc_vals + it.key_off;
  if (threadIdx.x == 0) {
    u32 node = b, a = 0;
    for (u32 i = 0; i < L; ++i) {
      // a chain shorter than blen (never, while blen matches the chain)
      // repeats its last node rather than leaving the pool
      if (node >= p.total_nodes) node = ids[i - 1];
      ids[i] = node;
      at[i] = a;
      a += p.count[node];
      node = p.next[node] - 1;
    }
  }
  __syncthreads();
  const u32 lane = lane_id();
  for (u32 i = threadIdx.x / kWave; i < L; i += kWaves) {
    const u32 node = ids[i];
    const u32 c = p.count[node];
    const u32 a = at[i];
    if (lane < c && a + lane < t0) {
      ak[a + lane] = p.keys[u64(node) * S + lane];
      av[a + lane] = p.vals[u64(node) * S + lane];
    }
  }
  __syncthreads();
  const u32 q = t / Lp, r = t % Lp;
  const unsigned long long first_new = *alloc_cursor + it.node_off;
  auto node_at = [&](u32 j) -> u32 {
    return j < L ? ids[j] : ring_ids[(first_new + (j - L)) % ring_cap];
  };
  // chain rows: rank = own rank + slice keys below
  for (u32 i = threadIdx.x; i < t0; i += 256) {
    const K key = ak[i];
    const u32 below = static_cast<u32>(lower_bound_g(bk, pos, pos + m, key) -
                                       pos);
    u32 j, slot;
    if (i + below >= t) continue;  // never, while bsize matches the chain
    spread_slot(i + below, q, r, j, slot);
    const u64 dst = u64(node_at(j)) * S + slot;
    p.keys[dst] = key;
    p.vals[dst] = av[i];
  }
  // slice rows: rank = own index + chain rows below
  for (u32 x = threadIdx.x; x < m; x += 256) {
    const K key = bk[pos + x];
    const u32 below = static_cast<u32>(lower_bound_g(ak, 0, t0, key));
    u32 j, slot;
    spread_slot(x + below, q, r, j, slot);
    const u64 dst = u64(node_at(j)) * S + slot;
    p.keys[dst] = key;
    p.vals[dst] = bv[pos + x];
  }
  for (u32 j = threadIdx.x; j < Lp; j += 256) {
    const u32 id = node_at(j);
    p.count[id] = q + (j < r ? 1u : 0u);
    p.next[id] = j + 1 < Lp ? node_at(j + 1) + 1 : 0u;
  }
  if (threadIdx.x == 0) {
    p.bsize[b] = t;
    p.blen[b] = Lp;
  }
}

// Moves the ring cursor past the insert's new nodes (after the apply
// passes read it).
static __global__ void ring_advance_kernel(unsigned long long *alloc_cursor,
                                           const InsertPlan *plan) {
  *alloc_cursor += plan->new_nodes;
}

// Clears an insert's plan counters.
static __global__ void plan_reset_kernel(InsertPlan *plan) {
  *plan = InsertPlan{0u, 0u, 0u, 0u, 0u, 0u, 0ull};
}

// Removes bucket b's slice dk[pos, end) of the sorted delete batch from its
// chain (one wave): a ballot keep-mask compacts each node, emptied non-head
// nodes are unlinked and returned to the ring, bsize/blen follow.  dk is the
// batch in global memory or the wave's copy of the slice in LDS (then pos =
// 0); every search of the slice reads dk.
template <class K, class V, u32 kLanes>
__device__ __forceinline__ void delete_bucket(Pool<K, V> &p, Ring &ring,
                                              const K *dk, u64 pos, u64 end,
                                              u32 b, u32 lane) {
  const u32 S = p.S;
  // the ballot bits of this lane's group (the wave, or its half holding
  // this bucket), lane = the slot index inside the group
  const u32 shift = kLanes == 64 ? 0u : (lane_id() & 32u);
  auto group_ballot = [&](bool v) {
    const u64 all = ballot(v);
    return kLanes == 64 ? all : (all >> shift) & 0xffffffffull;
  };
  u32 node = b;
  u32 prev = ~0u;
  u32 removed = 0, unlinked = 0;  // wave-uniform
  while (true) {
    const u32 nxt = p.next[node];
    // delete keys below the next node's first key (none when the slice's
    // first key is not below it)
    u64 q = end;
    if (nxt) {
      const K first_next = p.keys[u64(nxt - 1) * S];
      q = dk[pos] >= first_next ? pos : lower_bound_g(dk, pos, end, first_next);
    }
    bool emptied = false;
    if (q > pos) {
      const u32 c = p.count[node];
      const u64 base = u64(node) * S;
      K key = K(0);
      V val = V(0);
      bool hit = false;
      if (lane < c) {
        key = p.keys[base + lane];
        val = p.vals[base + lane];
        const u64 f = lower_bound_g(dk, pos, q, key);
        hit = f < q && dk[f] == key;
      }
      const u64 keep = group_ballot(lane < c && !hit);
      const u32 kept = popc64(keep);
      if (kept != c) {
        __builtin_amdgcn_wave_barrier();
        if ((keep >> lane) & 1ull) {
          const u32 o = popc64(keep & ((u64(1) << lane) - 1));
          p.keys[base + o] = key;
          p.vals[base + o] = val;
        }
        if (lane == 0) p.count[node] = kept;
      }
      removed += c - kept;
      emptied = kept == 0 && node != b;
    }
    pos = q;
    if (emptied) {
      // unlink the empty non-head node and give it back to the ring
      if (lane == 0) {
        p.next[prev] = nxt;
        p.next[node] = 0;
        const unsigned long long slot = atomicAdd(ring.free_count, 1ull);
        ring.ids[slot % ring.cap] = node;
      }
      ++unlinked;
    } else {
      prev = node;
    }
    if (pos >= end || !nxt) break;
    node = nxt - 1;
  }
  if (lane == 0 && (removed || unlinked)) {
    p.bsize[b] -= removed;
    p.blen[b] -= unlinked;
  }
}

// One wave per bucket (a half-wave when nodes have at most 32 slots) with a
// non-empty slice of the sorted delete batch;
// the waves stride over the work list, whose length the device holds (the
// grid is sized by an upper bound, so the host reads nothing back).
// A slice of at most kDeleteStage keys is copied to the wave's LDS first, so
// the per-node searches of the slice (one per lane and one for the next
// node's first key) read LDS instead of dependent global loads.
constexpr u32 kDeleteStage = 512;

template <class K, class V, u32 kLanes>
static __global__ __launch_bounds__(256) void delete_kernel(
    Pool<K, V> p, Ring ring, const K *dk, const u64 *begin, const u32 *work,
    const u32 *work_count) {
  static_assert(kLanes == 32 || kLanes == 64, "a node per wave or half-wave");
  constexpr u32 kPerBlock = 256 / kLanes;  // buckets in flight per block
  __shared__ K s_dk[kPerBlock][kDeleteStage];
  const u32 lane = lane_id() % kLanes;
  const u32 w = threadIdx.x / kLanes;
  const u32 items = *work_count;
  for (u32 item = blockIdx.x * kPerBlock + w; item < items;
       item += gridDim.x * kPerBlock) {
    const u32 b = work[item];
    const u64 pos = begin[b], end = begin[b + 1];
    const u64 m = end - pos;
    if (m <= kDeleteStage) {
      K *sk = s_dk[w];
      // the previous bucket's searches are done before the slot is reused
      __builtin_amdgcn_wave_barrier();
      for (u32 j = lane; j < m; j += kLanes) sk[j] = dk[pos + j];
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      delete_bucket<K, V, kLanes>(p, ring, static_cast<const K *>(sk), 0, m,
                                  b, lane);
    } else {
      delete_bucket<K, V, kLanes>(p, ring, dk, pos, end, b, lane);
    }
  }
}

// Point lookups (any order).  miss -> all-ones value.
template <class K, class V>
static __global__ void lookup_kernel(Pool<K, V> p, const K *q, u64 n, V *out) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const K key = q[i];
    const u32 b = bucket_of(p, key);
    const u32 node = owner_node(p, b, key);
    const u32 c = p.count[node];
    const K *nk = p.keys + u64(node) * p.S;
    const u32 lo = node_lower_bound(nk, c, key);
    out[i] = (lo < c && nk[lo] == key) ? p.vals[u64(node) * p.S + lo]
                                       : ~V(0);
  }
}

// Lookups of queries grouped by their top varying key byte (qk), results
// in grouped order.  Block b serves the contiguous chunk
// (b % 8) * (G / 8) + b / 8 of the grouped queries, so each XCD (blocks are
// dealt to the 8 XCDs in turn) walks its own eighth of the key space and
// the buckets and nodes it reads stay in its L2.  G is a multiple of 8.
template <class K, class V>
static __global__ void lookup_grouped_kernel(Pool<K, V> p, const K *qk, u64 n,
                                             V *res) {
  const u32 G = gridDim.x;
  const u32 logical = (blockIdx.x % 8u) * (G / 8u) + blockIdx.x / 8u;
  const u64 chunk = (n + G - 1) / G;
  const u64 begin = u64(logical) * chunk;
  const u64 end = begin + chunk < n ? begin + chunk : n;
  for (u64 i = begin + threadIdx.x; i < end; i += blockDim.x) {
    const K key = qk[i];
    const u32 b = bucket_of(p, key);
    const u32 node = owner_node(p, b, key);
    const u32 c = p.count[node];
    const K *nk = p.keys + u64(node) * p.S;
    const u32 lo = node_lower_bound_bisect(nk, c, key);
    res[i] = (lo < c && nk[lo] == key) ? p.vals[u64(node) * p.S + lo]
                                       : ~V(0);
  }
}

// out[i] = res[dest[i]]: the grouped results back in query order.
template <class V>
static __global__ void lookup_gather_kernel(const u32 *dest, const V *res,
                                            u64 n, V *out) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride)
    out[i] = res[dest[i]];
}

// Two-level grouping of any-order lookups (stable LSD passes: the digit
// below the queries' top varying one, then the top one), so that a group
// of queries covers 1/2^16 of the key spread instead of 1/2^8 and the
// nodes it reads stay in L2 while the group is looked up.
//
// shift_lo = the spread partition's shift - 8 (0 when the spread is
// narrower than 16 bits; the two digits then overlap, which changes the
// order only).
static __global__ void shift_below_kernel(const u32 *shift_top,
                                          u32 *shift_lo) {
  const u32 t = *shift_top;
  *shift_lo = t >= 8 ? t - 8 : 0u;
}

// rows[b % kHistRows][d] += the block's count of digit (key >> shift) & 0xff
// (the layout rs::scan_kernel sums, for one pass).
template <class K>
static __global__ __launch_bounds__(rs::kThreads) void digit_hist_kernel(
    const K *__restrict__ keys, u32 n, const u32 *shift_dev, u32 *rows) {
  static_assert(rs::kThreads == rs::kRadix, "a thread per digit");
  __shared__ u32 s_hist[rs::kRadix];
  s_hist[threadIdx.x] = 0;
  __syncthreads();
  const u32 shift = *shift_dev;
  const u64 round = u64(rs::kThreads) * rs::kHistItems;
  for (u64 base = u64(blockIdx.x) * round; base < n;
       base += u64(gridDim.x) * round) {
    K k[rs::kHistItems];
#pragma unroll
    for (int j = 0; j < rs::kHistItems; ++j) {
      const u64 i = base + u64(j) * rs::kThreads + threadIdx.x;
      k[j] = i < n ? keys[i] : K(0);
    }
#pragma unroll
    for (int j = 0; j < rs::kHistItems; ++j)
      if (base + u64(j) * rs::kThreads + threadIdx.x < n)
        atomicAdd(&s_hist[static_cast<u32>(k[j] >> shift) & 0xffu], 1u);
  }
  __syncthreads();
  const u32 c = s_hist[threadIdx.x];
  if (c)
    atomicAdd(&rows[u64(blockIdx.x % rs::kHistRows) * rs::kRadix +
                    threadIdx.x],
              c);
}

// out[i] = res[dest_hi[dest_lo[i]]]: the results of the two-level grouping
// back in query order.
template <class V>
static __global__ void lookup_gather2_kernel(const u32 *dest_lo,
                                             const u32 *dest_hi,
                                             const V *res, u64 n, V *out) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride)
    out[i] = res[dest_hi[dest_lo[i]]];
}

// First stored key >= key, starting at bucket b; (node, slot) or node = ~0.
template <class K, class V>
__device__ __forceinline__ void ceiling_position(const Pool<K, V> &p, K key,
                                                 u32 &node_out,
                                                 u32 &slot_out) {
  u32 b = bucket_of(p, key);
  u32 node = owner_node(p, b, key);
  while (true) {
    const u32 c = p.count[node];
    const K *nk = p.keys + u64(node) * p.S;
    const u32 lo = node_lower_bound(nk, c, key);
    if (lo < c) {
      node_out = node;
      slot_out = lo;
      return;
    }
    const u32 nxt = p.next[node];
    if (nxt) {
      node = nxt - 1;
      continue;
    }
    if (++b >= p.buckets) {
      node_out = ~0u;
      return;
    }
    node = b;
  }
}

template <class K, class V>
static __global__ void successor_kernel(Pool<K, V> p, const K *q, u64 n, K *out) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    u32 node, slot;
    ceiling_position(p, q[i], node, slot);
    // key_type(not_found): 0xFFFFFFFF zero-extended for 64-bit keys
    // (FliX/coarse_granular_lookups_tile_bulk.cuh:672,738)
    out[i] = node == ~0u ? static_cast<K>(0xffffffffu)
                         : p.keys[u64(node) * p.S + slot];
  }
}

// Keys of the index below `key` (kInclusive: up to and including it): the
// keys of the buckets before its bucket (bprefix, an exclusive scan of
// bsize), of its chain's nodes before the one that holds the rank's end,
// and of that node's slots.  Keys of node i are below the first key of
// node i + 1, so the walk stops at the first node whose successor starts
// at or above the key.
template <class K, class V, bool kInclusive>
__device__ __forceinline__ u64 rank_of(const Pool<K, V> &p, const u64 *bprefix,
                                       K key) {
  const u32 b = bucket_of(p, key);
  u64 r = bprefix[b];
  u32 node = b;
  while (true) {
    const u32 nxt = p.next[node];
    if (!nxt) break;
    const K first = p.keys[u64(nxt - 1) * p.S];
    if (kInclusive ? key < first : key <= first) break;
    r += p.count[node];
    node = nxt - 1;
  }
  const u32 c = p.count[node];
  const K *nk = p.keys + u64(node) * p.S;
  return r + (kInclusive ? node_upper_bound(nk, c, key)
                         : node_lower_bound(nk, c, key));
}

// Count pass of the range enumeration: |[lo, hi]| = rank(<= hi) - rank(< lo),
// two point searches per query instead of a walk over the range.
template <class K, class V>
static __global__ void range_count_kernel(Pool<K, V> p, const u64 *bprefix,
                                          const K *lo, const K *hi, u64 n,
                                          u64 *counts) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const K a = lo[i], z = hi[i];
    counts[i] = a <= z ? rank_of<K, V, true>(p, bprefix, z) -
                             rank_of<K, V, false>(p, bprefix, a)
                       : 0;
  }
}

// Fill pass of the range enumeration: a wave per query whose lanes hold one
// node's slots, so a node's keys in range leave with one ballot and
// coalesced stores, and a query walks its range a node per step rather than
// a key per step.  offsets[i] is where query i's pairs start; S <= 64.
// The walk starts at the node of lo's bucket chain that would own lo (the
// lanes below lo are masked, so no separate ceiling search), and each step
// issues the node's count, successor link, keys and values together: the
// slots past the count are in bounds (every node has S of them) and are
// masked, so a node costs one round of loads instead of three dependent
// ones (count, then keys and values, then the link).
template <class K, class V>
static __global__ __launch_bounds__(256) void range_wave_kernel(
    Pool<K, V> p, const K *lo, const K *hi, u64 n, const u64 *offsets,
    K *out_keys, V *out_vals) {
  const u32 lane = lane_id();
  const bool slot_lane = lane < p.S;
  const u64 waves = u64(gridDim.x) * (blockDim.x / kWave);
  for (u64 i = u64(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
       i < n; i += waves) {
    const K a = lo[i], z = hi[i];
    if (a > z) continue;
    u32 bucket = bucket_of(p, a);
    u32 node = owner_node(p, bucket, a);
    u64 w = offsets[i];
    while (true) {
      const u64 base = u64(node) * p.S;
      const u32 c = p.count[node];
      const u32 nxt = p.next[node];
      K k = K(0);
      V v = V(0);
      if (slot_lane) {
        k = p.keys[base + lane];
        v = p.vals[base + lane];
      }
      const bool live = lane < c;
      // keys ascend inside a node and across nodes: the lanes below lo
      // are a prefix, the lanes in range the run after it
      const u32 below = popc64(ballot(live && k < a));
      const bool in = live && k >= a && k <= z;
      const u32 m = popc64(ballot(in));
      if (in) {
        out_keys[w + (lane - below)] = k;
        out_vals[w + (lane - below)] = v;
      }
      w += m;
      if (below + m < c) break;  // a key above the range
      // next node in key order (an empty head has count 0 and passes)
      if (nxt) node = nxt - 1;
      else if (++bucket < p.buckets) node = bucket;
      else break;
    }
  }
}

// Live keys per bucket chain (for totals and rebuild offsets).
template <class K, class V>
static __global__ void chain_sizes_kernel(Pool<K, V> p, u64 *sizes) {
  const u32 b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= p.buckets) return;
  u64 s = 0;
  u32 node = b;
  while (true) {
    s += p.count[node];
    const u32 nxt = p.next[node];
    if (!nxt) break;
    node = nxt - 1;
  }
  sizes[b] = s;
}

// Exclusive scan of u64 values with a single-pass chained scan over tiles
// of kScanTile values, 16 per thread: coalesced loads into LDS, each
// thread sums 16 consecutive values, a block scan of the thread sums, the
// tile's look-back, then coalesced stores.  The LDS index is padded by one
// slot per 16 so a thread's consecutive values fall in distinct banks.
// out[n] receives the total (the value at n counts as 0).
constexpr u32 kScanItems = 16;
constexpr u32 kScanTile = 256 * kScanItems;
__device__ __forceinline__ u32 scan_pad(u32 i) { return i + i / kScanItems; }
static __global__ __launch_bounds__(256) void scan_u64_kernel(const u64 *in, u64 n,
                                                       u64 *out, u64 *status,
                                                       u32 *ticket) {
  __shared__ u32 s_tile;
  __shared__ u64 s_v[kScanTile + kScanTile / kScanItems];
  __shared__ u64 s_part[256 / kWave + 1];
  __shared__ u64 s_base;
  if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const u32 tile = s_tile;
  const u64 base = u64(tile) * kScanTile;
#pragma unroll
  for (u32 j = 0; j < kScanItems; ++j) {
    const u32 e = j * 256 + threadIdx.x;
    s_v[scan_pad(e)] = base + e < n ? in[base + e] : 0;
  }
  __syncthreads();
  u64 local[kScanItems];
  u64 sum = 0;
#pragma unroll
  for (u32 j = 0; j < kScanItems; ++j) {
    local[j] = s_v[scan_pad(threadIdx.x * kScanItems + j)];
    sum += local[j];
  }
  // block scan of the thread sums (u64)
  u64 inc = sum;
  const u32 lane = lane_id();
  for (int off = 1; off < kWave; off <<= 1) {
    const u64 o = __shfl_up(inc, off, kWave);
    if (lane >= static_cast<u32>(off)) inc += o;
  }
  const u32 wv = threadIdx.x / kWave;
  if (lane == kWave - 1) s_part[wv] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 run = 0;
    for (int k = 0; k < 256 / kWave; ++k) {
      const u64 t = s_part[k];
      s_part[k] = run;
      run += t;
    }
    s_part[256 / kWave] = run;
  }
  __syncthreads();
  if (threadIdx.x < kWave) {
    const u64 ex = wave_lookback(status, tile, s_part[256 / kWave]);
    if (threadIdx.x == 0) s_base = ex;
  }
  __syncthreads();
  u64 run = s_base + s_part[wv] + inc - sum;
#pragma unroll
  for (u32 j = 0; j < kScanItems; ++j) {
    s_v[scan_pad(threadIdx.x * kScanItems + j)] = run;
    run += local[j];
  }
  __syncthreads();
#pragma unroll
  for (u32 j = 0; j < kScanItems; ++j) {
    const u32 e = j * 256 + threadIdx.x;
    if (base + e <= n) out[base + e] = s_v[scan_pad(e)];
  }
}

// Compacts the marked rows of a sorted batch and gathers their values
// (sorted positions sp index the caller's value array; build passes values
// = positions themselves).
template <class K, class V>
static __global__ void compact_batch_kernel(const K *sk, const u32 *sp, u64 n,
                                     const u32 *keep, const u64 *offs,
                                     const V *values, int values_are_positions,
                                     K *ok, V *ov) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    if (!keep[i]) continue;
    const u64 o = offs[i];
    ok[o] = sk[i];
    ov[o] = values_are_positions ? static_cast<V>(sp[i]) : values[sp[i]];
  }
}

// Keeps the first occurrence of every key of a stably sorted batch.
template <class K>
static __global__ void mark_first_kernel(const K *sk, u64 n, u32 *keep) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    keep[i] = (i == 0 || sk[i - 1] != sk[i]) ? 1u : 0u;
}

// Buckets with a non-empty batch slice -> compact work list (in any
// order).  A block covers kWorkListBuckets consecutive buckets, 16 per
// thread, and takes its share of the list with one atomic, so the shared
// counter sees one update per 4096 buckets.
constexpr u32 kWorkListBuckets = 4096;
static __global__ __launch_bounds__(256) void work_list_kernel(
    const u64 *begin, u32 buckets, u32 *work, u32 *work_count) {
  constexpr u32 kPer = kWorkListBuckets / 256;
  __shared__ u32 s_scan[256 / kWave + 1];
  __shared__ u32 s_base;
  const u64 first = (u64(blockIdx.x) * 256 + threadIdx.x) * kPer;
  u32 mask = 0, c = 0;
#pragma unroll
  for (u32 j = 0; j < kPer; ++j) {
    const u64 b = first + j;
    if (b < buckets && begin[b + 1] > begin[b]) {
      mask |= 1u << j;
      ++c;
    }
  }
  const u32 ex = block_exclusive_scan<256>(c, s_scan);
  if (threadIdx.x == 0) s_base = atomicAdd(work_count, s_scan[256 / kWave]);
  __syncthreads();
  u32 o = s_base + ex;
#pragma unroll
  for (u32 j = 0; j < kPer; ++j)
    if ((mask >> j) & 1u) work[o++] = static_cast<u32>(first + j);
}

template <class K, class V>
static __global__ void node_total_kernel(Pool<K, V> p, unsigned long long *total) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  u64 local = 0;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < p.buckets;
       i += stride) {
    u32 node = static_cast<u32>(i);
    while (true) {
      local += p.count[node];
      const u32 nxt = p.next[node];
      if (!nxt) break;
      node = nxt - 1;
    }
  }
  local = wave_sum_u64(local);
  if (lane_id() == 0 && local) atomicAdd(total, (unsigned long long)local);
}

}  // namespace flix
}  // namespace gdict
