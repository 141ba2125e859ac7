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

namespace gdict {
namespace flix {

constexpr int kMaxSlots = 64;
constexpr u32 kPieceKeys = 2048;  // batch keys merged into a node per step
constexpr u32 kMaxOutNodes = 2 * (kPieceKeys + kMaxSlots) / 8 + 2;

template <class K, class V>
struct Pool {
  K *keys;
  V *vals;
  u32 *count;
  u32 *next;
  const K *maxv;
  u32 S;
  u32 buckets;
  u64 total_nodes;
};

// Ring of free node ids.  alloc_cursor / free_count grow monotonically;
// live ids sit at [alloc_cursor, free_count) modulo ring_cap.
struct Ring {
  u32 *ids;
  u64 cap;
  unsigned long long *alloc_cursor;
  unsigned long long *free_count;
  unsigned long long *first_fail;  // smallest cursor of a failed allocation
  u32 *fail_flag;
};

template <class K>
__device__ __forceinline__ u32 bucket_of(const K *maxv, u32 buckets, K key) {
  // first bucket whose max >= key
  u32 lo = 0, hi = buckets - 1;
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    if (maxv[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
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
      p.count[b] = first < n ? static_cast<u32>(end - first) : 0u;
      p.next[b] = 0;
      maxv_out[b] = (b + 1 == p.buckets) ? ~K(0) : keys[end - 1];
    }
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
                                     const K *batch, u64 n, u64 *begin) {
  const u32 b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > buckets) return;
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
      const u32 b = bucket_of(p.maxv, p.buckets, key);
      const u32 node = owner_node(p, b, key);
      const u32 c = p.count[node];
      const K *nk = p.keys + u64(node) * p.S;
      u32 lo = 0, hi = c;
      while (lo < hi) {
        const u32 mid = (lo + hi) >> 1;
        if (nk[mid] < key) lo = mid + 1;
        else hi = mid;
      }
      if (lo < c && nk[lo] == key) fresh = false;
    }
    keep[i] = fresh ? 1u : 0u;
  }
}

// Ring allocation of k nodes for one wave (lane 0 calls).  Returns the first
// ring position or ~0ull on failure.
__device__ __forceinline__ unsigned long long ring_take(const Ring &r,
                                                        u32 k) {
  const unsigned long long old = atomicAdd(r.alloc_cursor, (unsigned long long)k);
  const unsigned long long avail = *r.free_count;
  if (old + k > avail) {
    atomicMin(r.first_fail, old);
    atomicExch(r.fail_flag, 1u);
    return ~0ull;
  }
  return old;
}

// One wave per bucket with a non-empty slice of the (deduplicated, sorted)
// insert batch.  Keys of a node's range are merged with its slots; when the
// result overflows the node it is spread evenly over half-full nodes taken
// from the ring and linked in order (a node split,
// FliX/tile_utils.cuh:306-391, here for any number of keys at once).
template <class K, class V>
static __global__ __launch_bounds__(256) void insert_kernel(
    Pool<K, V> p, Ring ring, const K *bk, const V *bv, const u64 *begin,
    const u32 *work, u32 work_count) {
  constexpr int kWavesPerBlock = 4;
  __shared__ K s_keys[kWavesPerBlock][kMaxSlots];
  __shared__ u32 s_out[kWavesPerBlock][kMaxOutNodes];
  __shared__ unsigned long long s_pos[kWavesPerBlock];
  const u32 lane = lane_id();
  const u32 w = threadIdx.x / kWave;
  const u32 item = blockIdx.x * kWavesPerBlock + w;
  if (item >= work_count) return;
  const u32 b = work[item];
  u64 pos = begin[b];
  const u64 end = begin[b + 1];
  const u32 S = p.S;
  u32 node = b;
  while (pos < end) {
    const u32 nxt = p.next[node];
    // batch keys below the first key of the next node belong here; when
    // the slice's first key is not below it, none does (one comparison
    // instead of a search of the slice)
    u64 q = end;
    if (nxt) {
      const K first_next = p.keys[u64(nxt - 1) * S];
      if (bk[pos] >= first_next) {
        node = nxt - 1;
        continue;
      }
      q = lower_bound_g(bk, pos, end, first_next);
    }
    if (q - pos > kPieceKeys) q = pos + kPieceKeys;
    const u32 m = static_cast<u32>(q - pos);
    const u32 c = p.count[node];
    const u64 base = u64(node) * S;
    K ak = K(0);
    V av = V(0);
    if (lane < c) {
      ak = p.keys[base + lane];
      av = p.vals[base + lane];
      s_keys[w][lane] = ak;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const u32 t = c + m;
    u32 outs = 1, h = t;
    if (t > S) {
      outs = (2 * t + S - 1) / S;  // half-full nodes after the split
      h = (t + outs - 1) / outs;
      outs = (t + h - 1) / h;
    }
    // take outs-1 nodes from the ring
    if (lane == 0) s_pos[w] = outs > 1 ? ring_take(ring, outs - 1) : 0ull;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const unsigned long long rp = s_pos[w];
    if (rp == ~0ull) return;  // pool exhausted: this bucket keeps its state
    s_out[w][0] = node;
    for (u32 i = lane; i + 1 < outs; i += kWave)
      s_out[w][i + 1] = ring.ids[(rp + i) % ring.cap];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // A rows: rank = own index + number of batch keys below it
    if (lane < c) {
      const u64 below = lower_bound_g(bk, pos, q, ak) - pos;
      const u32 o = lane + static_cast<u32>(below);
      const u64 dst = u64(s_out[w][o / h]) * S + (o % h);
      p.keys[dst] = ak;
      p.vals[dst] = av;
    }
    // batch rows: rank = own index + number of node keys below it
    for (u32 j = lane; j < m; j += kWave) {
      const K key = bk[pos + j];
      u32 lo = 0, hi = c;
      while (lo < hi) {
        const u32 mid = (lo + hi) >> 1;
        if (s_keys[w][mid] < key) lo = mid + 1;
        else hi = mid;
      }
      const u32 o = j + lo;
      const u64 dst = u64(s_out[w][o / h]) * S + (o % h);
      p.keys[dst] = key;
      p.vals[dst] = bv[pos + j];
    }
    // counts and links: node -> new_1 -> ... -> new_{outs-1} -> nxt
    for (u32 i = lane; i < outs; i += kWave) {
      const u32 id = s_out[w][i];
      p.count[id] = (i + 1 < outs) ? h : t - h * (outs - 1);
      p.next[id] = (i + 1 < outs) ? s_out[w][i + 1] + 1 : nxt;
    }
    // make this step's slots, counts and links visible to the next step's
    // loads, which this wave makes on this CU: a workgroup-scope fence waits
    // for the stores (the CU's L1 serves its own work-groups coherently).
    // An agent-scope fence here would write back the L2 once per step.
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    pos = q;
    // stay on `node`: the loop re-walks the freshly linked nodes
  }
}

// One wave per bucket with a non-empty slice of the sorted delete batch.
template <class K, class V>
static __global__ __launch_bounds__(256) void delete_kernel(
    Pool<K, V> p, Ring ring, const K *dk, const u64 *begin, const u32 *work,
    u32 work_count) {
  constexpr int kWavesPerBlock = 4;
  const u32 lane = lane_id();
  const u32 w = threadIdx.x / kWave;
  const u32 item = blockIdx.x * kWavesPerBlock + w;
  if (item >= work_count) return;
  const u32 b = work[item];
  u64 pos = begin[b];
  const u64 end = begin[b + 1];
  const u32 S = p.S;
  u32 node = b;
  u32 prev = ~0u;
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
      const u64 keep = ballot(lane < c && !hit);
      const u32 kept = popc64(keep);
      if (kept != c) {
        __builtin_amdgcn_wave_barrier();
        if ((keep >> lane) & 1ull) {
          const u32 o = popc64(keep & lanemask_lt());
          p.keys[base + o] = key;
          p.vals[base + o] = val;
        }
        if (lane == 0) p.count[node] = kept;
      }
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
    } else {
      prev = node;
    }
    if (pos >= end || !nxt) break;
    node = nxt - 1;
  }
}

// Point lookups (any order).  miss -> all-ones value.
template <class K, class V>
static __global__ void lookup_kernel(Pool<K, V> p, const K *q, u64 n, V *out) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const K key = q[i];
    const u32 b = bucket_of(p.maxv, p.buckets, key);
    const u32 node = owner_node(p, b, key);
    const u32 c = p.count[node];
    const K *nk = p.keys + u64(node) * p.S;
    u32 lo = 0, hi = c;
    while (lo < hi) {
      const u32 mid = (lo + hi) >> 1;
      if (nk[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    out[i] = (lo < c && nk[lo] == key) ? p.vals[u64(node) * p.S + lo]
                                       : ~V(0);
  }
}

// First stored key >= key, starting at bucket b; (node, slot) or node = ~0.
template <class K, class V>
__device__ __forceinline__ void ceiling_position(const Pool<K, V> &p, K key,
                                                 u32 &node_out,
                                                 u32 &slot_out) {
  u32 b = bucket_of(p.maxv, p.buckets, key);
  u32 node = owner_node(p, b, key);
  while (true) {
    const u32 c = p.count[node];
    const K *nk = p.keys + u64(node) * p.S;
    u32 lo = 0, hi = c;
    while (lo < hi) {
      const u32 mid = (lo + hi) >> 1;
      if (nk[mid] < key) lo = mid + 1;
      else hi = mid;
    }
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

// Next (node, slot) in key order, node = ~0 at the end.
template <class K, class V>
__device__ __forceinline__ void advance(const Pool<K, V> &p, u32 &node,
                                        u32 &slot, u32 &bucket) {
  if (slot + 1 < p.count[node]) {
    ++slot;
    return;
  }
  slot = 0;
  while (true) {
    const u32 nxt = p.next[node];
    if (nxt) node = nxt - 1;
    else if (++bucket < p.buckets) node = bucket;
    else {
      node = ~0u;
      return;
    }
    if (p.count[node]) return;
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

// Range [lo, hi]: count pass (out_keys == nullptr) or fill pass.
template <class K, class V>
static __global__ void range_kernel(Pool<K, V> p, const K *lo, const K *hi, u64 n,
                             u64 *offsets, K *out_keys, V *out_vals) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const K a = lo[i], z = hi[i];
    u64 cnt = 0;
    if (a <= z) {
      u32 node, slot;
      ceiling_position(p, a, node, slot);
      u32 bucket = node == ~0u ? 0 : bucket_of(p.maxv, p.buckets, a);
      // the bucket index only matters for crossing to the next bucket:
      // recover the bucket that owns `node` by its first key
      if (node != ~0u) bucket = bucket_of(p.maxv, p.buckets,
                                          p.keys[u64(node) * p.S + slot]);
      u64 w = out_keys ? offsets[i] : 0;
      while (node != ~0u) {
        const K k = p.keys[u64(node) * p.S + slot];
        if (k > z) break;
        if (out_keys) {
          out_keys[w] = k;
          out_vals[w] = p.vals[u64(node) * p.S + slot];
          ++w;
        }
        ++cnt;
        advance(p, node, slot, bucket);
      }
    }
    if (!out_keys) offsets[i] = cnt;
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

// Writes every chain into a flat sorted array at its offset (wave/bucket).
template <class K, class V>
static __global__ __launch_bounds__(256) void flatten_kernel(Pool<K, V> p,
                                                      const u64 *offsets,
                                                      K *ok, V *ov) {
  const u32 lane = lane_id();
  const u32 b = blockIdx.x * 4 + threadIdx.x / kWave;
  if (b >= p.buckets) return;
  u64 at = offsets[b];
  u32 node = b;
  while (true) {
    const u32 c = p.count[node];
    if (lane < c) {
      ok[at + lane] = p.keys[u64(node) * p.S + lane];
      ov[at + lane] = p.vals[u64(node) * p.S + lane];
    }
    at += c;
    const u32 nxt = p.next[node];
    if (!nxt) break;
    node = nxt - 1;
  }
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
