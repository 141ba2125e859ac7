// Host runtime of the FliX updatable index and its C ABI (gd_flix_*).
//
// Replaces the non-OptiX members of cg_rtx_index_updates
// (FliX/impl_cg_rtx_index_updates.cuh:1422-2693): build_bucket_layer_only
// (:1711), insert (:2195), remove (:2253), lookups_ordered (:2485),
// lookups_successor (:2647), rebuild_bucket_layer_only (:1821),
// compute_total_size (:1657), gpu_resident_bytes (:1600).  Batches are
// sorted internally (onesweep radix sort over 32- or 64-bit keys), so unsorted
// input is accepted; sorted input behaves as in the reference.
#include <algorithm>
#include <memory>
#include <string>

#include "flix_kernels.hpp"
#include "host_common.hpp"
#include "radix_sort.hpp"
#include "seg_admission.hpp"

namespace gdict {
namespace flix {

// Ring cursors and the plan of the current insert, read back together.
struct RingCounters {
  unsigned long long alloc_cursor;
  unsigned long long free_count;
  InsertPlan plan;
};

__global__ void ring_counters_init_kernel(RingCounters *c, u64 free_count) {
  c->alloc_cursor = 0;
  c->free_count = free_count;
  c->plan = InsertPlan{0u, 0u, 0u, 0u, 0u, 0u, 0ull};
}

__global__ void widen_kernel(const u32 *in, u64 n, u64 *out) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = in[i];
}

class FlixBase {
 public:
  virtual ~FlixBase() = default;
  virtual void build(const void *keys, u64 n, u64 max_size, hipStream_t s) = 0;
  virtual void insert(const void *keys, const void *vals, u64 n,
                      hipStream_t s) = 0;
  virtual void remove(const void *keys, u64 n, bool sorted, hipStream_t s) = 0;
  // sorted: the queries ascend (lookups_ordered), so they are not grouped
  virtual void lookup(const void *q, u64 n, void *out, bool sorted,
                      hipStream_t s) = 0;
  virtual void successor(const void *q, u64 n, void *out, hipStream_t s) = 0;
  virtual void range(const void *lo, const void *hi, u64 n, u64 *offsets,
                     void *ok, void *ov, hipStream_t s) = 0;
  virtual void rebuild(hipStream_t s) = 0;
  virtual u64 total_keys(hipStream_t s) = 0;
  virtual u64 resident_bytes() const = 0;
  virtual void memory(gd_memory_stats *out) const = 0;
  virtual void trim() = 0;
};

inline u64 pow2_at_least(u64 n) {
  u64 p = 1;
  while (p < n) p <<= 1;
  return p;
}

// Any-order lookups of this many queries or more are grouped first.
constexpr u64 kGroupedLookups = u64(1) << 18;

template <class K, class V>
class FlixImpl : public FlixBase {
 public:
  FlixImpl(u32 slots, u32 fill_percent, u32 rebuild_chain_nodes)
      : S_(slots),
        fill_percent_(fill_percent),
        rebuild_chain_nodes_(rebuild_chain_nodes) {}

  ~FlixImpl() override {
    try {
      order_.wait_host();
    } catch (...) {
    }
    release_structure();
  }

  // ----------------------------------------------------------- build
  void build(const void *keys_v, u64 n, u64 max_size, hipStream_t s) override {
    const K *keys = static_cast<const K *>(keys_v);
    if (n && !keys) throw Error(GD_INVALID_ARGUMENT, "null build keys");
    if (n > 0xffffffffull)
      throw Error(GD_INVALID_ARGUMENT, "build size exceeds 2^32-1");
    order_.begin(s);
    try {
      max_size_ = std::max<u64>(max_size ? max_size : 2 * n, n);
      Lease<K> uk(pool_, std::max<u64>(n, 1));
      Lease<V> uv(pool_, std::max<u64>(n, 1));
      u64 m = 0;
      if (n) {
        Lease<u64> offs;
        sort_unique(keys, nullptr, n, /*positions=*/true, nullptr, uk.get(),
                    uv.get(), offs, s);
        GD_HIP(hipMemcpyAsync(&m, offs.get() + n, sizeof(u64),
                              hipMemcpyDeviceToHost, s));
        GD_HIP(hipStreamSynchronize(s));
      }
      layout(uk.get(), uv.get(), m, s);
    } catch (...) {
      order_.end(s);
      throw;
    }
    order_.end(s);
  }

  // ---------------------------------------------------------- updates
  // Insert-if-absent of a batch (impl_cg_rtx_index_updates.cuh:2195).  The
  // batch is sorted and reduced to the keys not yet stored; the plan sizes
  // every touched chain and reserves its new nodes; when the ring cannot
  // hold them the index is rebuilt with a larger allocation region first,
  // so a batch is applied whole or (GD_CAPACITY, GD_OUT_OF_MEMORY) not at
  // all.  With a chain-length threshold set, an insert whose longest chain
  // exceeds it rebuilds the bucket layer afterwards.
  void insert(const void *keys_v, const void *vals_v, u64 n,
              hipStream_t s) override {
    require_built();
    if (!n) return;
    const K *keys = static_cast<const K *>(keys_v);
    const V *vals = static_cast<const V *>(vals_v);
    if (!keys || !vals) throw Error(GD_INVALID_ARGUMENT, "null insert batch");
    if (n > 0xffffffffull)
      throw Error(GD_INVALID_ARGUMENT, "batch size exceeds 2^32-1");
    order_.begin(s);
    try {
      Lease<K> bk(pool_, n);
      Lease<V> bv(pool_, n);
      const Pool<K, V> existing = pool_view();
      // the new keys' count stays on the device (offs[n]): the plan reads
      // it there, and the plan's own read-back is the call's only wait
      Lease<u64> offs;
      sort_unique(keys, vals, n, false, &existing, bk.get(), bv.get(), offs,
                  s);
      apply_insert(bk.get(), bv.get(), offs.get() + n, n, s);
    } catch (...) {
      order_.end(s);
      throw;
    }
    order_.end(s);
  }

  // sorted: the caller's keys are ascending already (the reference's
  // remove(sorted_keys, ...) contract), so they are not sorted again.
  void remove(const void *keys_v, u64 n, bool sorted, hipStream_t s) override {
    require_built();
    if (!n) return;
    const K *keys = static_cast<const K *>(keys_v);
    if (!keys) throw Error(GD_INVALID_ARGUMENT, "null delete batch");
    if (n > 0xffffffffull)
      throw Error(GD_INVALID_ARGUMENT, "batch size exceeds 2^32-1");
    order_.begin(s);
    try {
      const u32 nn = static_cast<u32>(n);
      Lease<K> ka, kb;
      Lease<u32> va, vb, ctrl;
      const K *batch = keys;
      if (!sorted) {
        ka = Lease<K>(pool_, n);
        kb = Lease<K>(pool_, n);
        va = Lease<u32>(pool_, n);
        vb = Lease<u32>(pool_, n);
        ctrl = Lease<u32>(pool_, rs::control_words(n, rs::passes_for<K>()));
        GD_HIP(rs::sort_pairs<K>(keys, nn, ka.get(), va.get(), kb.get(),
                                 vb.get(), ctrl.get(), s));
        batch = ka.get();
      }
      Lease<u64> begin(pool_, u64(P_) + 1);
      const u64 most = std::min<u64>(P_, n);  // touched buckets, at most
      Lease<u32> work(pool_, most + 1);
      const u32 *count = slices(batch, n, begin.get(), work.get(), s);
      // a wave per touched bucket, the grid sized by the upper bound (capped;
      // the waves stride over the list)
      if (S_ <= 32) {  // a half-wave per bucket
        const u64 blocks = std::min<u64>((most + 7) / 8, u64(1) << 20);
        delete_kernel<K, V, 32><<<static_cast<u32>(blocks), 256, 0, s>>>(
            pool_view(), ring_view(), batch, begin.get(), work.get(), count);
      } else {
        const u64 blocks = std::min<u64>((most + 3) / 4, u64(1) << 20);
        delete_kernel<K, V, 64><<<static_cast<u32>(blocks), 256, 0, s>>>(
            pool_view(), ring_view(), batch, begin.get(), work.get(), count);
      }
      GD_LAUNCH_CHECK();
    } catch (...) {
      order_.end(s);
      throw;
    }
    order_.end(s);
  }

  // ------------------------------------------------------------ reads
  void lookup(const void *q, u64 n, void *out, bool sorted,
              hipStream_t s) override {
    require_built();
    if (!n) return;
    if (!q || !out) throw Error(GD_INVALID_ARGUMENT, "null lookup buffers");
    order_.begin(s);
    try {
      if (!sorted && n >= kGroupedLookups && n <= 0xffffffffull) {
        // group the queries by the 16 bits below their top varying one
        // (two stable passes: the lower digit, then the top digit), look
        // them up in grouped order (each XCD on its own share of the
        // buckets, each group's nodes in L2), then bring the results back
        // to query order through both passes' destinations
        const u32 nn = static_cast<u32>(n);
        const K *qk = static_cast<const K *>(q);
        const lsm::SpreadLayout<K> L(n);
        const u64 lo_words = rs::control_words(n, 1);
        Lease<K> qa(pool_, n), qb(pool_, n);
        Lease<u32> dest_lo(pool_, n), dest_hi(pool_, n);
        Lease<V> res(pool_, n);
        Lease<u32> ctrl(pool_, L.words);
        Lease<u32> lo_ctrl(pool_, lo_words + 1);
        u32 *shift_lo = lo_ctrl.get() + lo_words;
        const rs::Control lc = rs::control_layout(lo_ctrl.get(), 1);
        // the top digit's shift, then the lower digit's counts and pass
        lsm::spread_counts<K>(qk, nn, ctrl.get(), s);
        rs::clear_control(lo_ctrl.get(), lo_words, s);
        shift_below_kernel<<<1, 1, 0, s>>>(ctrl.get() + L.shift_at, shift_lo);
        digit_hist_kernel<K><<<rs::hist_blocks_for(n), rs::kThreads, 0, s>>>(
            qk, nn, shift_lo, lc.rows);
        rs::scan_kernel<<<1, rs::kRadix, 0, s>>>(lc.rows, 1, lc.offsets);
        rs::launch_pass<K>(rs::block_for(n), rs::tiles_for(n), qk,
                           rs::PayloadPos{}, qa.get(), nullptr, nn, 0,
                           lc.offsets, lc.status, lc.tickets, s,
                           dest_lo.get(), shift_lo, nullptr);
        // the top digit, stable over the lower one
        lsm::partition_by_spread<K>(qa.get(), nn, qb.get(), nullptr,
                                    dest_hi.get(), rs::PayloadPos{},
                                    ctrl.get(), s, 0);
        lookup_grouped_kernel<K, V><<<lsm::grouped_lookup_blocks(n), 256, 0,
                                      s>>>(pool_view(), qb.get(), n,
                                           res.get());
        lookup_gather2_kernel<V><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(
            dest_lo.get(), dest_hi.get(), res.get(), n,
            static_cast<V *>(out));
      } else {
        lookup_kernel<K, V><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(
            pool_view(), static_cast<const K *>(q), n, static_cast<V *>(out));
      }
      GD_LAUNCH_CHECK();
    } catch (...) {
      order_.end(s);
      throw;
    }
    order_.end(s);
  }

  void successor(const void *q, u64 n, void *out, hipStream_t s) override {
    require_built();
    if (!n) return;
    if (!q || !out) throw Error(GD_INVALID_ARGUMENT, "null successor buffers");
    order_.begin(s);
    successor_kernel<K, V><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(
        pool_view(), static_cast<const K *>(q), n, static_cast<K *>(out));
    GD_LAUNCH_CHECK();
    order_.end(s);
  }

  void range(const void *lo, const void *hi, u64 n, u64 *offsets, void *ok,
             void *ov, hipStream_t s) override {
    require_built();
    if (!n) return;
    if (!lo || !hi || !offsets || (ok && !ov))
      throw Error(GD_INVALID_ARGUMENT, "null range buffers");
    order_.begin(s);
    try {
      const Pool<K, V> pv = pool_view();
      // a wave per query (4 per block)
      const u32 blocks = grid_for(n, 4, 1u << 16);
      if (!ok) {
        // counts from ranks: the keys before each bucket (a scan of the
        // chains' live keys), then two point searches per query
        Lease<u64> sizes(pool_, u64(P_) + 1);
        Lease<u64> bprefix(pool_, u64(P_) + 1);
        widen(bsize_, P_, sizes.get(), s);
        exclusive_scan(sizes.get(), P_, bprefix.get(), s);
        Lease<u64> counts(pool_, n + 1);
        range_count_kernel<K, V><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(
            pv, bprefix.get(), static_cast<const K *>(lo),
            static_cast<const K *>(hi), n, counts.get());
        GD_LAUNCH_CHECK();
        exclusive_scan(counts.get(), n, offsets, s);
      } else {
        range_wave_kernel<K, V><<<blocks, 256, 0, s>>>(
            pv, static_cast<const K *>(lo), static_cast<const K *>(hi), n,
            offsets, static_cast<K *>(ok), static_cast<V *>(ov));
        GD_LAUNCH_CHECK();
      }
    } catch (...) {
      order_.end(s);
      throw;
    }
    order_.end(s);
  }

  // ---------------------------------------------------------- rebuild
  void rebuild(hipStream_t s) override {
    require_built();
    order_.begin(s);
    try {
      rebuild_into(0, s);
    } catch (...) {
      order_.end(s);
      throw;
    }
    order_.end(s);
  }

  u64 total_keys(hipStream_t s) override {
    require_built();
    order_.begin(s);
    u64 total = 0;
    try {
      Lease<unsigned long long> t(pool_, 1);
      rs::zero_u32_kernel<<<1, 64, 0, s>>>(reinterpret_cast<u32 *>(t.get()), 2);
      node_total_kernel<K, V><<<grid_for(P_, 256, 4096), 256, 0, s>>>(
          pool_view(), t.get());
      GD_LAUNCH_CHECK();
      GD_HIP(hipMemcpyAsync(&total, t.get(), sizeof(u64),
                            hipMemcpyDeviceToHost, s));
      GD_HIP(hipStreamSynchronize(s));
    } catch (...) {
      order_.end(s);
      throw;
    }
    order_.end(s);
    return total;
  }

  u64 resident_bytes() const override { return pool_.held_bytes(); }
  void memory(gd_memory_stats *out) const override {
    out->held_bytes = pool_.held_bytes();
    out->cached_bytes = pool_.cached_bytes();
    out->live_bytes = pool_.live_bytes();
    out->peak_held_bytes = pool_.peak_bytes();
    out->trimmed_blocks = pool_.trimmed_blocks();
    out->row_bytes_ub =
        total_nodes_ * (u64(S_) * (sizeof(K) + sizeof(V)) + 2 * sizeof(u32));
  }
  void trim() override {
    order_.wait_host();
    pool_.trim();
  }

 private:
  void require_built() const {
    if (!node_keys_)
      throw Error(GD_INVALID_ARGUMENT, "FliX index used before build");
  }

  Pool<K, V> pool_view() const {
    return Pool<K, V>{node_keys_, node_vals_, node_count_, node_next_,
                      maxv_,      bsize_,     blen_,       S_,
                      P_,         total_nodes_, bdir_,     bdir_shift_};
  }

  Ring ring_view() const {
    return Ring{ring_ids_, ring_cap_, &counters_->alloc_cursor,
                &counters_->free_count};
  }

  void release_structure() {
    pool_.put(node_keys_);
    pool_.put(node_vals_);
    pool_.put(node_count_);
    pool_.put(node_next_);
    pool_.put(maxv_);
    pool_.put(bdir_);
    pool_.put(bsize_);
    pool_.put(blen_);
    pool_.put(ring_ids_);
    pool_.put(counters_);
    node_keys_ = nullptr;
    node_vals_ = nullptr;
    node_count_ = nullptr;
    node_next_ = nullptr;
    maxv_ = nullptr;
    bdir_ = nullptr;
    bsize_ = nullptr;
    blen_ = nullptr;
    ring_ids_ = nullptr;
    counters_ = nullptr;
  }

  // out[0..n) = exclusive scan of in[0..n), out[n] = total.
  void exclusive_scan(const u64 *in, u64 n, u64 *out, hipStream_t s) {
    const u64 tiles = (n + 1 + kScanTile - 1) / kScanTile;
    Lease<u32> ctrl(pool_, 64 + 2 * tiles);
    rs::zero_u32_kernel<<<grid_for(64 + 2 * tiles, 256, 1024), 256, 0, s>>>(
        ctrl.get(), 64 + 2 * tiles);
    scan_u64_kernel<<<static_cast<u32>(tiles), 256, 0, s>>>(
        in, n, out, reinterpret_cast<u64 *>(ctrl.get() + 64), ctrl.get());
    GD_LAUNCH_CHECK();
  }

  // Sorts (key, position), keeps the first occurrence of each key (and, when
  // `existing` is given, only keys not yet stored), compacts into (ok, ov)
  // with values gathered from `vals` (or = input positions).  offs receives
  // n + 1 words, the rows' positions; offs[n] = the number of rows (on the
  // device: nothing here waits for it).
  void sort_unique(const K *keys, const V *vals, u64 n, bool positions,
                   const Pool<K, V> *existing, K *ok, V *ov, Lease<u64> &offs,
                   hipStream_t s) {
    const u32 nn = static_cast<u32>(n);
    Lease<K> ka(pool_, n), kb(pool_, n);
    Lease<u32> va(pool_, n), vb(pool_, n);
    Lease<u32> ctrl(pool_, rs::control_words(n, rs::passes_for<K>()));
    GD_HIP(rs::sort_pairs<K>(keys, nn, ka.get(), va.get(), kb.get(), vb.get(),
                             ctrl.get(), s));
    Lease<u32> keep(pool_, n);
    if (existing) {
      mark_new_kernel<K, V><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(
          *existing, ka.get(), n, keep.get());
    } else {
      mark_first_kernel<K><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(
          ka.get(), n, keep.get());
    }
    GD_LAUNCH_CHECK();
    Lease<u64> keep64(pool_, n + 1);
    widen(keep.get(), n, keep64.get(), s);
    offs = Lease<u64>(pool_, n + 1);
    exclusive_scan(keep64.get(), n, offs.get(), s);
    compact_batch_kernel<K, V><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(
        ka.get(), va.get(), n, keep.get(), offs.get(), vals,
        positions ? 1 : 0, ok, ov);
    GD_LAUNCH_CHECK();
  }

  void widen(const u32 *in, u64 n, u64 *out, hipStream_t s) {
    widen_kernel<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(in, n, out);
    GD_LAUNCH_CHECK();
  }

  // rebuild_bucket_layer_only (impl_cg_rtx_index_updates.cuh:1821-1881):
  // the chains' keys, in key order, are laid out as a fresh bucket layer
  // (each chain writes its rows there directly).  grow_for > 0 (an insert of that many keys that the ring cannot
  // hold) sizes the new allocation region for twice the keys after it.
  void rebuild_into(u64 grow_for, hipStream_t s) {
    Lease<u64> sizes(pool_, u64(P_) + 1);
    Lease<u64> offs(pool_, u64(P_) + 1);
    chain_sizes_kernel<K, V><<<grid_for(P_, 256), 256, 0, s>>>(pool_view(),
                                                               sizes.get());
    GD_LAUNCH_CHECK();
    exclusive_scan(sizes.get(), P_, offs.get(), s);
    u64 total = 0;
    GD_HIP(hipMemcpyAsync(&total, offs.get() + P_, sizeof(u64),
                          hipMemcpyDeviceToHost, s));
    GD_HIP(hipStreamSynchronize(s));
    max_size_ = std::max(max_size_, total);
    if (grow_for) max_size_ = std::max(max_size_, 2 * (total + grow_for));
    relayout(offs.get(), total, s);
  }

  // Applies a sorted batch of keys absent from the index (see insert()):
  // *m_dev of them (device memory), at most n_ub.
  void apply_insert(const K *bk, const V *bv, const u64 *m_dev, u64 n_ub,
                    hipStream_t s) {
    RingCounters rc{};
    for (int attempt = 0;; ++attempt) {
      Lease<u64> begin(pool_, u64(P_) + 1);
      Lease<uint2> fast(pool_, std::min<u64>(n_ub, u64(P_)) + 1);
      Lease<BigItem> big(pool_, std::min<u64>(n_ub, u64(P_)) + 1);
      bucket_slices_kernel<K><<<grid_for(u64(P_) + 1, 256), 256, 0, s>>>(
          maxv_, P_, bk, 0, m_dev, begin.get());
      plan_reset_kernel<<<1, 1, 0, s>>>(&counters_->plan);
      insert_plan_kernel<<<static_cast<u32>((u64(P_) + kPlanBuckets - 1) /
                                            kPlanBuckets),
                           256, 0, s>>>(begin.get(), bsize_, blen_, P_, S_,
                                        fast.get(), big.get(),
                                        &counters_->plan);
      GD_LAUNCH_CHECK();
      GD_HIP(hipMemcpyAsync(&rc, counters_, sizeof(rc), hipMemcpyDeviceToHost,
                            s));
      GD_HIP(hipStreamSynchronize(s));
      const u64 avail = rc.free_count - rc.alloc_cursor;
      if (rc.plan.new_nodes <= avail) {
        launch_apply(bk, bv, begin.get(), fast.get(), big.get(), rc.plan, s);
        break;
      }
      if (attempt > 0)
        throw Error(GD_CAPACITY,
                    "FliX allocation region of " + std::to_string(ring_cap_) +
                        " nodes cannot hold the batch even after a rebuild; "
                        "no key was inserted");
      // grow: rebuild with room for twice the keys after this batch (sized
      // by the batch's length, an upper bound of its new keys)
      rebuild_into(n_ub, s);
    }
    if (rebuild_chain_nodes_ && rc.plan.max_nodes > rebuild_chain_nodes_)
      rebuild_into(0, s);
  }

  void launch_apply(const K *bk, const V *bv, const u64 *begin,
                    const uint2 *fast, const BigItem *big,
                    const InsertPlan &plan, hipStream_t s) {
    const Pool<K, V> pv = pool_view();
    const unsigned long long *cursor = &counters_->alloc_cursor;
    if (plan.n_fast) {
      if (S_ <= 32)  // a half-wave per chain, eight chains per block
        insert_fast_kernel<K, V, 32><<<(plan.n_fast + 7) / 8, 256, 0, s>>>(
            pv, ring_ids_, ring_cap_, cursor, bk, bv, begin, fast,
            plan.n_fast);
      else
        insert_fast_kernel<K, V, 64><<<(plan.n_fast + 3) / 4, 256, 0, s>>>(
            pv, ring_ids_, ring_cap_, cursor, bk, bv, begin, fast,
            plan.n_fast);
      GD_LAUNCH_CHECK();
    }
    if (plan.n_big) {
      // scratch sizes follow the data from batch to batch; rounded up to a
      // power of two they recur, so the pool hands back cached blocks
      // instead of allocating anew
      const u64 nodes = pow2_at_least(u64(plan.big_nodes) + 1);
      const u64 keys = pow2_at_least(plan.big_keys + 1);
      Lease<u32> ids(pool_, nodes);
      Lease<u32> at(pool_, nodes);
      Lease<K> sk(pool_, keys);
      Lease<V> sv(pool_, keys);
      insert_big_kernel<K, V><<<plan.n_big, 256, 0, s>>>(
          pv, ring_ids_, ring_cap_, cursor, bk, bv, begin, big, ids.get(),
          at.get(), sk.get(), sv.get());
      GD_LAUNCH_CHECK();
    }
    ring_advance_kernel<<<1, 1, 0, s>>>(&counters_->alloc_cursor,
                                        &counters_->plan);
    GD_LAUNCH_CHECK();
  }

  // Bucket slices of a sorted batch + compact list of touched buckets.
  // Returns the device word holding the list's length (behind the list).
  const u32 *slices(const K *batch, u64 m, u64 *begin, u32 *work,
                    hipStream_t s) {
    bucket_slices_kernel<K><<<grid_for(u64(P_) + 1, 256), 256, 0, s>>>(
        maxv_, P_, batch, m, nullptr, begin);
    GD_LAUNCH_CHECK();
    u32 *count = work + std::min<u64>(P_, m);
    rs::zero_u32_kernel<<<1, 64, 0, s>>>(count, 1);
    work_list_kernel<<<static_cast<u32>(
                           (u64(P_) + kWorkListBuckets - 1) / kWorkListBuckets),
                       256, 0, s>>>(begin, P_, work, count);
    GD_LAUNCH_CHECK();
    return count;
  }

  // Lays sorted unique (key, value) rows into a fresh bucket layer.
  void layout(const K *uk, const V *uv, u64 m, hipStream_t s) {
    const u32 fill = layer_fill();
    const Layer L = new_layer(m);
    install(L);
    clear_headers(s);
    layout_kernel<K, V><<<grid_for(u64(P_) * S_, 256, 1u << 16), 256, 0, s>>>(
        pool_view(), maxv_, uk, uv, m, fill);
    GD_LAUNCH_CHECK();
    finish_layer(s);
  }

  // Rebuild of the current layer's m live keys into a fresh layer, straight
  // from the chains (offsets = the live keys before each chain) with no
  // flattened copy in between.
  void relayout(const u64 *offsets, u64 m, hipStream_t s) {
    const u32 fill = layer_fill();
    const Pool<K, V> old = pool_view();
    const Layer L = new_layer(m);
    const Pool<K, V> nv{L.keys, L.vals, L.count, L.next, L.maxv,
                        L.bsize, L.blen, S_, static_cast<u32>(L.P), L.total,
                        L.bdir, static_cast<u32>(8 * sizeof(K)) - L.d};
    rs::zero_u32_kernel<<<grid_for(L.total, 256, 1u << 16), 256, 0, s>>>(
        L.count, L.total);
    rs::zero_u32_kernel<<<grid_for(L.total, 256, 1u << 16), 256, 0, s>>>(
        L.next, L.total);
    relayout_kernel<K, V><<<(old.buckets + 3) / 4, 256, 0, s>>>(old, offsets,
                                                                nv, fill);
    GD_LAUNCH_CHECK();
    install(L);  // the old arrays go back to the pool behind the kernel
    layout_heads_kernel<K, V><<<grid_for(u64(P_), 256, 1u << 16), 256, 0, s>>>(
        pool_view(), maxv_, m, fill);
    GD_LAUNCH_CHECK();
    finish_layer(s);
  }

  u32 layer_fill() const {
    return std::max<u32>(1, std::min<u32>(
        S_, static_cast<u32>((u64(S_) * fill_percent_ + 99) / 100)));
  }

  // The arrays of a bucket layer for m keys, allocated but not installed.
  struct Layer {
    K *keys;
    V *vals;
    u32 *count, *next;
    K *maxv;
    u32 *bdir, *bsize, *blen, *ring;
    RingCounters *ctr;
    u64 P, total, region;
    u32 d;
  };

  Layer new_layer(u64 m) {
    const u32 fill = layer_fill();
    Layer L{};
    L.P = std::max<u64>(1, (m + fill - 1) / fill);
    const u64 need = (2 * std::max(max_size_, m) + S_ - 1) / S_;
    L.region = std::max<u64>(need > L.P ? need - L.P : 0, L.P / 4) + 1024;
    L.total = L.P + L.region;
    if (L.total >= 0xffffffffull)
      throw Error(GD_INVALID_ARGUMENT, "FliX node pool exceeds 2^32-1 nodes");
    // bucket directory: 2^d slices, 2^d <= P (at least 2)
    L.d = 1;
    while (L.d < 24 && (u64(1) << (L.d + 1)) <= L.P) ++L.d;
    // a failed allocation returns what was taken so far
    std::vector<void *> got;
    try {
      L.keys = pool_.get_n<K>(L.total * S_);
      got.push_back(L.keys);
      L.vals = pool_.get_n<V>(L.total * S_);
      got.push_back(L.vals);
      L.count = pool_.get_n<u32>(L.total);
      got.push_back(L.count);
      L.next = pool_.get_n<u32>(L.total);
      got.push_back(L.next);
      L.maxv = pool_.get_n<K>(L.P);
      got.push_back(L.maxv);
      L.bdir = pool_.get_n<u32>((u64(1) << L.d) + 1);
      got.push_back(L.bdir);
      L.bsize = pool_.get_n<u32>(L.P);
      got.push_back(L.bsize);
      L.blen = pool_.get_n<u32>(L.P);
      got.push_back(L.blen);
      L.ring = pool_.get_n<u32>(L.region);
      got.push_back(L.ring);
      L.ctr = pool_.get_n<RingCounters>(1);
    } catch (...) {
      for (void *b : got) pool_.put(b);
      throw;
    }
    return L;
  }

  // Makes L the current layer; the previous layer's arrays return to the
  // pool (work enqueued later on the stream runs after every use of them).
  void install(const Layer &L) {
    release_structure();
    node_keys_ = L.keys;
    node_vals_ = L.vals;
    node_count_ = L.count;
    node_next_ = L.next;
    maxv_ = L.maxv;
    bdir_ = L.bdir;
    bdir_d_ = L.d;
    bdir_shift_ = static_cast<u32>(8 * sizeof(K)) - L.d;
    bsize_ = L.bsize;
    blen_ = L.blen;
    ring_ids_ = L.ring;
    counters_ = L.ctr;
    P_ = static_cast<u32>(L.P);
    total_nodes_ = L.total;
    ring_cap_ = L.region;
  }

  // Allocation-region headers start empty.
  void clear_headers(hipStream_t s) {
    rs::zero_u32_kernel<<<grid_for(total_nodes_, 256, 1u << 16), 256, 0, s>>>(
        node_count_, total_nodes_);
    rs::zero_u32_kernel<<<grid_for(total_nodes_, 256, 1u << 16), 256, 0, s>>>(
        node_next_, total_nodes_);
  }

  // Bucket directory and ring of a layer whose heads are written.
  void finish_layer(hipStream_t s) {
    const u64 slices = u64(1) << bdir_d_;
    bucket_dir_clear_kernel<<<grid_for(slices, 256, 1u << 16), 256, 0, s>>>(
        bdir_, slices);
    bucket_dir_kernel<K><<<grid_for(P_, 256), 256, 0, s>>>(
        maxv_, P_, bdir_, bdir_d_, bdir_shift_);
    bucket_dir_fill_kernel<K><<<grid_for(slices, 256, 1u << 16), 256, 0, s>>>(
        maxv_, P_, bdir_, bdir_d_, bdir_shift_);
    ring_init_kernel<<<grid_for(ring_cap_, 256, 1u << 16), 256, 0, s>>>(
        ring_ids_, P_, ring_cap_);
    ring_counters_init_kernel<<<1, 1, 0, s>>>(counters_, ring_cap_);
    GD_LAUNCH_CHECK();
  }

  DevicePool pool_;
  OperationOrder order_;
  u32 S_;
  u32 fill_percent_;
  u32 rebuild_chain_nodes_;  // rebuild after an insert leaves a longer chain
  u64 max_size_ = 0;
  u32 P_ = 0;
  u64 total_nodes_ = 0;
  u64 ring_cap_ = 0;
  K *node_keys_ = nullptr;
  V *node_vals_ = nullptr;
  u32 *node_count_ = nullptr;
  u32 *node_next_ = nullptr;
  K *maxv_ = nullptr;
  u32 *bdir_ = nullptr;  // bucket directory (2^bdir_d_ + 1 entries)
  u32 bdir_d_ = 1, bdir_shift_ = 0;
  u32 *bsize_ = nullptr;  // live keys of each bucket's chain
  u32 *blen_ = nullptr;   // nodes of each bucket's chain
  u32 *ring_ids_ = nullptr;
  RingCounters *counters_ = nullptr;
};

}  // namespace flix
}  // namespace gdict

using gdict::guarded;
using gdict::flix::FlixBase;

// The error header (HandleBase) sits at offset 0 of both handle types, which
// is what gd_last_error_message relies on.
struct gd_flix : gdict::HandleBase {
  std::unique_ptr<FlixBase> impl;
};

extern "C" {

gd_status gd_flix_create(const gd_flix_config *config, gd_flix **out) {
  if (!config || !out) {
    gdict::set_create_error("null config or output handle");
    return GD_INVALID_ARGUMENT;
  }
  *out = nullptr;
  return guarded(nullptr, [&] {
    const uint32_t kb = config->key_bytes, vb = config->value_bytes;
    const uint32_t slots = config->node_slots ? config->node_slots : 32;
    const uint32_t fill = config->fill_percent ? config->fill_percent : 50;
    if ((kb != 4 && kb != 8) || (vb != 4 && vb != 8))
      throw gdict::Error(GD_INVALID_ARGUMENT,
                         "key_bytes and value_bytes must be 4 or 8");
    if (slots < 8 || slots > 64 || fill > 100)
      throw gdict::Error(GD_INVALID_ARGUMENT,
                         "node_slots must be in [8, 64], fill_percent <= 100");
    auto h = std::make_unique<gd_flix>();
    using gdict::u32;
    using gdict::u64;
    using gdict::flix::FlixImpl;
    const uint32_t chain = config->rebuild_chain_nodes;
    if (kb == 4 && vb == 4)
      h->impl.reset(new FlixImpl<u32, u32>(slots, fill, chain));
    else if (kb == 4)
      h->impl.reset(new FlixImpl<u32, u64>(slots, fill, chain));
    else if (vb == 4)
      h->impl.reset(new FlixImpl<u64, u32>(slots, fill, chain));
    else
      h->impl.reset(new FlixImpl<u64, u64>(slots, fill, chain));
    *out = h.release();
  });
}

gd_status gd_flix_destroy(gd_flix *h) {
  delete h;
  return GD_OK;
}

#define GD_FLIX_CALL(h, body)                                   \
  do {                                                          \
    if (!(h) || !(h)->impl) return GD_INVALID_ARGUMENT;         \
    FlixBase *f = (h)->impl.get();                              \
    std::lock_guard<std::mutex> lock__((h)->mutex);             \
    return guarded((h), [&] { body; });                         \
  } while (0)

gd_status gd_flix_build(gd_flix *h, const void *keys, uint64_t count,
                        uint64_t max_size, hipStream_t s) {
  GD_FLIX_CALL(h, f->build(keys, count, max_size, s));
}

gd_status gd_flix_insert(gd_flix *h, const void *keys, const void *values,
                         uint64_t count, hipStream_t s) {
  GD_FLIX_CALL(h, f->insert(keys, values, count, s));
}

gd_status gd_flix_remove(gd_flix *h, const void *keys, uint64_t count,
                         hipStream_t s) {
  GD_FLIX_CALL(h, f->remove(keys, count, false, s));
}

gd_status gd_flix_remove_sorted(gd_flix *h, const void *sorted_keys,
                                uint64_t count, hipStream_t s) {
  GD_FLIX_CALL(h, f->remove(sorted_keys, count, true, s));
}

gd_status gd_flix_lookup_sorted(gd_flix *h, const void *sorted_queries,
                                uint64_t count, void *out_values,
                                hipStream_t s) {
  GD_FLIX_CALL(h, f->lookup(sorted_queries, count, out_values, true, s));
}

gd_status gd_flix_lookup(gd_flix *h, const void *queries, uint64_t count,
                         void *out_values, hipStream_t s) {
  GD_FLIX_CALL(h, f->lookup(queries, count, out_values, false, s));
}

gd_status gd_flix_successor_sorted(gd_flix *h, const void *sorted_queries,
                                   uint64_t count, void *out_keys,
                                   hipStream_t s) {
  GD_FLIX_CALL(h, f->successor(sorted_queries, count, out_keys, s));
}

gd_status gd_flix_rebuild(gd_flix *h, hipStream_t s) {
  GD_FLIX_CALL(h, f->rebuild(s));
}

gd_status gd_flix_range(gd_flix *h, const void *lo, const void *hi,
                        uint64_t count, uint64_t *out_offsets, void *out_keys,
                        void *out_values, hipStream_t s) {
  GD_FLIX_CALL(h, f->range(lo, hi, count, out_offsets, out_keys, out_values,
                           s));
}

gd_status gd_flix_total_keys(gd_flix *h, uint64_t *out_keys, hipStream_t s) {
  if (!out_keys) return GD_INVALID_ARGUMENT;
  GD_FLIX_CALL(h, *out_keys = f->total_keys(s));
}

gd_status gd_flix_resident_bytes(gd_flix *h, uint64_t *out_bytes) {
  if (!out_bytes) return GD_INVALID_ARGUMENT;
  GD_FLIX_CALL(h, *out_bytes = f->resident_bytes());
}

gd_status gd_flix_memory(gd_flix *h, gd_memory_stats *out) {
  if (!out) return GD_INVALID_ARGUMENT;
  GD_FLIX_CALL(h, f->memory(out));
}

gd_status gd_flix_trim(gd_flix *h) { GD_FLIX_CALL(h, f->trim()); }

}  // extern "C"
