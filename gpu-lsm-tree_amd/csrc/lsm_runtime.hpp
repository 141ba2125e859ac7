// Host runtime of the LSM dictionary (class Lsm), shared by the C ABI
// (lsm.hip) and the serving engine (serve.hip).
//
// Replaces class GPULSMOpt (GPULSMOpt.cuh:5599-7805).  The observable
// contract is the reference's; the structure is this library's own:
//   * every admitted tile (<= batch_capacity keys) is sorted on the GPU by a
//     onesweep LDS radix sort (one fused workgroup for <= 4096 keys) and
//     becomes an immutable sorted run, so reads never scan unsorted rows;
//   * runs merge pairwise (merge-path partitioned, streamed through LDS)
//     whenever the older neighbour is no larger than the newer one, which is
//     a binary counter for equal batches: log2(N/b) merges per row and at
//     most log2(N/b) + 1 runs to search;
//   * the reference's epoch of 16 batches survives as the capacity check
//     point (GPULSMOpt.cuh:6971-6972, :7350-7363).
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <unordered_map>
#include <string>
#include <vector>

#include "host_common.hpp"
#include "kway_merge.hpp"
#include "lsm_kernels.hpp"
#include "radix_sort.hpp"
#include "seg_admission.hpp"

namespace gdict {
namespace lsm {

constexpr u64 kMaximumOperationTile = u64(1) << 20;  // GPULSMOpt.cuh:56
constexpr u32 kBatchesPerEpoch = 16;                  // GPULSMOpt.cuh:42
constexpr u64 kMaximumRows = 0xffffffffull;           // GPULSMOpt.cuh:59-60


template <class T>
struct RunT {
  using Row = typename T::Row;
  Row *rows = nullptr;
  u32 *dir = nullptr;  // RunMeta header, then the directory entries
  u64 rows_ub = 0;     // upper bound on the rows (exact after a refresh)
  u32 max_bits = 1;

  RunMeta *meta() const { return reinterpret_cast<RunMeta *>(dir); }
  u32 *entries() const { return dir + kMetaBytes / sizeof(u32); }
  RunView<T> view() const { return {rows, entries(), meta()}; }
};

using rs::zero_two_kernel;

template <class T>
class LsmT : public HandleBase {
 public:
  using K = typename T::K;
  using V = typename T::V;
  using Row = typename T::Row;
  using Run = RunT<T>;

  // GPULSMOpt.cuh:5606-5616 and :255-270
  static u64 batch_capacity_for(const gd_config &c) {
    return std::min<u64>(kMaximumOperationTile,
                         std::max<u64>(1, c.batch_capacity));
  }
  static u64 capacity_for(const gd_config &c) {
    const u64 cap =
        std::max<u64>(c.max_elements, batch_capacity_for(c) * kBatchesPerEpoch);
    if (cap > kMaximumRows)
      throw Error(GD_INVALID_ARGUMENT,
                  "GPULSMOpt capacity exceeds 32-bit key space");
    return cap;
  }

  explicit LsmT(const gd_config &c)
      : batch_capacity_(batch_capacity_for(c)), capacity_(capacity_for(c)) {}

  ~LsmT() {
    try {
      order_.wait_host();
      reclaim(/*wait=*/true);
    } catch (...) {
    }
    clear_profile();
    if (prof_meta_) (void)hipHostFree(prof_meta_);
    for (auto &kv : pending_) (void)hipEventDestroy(kv.second.done);
    pending_.clear();
    defer_ = false;
    for (auto &r : runs_) free_run(r);
    runs_.clear();
  }

  // ------------------------------------------------------------- updates
  void clear(hipStream_t s) {
    order_.begin(s);
    drop_all_runs();
    reset_state();
    order_.end(s);
  }

  void bulk_build(const K *keys, const V *values, u64 n, hipStream_t s) {
    if ((n && (!keys || !values)) || n > kMaximumRows)
      throw Error(GD_INVALID_ARGUMENT, "invalid GPULSMOpt initial input");
    order_.begin(s);
    drop_all_runs();
    reset_state();
    if (n) {
      Run r = sorted_run(keys, values, nullptr, static_cast<u32>(n),
                         kModeInsert, s);
      runs_.insert(runs_.begin(), r);
      const u32 unique = read_count(runs_.front(), s);
      runs_.front().rows_ub = unique;
      if (unique > capacity_) {
        drop_all_runs();
        order_.end(s);
        throw Error(GD_OUT_OF_MEMORY,
                    "bad_alloc: bulk_build of " + std::to_string(unique) +
                        " unique keys exceeds capacity " +
                        std::to_string(capacity_));
      }
    }
    order_.end(s);
  }

  void admit(const K *keys, const V *values, const u8 *flags, u64 n,
             int mode, hipStream_t s) {
    order_.begin(s);
    try {
      admit_tiles(keys, values, flags, n, mode, s);
    } catch (...) {
      order_.end(s);
      throw;
    }
    order_.end(s);
  }

  void admit_tiles(const K *keys, const V *values, const u8 *flags, u64 n,
                   int mode, hipStream_t s) {
    if (failed_) {
      throw Error(GD_PUBLICATION_FAILED,
                  "GPULSMOpt canonical carry capacity exhausted with status " +
                      std::to_string(status_) + "; capacity=" +
                      std::to_string(capacity_) +
                      "; pending updates were preserved");
    }
    if (!n) return;
    if (!keys || (mode != kModeErase && !values) ||
        (mode == kModeMixed && !flags))
      throw Error(GD_INVALID_ARGUMENT, "invalid GPULSMOpt update batch");
    u64 consumed = 0;
    bool incomplete = false;
    while (consumed < n) {
      const u32 tile =
          static_cast<u32>(std::min<u64>(n - consumed, batch_capacity_));
      ProfRec rec{};
      if (profiling_) rec = prof_begin(2, s, tile);
      Run r = sorted_run(keys + consumed, values ? values + consumed : nullptr,
                         flags ? flags + consumed : nullptr, tile, mode, s);
      if (profiling_) prof_end(rec, s);
      runs_.insert(runs_.begin(), r);
      consumed += tile;
      hipStream_t ms = s;
      if (ms_) {
        // merges wait for this admission, admissions never wait for merges
        hipEvent_t admitted;
        GD_HIP(hipEventCreateWithFlags(&admitted, hipEventDisableTiming));
        GD_HIP(hipEventRecord(admitted, s));
        GD_HIP(hipStreamWaitEvent(ms_, admitted, 0));
        (void)hipEventDestroy(admitted);
        ms = ms_;
      }
      rebalance(ms);
      if (++pending_batches_ == kBatchesPerEpoch) {
        pending_batches_ = 0;
        epoch_capacity_check(ms);
      }
      if (consumed < n && failed_) {
        incomplete = true;
        break;
      }
    }
    if (incomplete)
      throw Error(GD_PUBLICATION_FAILED,
                  "GPULSMOpt publication failed while tiling an update "
                  "(capacity=" + std::to_string(capacity_) +
                      "); accepted pending records were preserved");
  }

  // -------------------------------------------------------------- reads
  void lookup(const K *q, u64 n, V *out_values, u8 *out_found,
              hipStream_t s) {
    if (!n) return;
    if (!q || !out_values)
      throw Error(GD_INVALID_ARGUMENT, "invalid GPULSMOpt lookup");
    order_.begin(s);
    try {
      const RunSet<T> set = run_set();
      if (n >= grouped_lookup_min() && n <= 0xffffffffull && !runs_.empty()) {
        // group the queries by their top varying key byte first (one radix
        // pass that also records where each query went), look up in
        // grouped order, then gather the results back to query order
        Lease<K> qk(pool_, n);
        Lease<u32> dest(pool_, n);
        Lease<GroupedResult<T>> res(pool_, n);
        Lease<u32> ctrl(pool_, SpreadLayout<K>(n).words);
        grouped_lookup(q, n, out_values, out_found,
                       LookupScratch{qk.get(), dest.get(), res.get(),
                                     ctrl.get(), 0},
                       set, s, profiling_);
      } else {
        ProfRec rec{};
        if (profiling_) rec = prof_begin(1, s, n);
        lookup_kernel<T><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(
            PlainLookup<T>{q, out_values, out_found}, n, set);
        if (profiling_) prof_end(rec, s);
      }
      GD_LAUNCH_CHECK();
    } catch (...) {
      order_.end(s);
      throw;
    }
    order_.end(s);
  }

  // Device scratch of a grouped lookup of n queries: the grouped keys, each
  // query's grouped position, the results in grouped order and the
  // partition's control words.  carve() lays them out in one caller-owned
  // buffer (parts 256-byte aligned); bytes = the buffer size they need.
  struct LookupScratch {
    K *qk;
    u32 *dest;
    GroupedResult<T> *res;
    u32 *ctrl;
    u64 bytes;

    static LookupScratch carve(char *base, u64 n) {
      auto up = [](u64 b) { return (b + 255) & ~u64(255); };
      LookupScratch L{};
      u64 off = 0;
      L.qk = reinterpret_cast<K *>(base + off);
      off += up(n * sizeof(K));
      L.dest = reinterpret_cast<u32 *>(base + off);
      off += up(n * sizeof(u32));
      L.res = reinterpret_cast<GroupedResult<T> *>(base + off);
      off += up(n * sizeof(GroupedResult<T>));
      L.ctrl = reinterpret_cast<u32 *>(base + off);
      off += up(SpreadLayout<K>(n).words * sizeof(u32));
      L.bytes = off;
      return L;
    }
  };

  // lookup() with the grouped path's buffers taken from `scratch`
  // (scratch_bytes long, 256-byte aligned device memory) instead of the
  // handle's pool; allocates nothing.  Same kernels and results as lookup().
  void lookup_in_scratch(const K *q, u64 n, V *out_values, u8 *out_found,
                         char *scratch, u64 scratch_bytes, hipStream_t s) {
    if (!n) return;
    if (!q || !out_values)
      throw Error(GD_INVALID_ARGUMENT, "invalid GPULSMOpt lookup");
    const bool grouped =
        n >= grouped_lookup_min() && n <= 0xffffffffull && !runs_.empty();
    LookupScratch L{};
    if (grouped) {
      if (!scratch)
        throw Error(GD_INVALID_ARGUMENT, "no lookup scratch buffer");
      L = LookupScratch::carve(scratch, n);
      if (L.bytes > scratch_bytes)
        throw Error(GD_INVALID_ARGUMENT,
                    "a grouped lookup of " + std::to_string(n) +
                        " queries needs " + std::to_string(L.bytes) +
                        " scratch bytes, " + std::to_string(scratch_bytes) +
                        " are available");
    }
    order_.begin(s);
    try {
      const RunSet<T> set = run_set();
      if (grouped) {
        grouped_lookup(q, n, out_values, out_found, L, set, s);
      } else {
        lookup_kernel<T><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(
            PlainLookup<T>{q, out_values, out_found}, n, set);
      }
      GD_LAUNCH_CHECK();
    } catch (...) {
      order_.end(s);
      throw;
    }
    order_.end(s);
  }

  // A large range batch over a stack whose newer runs hold a large share
  // of the rows first merges the stack into one run: a range visits the
  // rows of the newer runs that fall in it (each tested for being the
  // newest version), so with newer rows at 1/8 of the oldest run's or more
  // and at least kRangeCompactQueries ranges, one merge pass (a few
  // hundred microseconds per 10^8 rows) costs less than the visits.  The
  // contents and every answer are unchanged.
  static constexpr u64 kRangeCompactQueries = u64(1) << 14;
  static constexpr u32 kRangeCompactShift = 3;

  bool range_compaction_pays(u64 n) const {
    if (n < kRangeCompactQueries || runs_.size() < 2) return false;
    u64 newer = 0;
    for (std::size_t i = 0; i + 1 < runs_.size(); ++i) newer += runs_[i].rows_ub;
    return newer >= (runs_.back().rows_ub >> kRangeCompactShift);
  }

  // Merges every run into one (tombstones dropped): one k-way pass when the
  // stack fits one, else pairwise.
  void compact_runs(hipStream_t s) {
    if (runs_.size() < 2) return;
    Run merged;
    if (runs_.size() <= static_cast<std::size_t>(kMaxWay) && runs_.size() > 2) {
      merged = merge_many(runs_.data(), static_cast<u32>(runs_.size()), true,
                          s);
      for (auto &r : runs_) free_run(r);
    } else {
      merged = merge_all(runs_, /*consume=*/true, s);
    }
    runs_.clear();
    runs_.push_back(merged);
  }

  void range(const K *lo, const K *hi, u64 n, V *out, bool count_only,
             hipStream_t s) {
    if (!n) return;
    if (!lo || !hi || !out)
      throw Error(GD_INVALID_ARGUMENT, "invalid GPULSMOpt range");
    order_.begin(s);
    if (range_compaction_pays(n)) compact_runs(s);
    const RunSet<T> set = run_set();
    const u64 waves = (n + range_queries_per_wave(set.count) - 1) /
                      range_queries_per_wave(set.count);
    ProfRec rec{};
    if (profiling_) rec = prof_begin(3, s, n);
    range_kernel<T><<<grid_for(waves, kRangeWaves, 1u << 16), kRangeThreads, 0,
                      s>>>(PlainRange<T>{lo, hi, out, count_only ? 1 : 0}, n,
                           set);
    GD_LAUNCH_CHECK();
    if (profiling_) {
      rec.runs = set.count;
      prof_end(rec, s);
    }
    order_.end(s);
  }

  void successor(const K *q, u64 n, K *out, hipStream_t s) {
    if (!n) return;
    if (!q || !out)
      throw Error(GD_INVALID_ARGUMENT, "invalid GPULSMOpt successor");
    order_.begin(s);
    const RunSet<T> set = run_set();
    successor_kernel<T><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(q, n, out,
                                                                    set);
    GD_LAUNCH_CHECK();
    order_.end(s);
  }

  u64 dump(K *out_keys, V *out_vals, u64 capacity, hipStream_t s) {
    order_.begin(s);
    u64 count = 0;
    try {
      if (!runs_.empty()) {
        Run merged = merge_all(runs_, /*consume=*/false, s);
        count = read_count(merged, s);
        if (count <= capacity && count) {
          if (!out_keys || !out_vals) {
            free_run(merged);
            throw Error(GD_INVALID_ARGUMENT, "null dump output");
          }
          decode_run_kernel<T><<<grid_for(count, 256, 1u << 16), 256, 0, s>>>(
              merged.rows, merged.meta(), out_keys, out_vals);
          GD_LAUNCH_CHECK();
        }
        free_run(merged);
        GD_HIP(hipStreamSynchronize(s));
      }
    } catch (...) {
      order_.end(s);
      throw;
    }
    order_.end(s);
    return count;
  }

  void compact(hipStream_t s) {
    order_.begin(s);
    if (!runs_.empty()) {
      Run merged = merge_all(runs_, /*consume=*/true, s);
      runs_.clear();
      runs_.push_back(merged);
    }
    order_.end(s);
  }

  // ------------------------------------------------- snapshot / restore
  // File format (little endian):
  //   header  magic "GDLSMSN1", u32 version (1), u32 key bytes (4 | 8),
  //           u64 live count, u64 digest of the keys, u64 digest of the
  //           values (word-wise FNV-1a over the 64-bit words of each array,
  //           a zero-padded tail word for 32-bit arrays of odd length)
  //   body    keys[count] ascending, then values[count]
  // A snapshot holds the live contents only (tombstones and shadowed
  // versions are resolved), so restoring it is a bulk build.
  struct SnapshotHeader {
    char magic[8];
    u32 version;
    u32 key_bytes;
    u64 count;
    u64 key_digest;
    u64 value_digest;
  };

  u64 save(const char *path, hipStream_t s) {
    if (!path) throw Error(GD_INVALID_ARGUMENT, "null snapshot path");
    order_.begin(s);
    u64 count = 0;
    // the merged copy goes back to the pool on every exit path
    Run merged{};
    bool have = false;
    try {
      if (!runs_.empty()) {
        merged = merge_all(runs_, /*consume=*/false, s);
        have = true;
        count = read_count(merged, s);
      }
      Lease<K> k(pool_, std::max<u64>(count, 1));
      Lease<V> v(pool_, std::max<u64>(count, 1));
      if (count) {
        decode_run_kernel<T><<<grid_for(count, 256, 1u << 16), 256, 0, s>>>(
            merged.rows, merged.meta(), k.get(), v.get());
        GD_LAUNCH_CHECK();
      }
      FileCloser f(std::fopen(path, "wb"));
      if (!f.p)
        throw Error(GD_INVALID_ARGUMENT,
                    std::string("cannot open snapshot for writing: ") + path);
      SnapshotHeader h{};
      std::memcpy(h.magic, "GDLSMSN1", 8);
      h.version = 1;
      h.key_bytes = sizeof(K);
      h.count = count;
      // header first with zero digests, body, then the final header
      write_all(f.p, &h, sizeof(h));
      h.key_digest = stream_out(f.p, k.get(), count, s);
      h.value_digest = stream_out(f.p, v.get(), count, s);
      if (std::fseek(f.p, 0, SEEK_SET) != 0)
        throw Error(GD_INVALID_ARGUMENT, "snapshot seek failed");
      write_all(f.p, &h, sizeof(h));
      // a failed final flush (a full disk) must not report success
      f.close_checked();
      if (have) free_run(merged);
      have = false;
    } catch (...) {
      if (have) free_run(merged);
      order_.end(s);
      throw;
    }
    order_.end(s);
    return count;
  }

  u64 restore(const char *path, hipStream_t s) {
    if (!path) throw Error(GD_INVALID_ARGUMENT, "null snapshot path");
    FileCloser f(std::fopen(path, "rb"));
    if (!f.p)
      throw Error(GD_INVALID_ARGUMENT,
                  std::string("cannot open snapshot: ") + path);
    SnapshotHeader h{};
    if (std::fread(&h, sizeof(h), 1, f.p) != 1 ||
        std::memcmp(h.magic, "GDLSMSN1", 8) != 0 || h.version != 1)
      throw Error(GD_INVALID_ARGUMENT, "not a gdict LSM snapshot");
    if (h.key_bytes != sizeof(K))
      throw Error(GD_INVALID_ARGUMENT,
                  "snapshot key width " + std::to_string(h.key_bytes) +
                      " does not match this dictionary");
    if (h.count > capacity_)
      throw Error(GD_OUT_OF_MEMORY,
                  "bad_alloc: snapshot of " + std::to_string(h.count) +
                      " keys exceeds capacity " + std::to_string(capacity_));
    Lease<K> k(pool_, std::max<u64>(h.count, 1));
    Lease<V> v(pool_, std::max<u64>(h.count, 1));
    order_.begin(s);
    try {
      if (stream_in(f.p, k.get(), h.count, s) != h.key_digest ||
          stream_in(f.p, v.get(), h.count, s) != h.value_digest)
        throw Error(GD_INVALID_ARGUMENT, "snapshot digest mismatch");
    } catch (...) {
      order_.end(s);
      throw;
    }
    order_.end(s);
    bulk_build(k.get(), v.get(), h.count, s);
    return h.count;
  }

  // ----------------------------------------------- serving-engine access
  // The serving engine owns its dictionary exclusively and orders work with
  // its own streams and events, so these entry points skip the per-call
  // event chaining.  Updates run on one stream while earlier queries may
  // still read the previous runs on another: runs released by updates are
  // kept (deferred) until an event recorded on the query stream completes.
  void admit_on(const K *keys, const V *values, const u8 *flags, u64 n,
                int mode, hipStream_t s) {
    admit_tiles(keys, values, flags, n, mode, s);
  }

  RunSet<T> snapshot() const { return run_set(); }
  bool empty() const { return runs_.empty(); }

  // Serving mode.  Admissions stay on the caller's stream; the merges that
  // rebalance the stack run on `merge_stream` (after an event of the
  // admission), and queries on `query_stream`.  A query snapshot reads the
  // two inputs of every merge that has not completed instead of its output
  // (snapshot_for), so queries never wait for merges.  Released runs stay
  // allocated until (1) every merge enqueued before their release has
  // completed, after which no new snapshot can name them, and then (2) every
  // query enqueued up to that moment has completed.  Merges allocate from
  // their own pool: a pool hands out buffers whose last use was enqueued on
  // its own stream only.
  void set_async(hipStream_t merge_stream, hipStream_t query_stream) {
    ms_ = merge_stream;
    qs_ = query_stream;
    defer_ = true;
  }

  // Queries wait for the merges enqueued so far when their inputs hold at
  // most this many rows (a wait of well under a millisecond), and read the
  // pending merges' inputs otherwise.  Measured on MI355X: a query launch
  // over the 9-12 runs of expanded pending merges cost more than the wait
  // (uneven requests of 32 to 2^18 keys, Zipf(0.99) over 2^30, on a 2^26-key
  // index: 2.1x the saturated throughput with waiting).
  static constexpr u64 kQueryWaitRows = u64(1) << 26;

  // ------------------------------------------ serving-engine step graphs
  // A step whose updates fit one fused small-batch admission runs as a
  // captured graph: the engine takes the run's buffers here, the graph's
  // admission kernel fills them, and adopt_run then places the run on top
  // of the stack exactly as admit_tiles does after sorted_run.
  u32 small_batch_limit() const { return tuning_.small_batch_max; }
  u32 dir_rows_log() const { return tuning_.dir_rows_log; }
  Run alloc_small_run(u32 n) { return alloc_run(n); }
  void release_run(Run &r) { free_run(r); }

  void adopt_run(const Run &r, hipStream_t s) {
    runs_.insert(runs_.begin(), r);
    hipEvent_t admitted;
    GD_HIP(hipEventCreateWithFlags(&admitted, hipEventDisableTiming));
    GD_HIP(hipEventRecord(admitted, s));
    GD_HIP(hipStreamWaitEvent(ms_, admitted, 0));
    (void)hipEventDestroy(admitted);
    rebalance(ms_);
    if (++pending_batches_ == kBatchesPerEpoch) {
      pending_batches_ = 0;
      epoch_capacity_check(ms_);
    }
  }

  // The runs a step's queries read, without any wait: `newest` (a run the
  // step's own graph admits, nullable) first, then the stack with pending
  // merges read as their inputs.  False when they exceed a snapshot.
  bool snapshot_nowait(RunSet<T> &set, const Run *newest) {
    set = RunSet<T>{};
    bool fits = true;
    auto push = [&](const Run &r) {
      if (set.count < static_cast<u32>(kMaxRuns)) set.run[set.count] = r.view();
      else fits = false;
      ++set.count;
    };
    if (newest) push(*newest);
    std::function<void(const Run &)> add = [&](const Run &r) {
      auto it = pending_.find(r.rows);
      if (it != pending_.end()) {
        const hipError_t q = hipEventQuery(it->second.done);
        if (q == hipSuccess) {
          (void)hipEventDestroy(it->second.done);
          pending_.erase(it);
        } else {
          if (q != hipErrorNotReady) GD_HIP(q);
          const std::vector<Run> in = it->second.in;
          for (const Run &x : in) add(x);
          return;
        }
      }
      push(r);
    };
    for (const Run &r : runs_) add(r);
    return fits;
  }

  // Serving mode: merges the whole stack into one run on the merge stream,
  // after the work enqueued on `after` so far (a step's admission), as a
  // merge of the stack's own (readers of older snapshots keep its inputs).
  void compact_async(hipStream_t after) {
    if (runs_.size() < 2) return;
    hipEvent_t admitted;
    GD_HIP(hipEventCreateWithFlags(&admitted, hipEventDisableTiming));
    GD_HIP(hipEventRecord(admitted, after));
    GD_HIP(hipStreamWaitEvent(ms_, admitted, 0));
    (void)hipEventDestroy(admitted);
    PendingMerge pm{nullptr, runs_};
    compact_runs(ms_);
    GD_HIP(hipEventCreateWithFlags(&pm.done, hipEventDisableTiming));
    GD_HIP(hipEventRecord(pm.done, ms_));
    auto old = pending_.find(runs_.front().rows);  // a stale, completed record
    if (old != pending_.end()) {
      (void)hipEventDestroy(old->second.done);
      pending_.erase(old);
    }
    pending_.emplace(runs_.front().rows, pm);
  }

  // wait_merges: the queries wait for every merge enqueued so far and read
  // the stack as it is (used after compact_async).
  RunSet<T> snapshot_for(hipStream_t qs, bool wait_merges = false) {
    if (wait_merges) {
      hipEvent_t merged;
      GD_HIP(hipEventCreateWithFlags(&merged, hipEventDisableTiming));
      GD_HIP(hipEventRecord(merged, ms_));
      GD_HIP(hipStreamWaitEvent(qs, merged, 0));
      (void)hipEventDestroy(merged);
      return run_set();
    }
    u64 backlog = 0;
    for (auto it = pending_.begin(); it != pending_.end();) {
      const hipError_t q = hipEventQuery(it->second.done);
      if (q == hipSuccess) {
        (void)hipEventDestroy(it->second.done);
        it = pending_.erase(it);
        continue;
      }
      if (q != hipErrorNotReady) GD_HIP(q);
      for (const Run &r : it->second.in) backlog += r.rows_ub;
      ++it;
    }
    if (backlog && backlog <= kQueryWaitRows) {
      hipEvent_t merged;
      GD_HIP(hipEventCreateWithFlags(&merged, hipEventDisableTiming));
      GD_HIP(hipEventRecord(merged, ms_));
      GD_HIP(hipStreamWaitEvent(qs, merged, 0));
      (void)hipEventDestroy(merged);
      return run_set();
    }
    RunSet<T> set{};
    bool too_many = false;
    // newest first: a pending merge reads as its inputs, newest first
    std::function<void(const Run &)> add = [&](const Run &r) {
      auto it = pending_.find(r.rows);
      if (it != pending_.end()) {
        const hipError_t q = hipEventQuery(it->second.done);
        if (q == hipSuccess) {
          (void)hipEventDestroy(it->second.done);
          pending_.erase(it);
        } else {
          if (q != hipErrorNotReady) GD_HIP(q);
          const std::vector<Run> in = it->second.in;
          for (const Run &r : in) add(r);
          return;
        }
      }
      if (set.count < static_cast<u32>(kMaxRuns)) set.run[set.count] = r.view();
      else too_many = true;
      ++set.count;
    };
    for (const Run &r : runs_) add(r);
    if (too_many) {
      // more runs than a snapshot holds: wait for the merges instead
      hipEvent_t merged;
      GD_HIP(hipEventCreateWithFlags(&merged, hipEventDisableTiming));
      GD_HIP(hipEventRecord(merged, ms_));
      GD_HIP(hipStreamWaitEvent(qs, merged, 0));
      (void)hipEventDestroy(merged);
      return run_set();
    }
    return set;
  }

  // Ends serving mode: completes every retirement while the streams still
  // exist (call after synchronising them).
  void detach_streams() {
    reclaim(/*wait=*/true);
    for (auto &kv : pending_) (void)hipEventDestroy(kv.second.done);
    pending_.clear();
    ms_ = qs_ = nullptr;
    defer_ = false;
  }

  // Moves the runs released so far into a retirement (serving mode).
  void retire_pending() {
    if (pending_free_.empty()) return;
    Deferred d;
    GD_HIP(hipEventCreateWithFlags(&d.merges_done, hipEventDisableTiming));
    GD_HIP(hipEventRecord(d.merges_done, ms_));
    d.buffers.swap(pending_free_);
    deferred_.push_back(std::move(d));
  }

  // Returns the buffers of completed retirements to their pools, and drops
  // the records of completed merges.
  void reclaim(bool wait = false) {
    while (!deferred_.empty()) {
      Deferred &d = deferred_.front();
      if (d.merges_done) {
        if (wait) {
          GD_HIP(hipEventSynchronize(d.merges_done));
        } else {
          const hipError_t q = hipEventQuery(d.merges_done);
          if (q == hipErrorNotReady) break;
          GD_HIP(q);
        }
        (void)hipEventDestroy(d.merges_done);
        d.merges_done = nullptr;
        GD_HIP(hipEventCreateWithFlags(&d.queries_done,
                                       hipEventDisableTiming));
        GD_HIP(hipEventRecord(d.queries_done, qs_));
      }
      if (wait) {
        GD_HIP(hipEventSynchronize(d.queries_done));
      } else {
        const hipError_t q = hipEventQuery(d.queries_done);
        if (q == hipErrorNotReady) break;
        GD_HIP(q);
      }
      for (void *p : d.buffers) {
        pool_.put(p);
        merge_pool_.put(p);  // each buffer belongs to exactly one of them
      }
      (void)hipEventDestroy(d.queries_done);
      deferred_.pop_front();
    }
    for (auto it = pending_.begin(); it != pending_.end();) {
      if (hipEventQuery(it->second.done) == hipSuccess) {
        (void)hipEventDestroy(it->second.done);
        it = pending_.erase(it);
      } else {
        ++it;
      }
    }
  }


  bool failed() const { return failed_; }

  void set_tuning(const gd_lsm_tuning &t) {
    gd_lsm_tuning v = tuning_;
    if (t.dir_rows_log) v.dir_rows_log = std::min<u32>(t.dir_rows_log, 8);
    if (t.merge_tile_log) {
      if (t.merge_tile_log != 11 && t.merge_tile_log != 12)
        throw Error(GD_INVALID_ARGUMENT, "merge_tile_log must be 11 or 12");
      v.merge_tile_log = t.merge_tile_log;
    }
    if (t.sort_block) {
      if (t.sort_block != 256 && t.sort_block != 512 && t.sort_block != 1024)
        throw Error(GD_INVALID_ARGUMENT,
                    "sort_block must be 256, 512 or 1024");
      v.sort_block = t.sort_block;
    }
    if (t.grouped_lookup_min_log)
      v.grouped_lookup_min_log = std::min<u32>(t.grouped_lookup_min_log, 63);
    if (t.small_batch_max)
      v.small_batch_max = std::min<u32>(t.small_batch_max, T::kSmallTile);
    if (t.max_runs)  // max_runs
      v.max_runs = std::max<u32>(2, std::min<u32>(t.max_runs,
                                                     kMaxRuns - 1));
    if (t.segmented_admission) {
      if (t.segmented_admission > 2)
        throw Error(GD_INVALID_ARGUMENT, "segmented_admission must be 1 or 2");
      v.segmented_admission = t.segmented_admission;
    }
    if (t.merge_block) {
      if (t.merge_block != 256 && t.merge_block != 512)
        throw Error(GD_INVALID_ARGUMENT, "merge_block must be 256 or 512");
      v.merge_block = t.merge_block;
    }
    if (t.merge_ways) {
      if (t.merge_ways < 2 || t.merge_ways > static_cast<u32>(kMaxWay))
        throw Error(GD_INVALID_ARGUMENT, "merge_ways must be 2..8");
      v.merge_ways = t.merge_ways;
    }
    if (t.merge_fanout) {
      if (t.merge_fanout < 2 || t.merge_fanout > static_cast<u32>(kMaxWay))
        throw Error(GD_INVALID_ARGUMENT, "merge_fanout must be 2..8");
      v.merge_fanout = t.merge_fanout;
    }
    tuning_ = v;
  }
  const gd_lsm_tuning &tuning() const { return tuning_; }
  u64 grouped_lookup_min() const {
    return tuning_.grouped_lookup_min_log >= 63
               ? ~0ull
               : (u64(1) << tuning_.grouped_lookup_min_log);
  }

  // ------------------------------------------------------ kernel profile
  // Opt-in: HIP events on the launch stream around every merge and lookup
  // kernel, and the merge's row counts (copied to pinned memory behind the
  // kernel) for its algorithmic bytes (S + M) * 8.
  void set_profiling(bool on) {
    if (on && !prof_meta_) {
      GD_HIP(hipHostMalloc(reinterpret_cast<void **>(&prof_meta_),
                           kProfSlots * kProfMetas * sizeof(RunMeta),
                           hipHostMallocDefault));
    }
    profiling_ = on;
  }

  void read_profile(gd_lsm_profile *out, bool reset) {
    order_.wait_host();
    *out = gd_lsm_profile{};
    for (std::size_t i = 0; i < prof_.size(); ++i) {
      ProfRec &r = prof_[i];
      GD_HIP(hipEventSynchronize(r.end));
      float ms = 0;
      GD_HIP(hipEventElapsedTime(&ms, r.start, r.end));
      if (r.kind == 0) {
        const RunMeta *m = prof_meta_ + kProfMetas * (r.slot % kProfSlots);
        out->merge_launches += 1;
        out->merge_ms += ms;
        for (u64 i = 0; i < r.n; ++i) out->merge_rows_in += m[i].count;
        out->merge_rows_out += m[kMaxWay].count;
      } else if (r.kind == 1) {
        out->lookup_launches += 1;
        out->lookup_ms += ms;
        out->lookup_queries += r.n;
      } else if (r.kind == 3) {
        out->range_launches += 1;
        out->range_ms += ms;
        out->range_queries += r.n;
        out->range_runs += r.runs;
      } else {
        out->admission_batches += 1;
        out->admission_ms += ms;
        out->admission_keys += r.n;
      }
    }
    if (reset) clear_profile();
  }

  u32 carry_status() const { return status_; }
  u64 resident_bytes() const {
    return pool_.held_bytes() + merge_pool_.held_bytes();
  }
  void memory(gd_memory_stats *out) const {
    out->held_bytes = resident_bytes();
    out->cached_bytes = pool_.cached_bytes() + merge_pool_.cached_bytes();
    out->live_bytes = out->held_bytes - out->cached_bytes;
    out->peak_held_bytes = usage_.peak;
    out->trimmed_blocks = pool_.trimmed_blocks() + merge_pool_.trimmed_blocks();
    u64 rows = 0;
    for (const auto &r : runs_) rows += r.rows_ub;
    out->row_bytes_ub = rows * sizeof(Row);
  }
  // Returns every cached block to HIP (waits for the device).
  void trim() {
    order_.wait_host();
    reclaim(/*wait=*/true);
    pool_.trim();
    merge_pool_.trim();
  }
  // Synchronous: exact rows per run (refreshes the host upper bounds).
  void shape(u32 *runs, u64 *rows) {
    order_.wait_host();
    *runs = static_cast<u32>(runs_.size());
    u64 t = 0;
    for (auto &r : runs_) {
      RunMeta m{};
      GD_HIP(hipMemcpy(&m, r.meta(), sizeof(m), hipMemcpyDeviceToHost));
      r.rows_ub = m.count;
      t += m.count;
    }
    *rows = t;
  }

 private:
  // ----------------------------------------------------------- helpers
  // The grouped lookup path: one spread partition of the queries (which
  // also records each query's grouped position), the lookups in grouped
  // order, then the gather back to query order.  The kernel profile, when
  // on, times the lookup kernel (lookup() only: its records use pinned
  // host memory).
  void grouped_lookup(const K *q, u64 n, V *out_values, u8 *out_found,
                      const LookupScratch &L, const RunSet<T> &set,
                      hipStream_t s, bool profile = false) {
    partition_by_spread<K>(q, static_cast<u32>(n), L.qk, nullptr, L.dest,
                           rs::PayloadPos{}, L.ctrl, s, tuning_.sort_block);
    ProfRec rec{};
    if (profile) rec = prof_begin(1, s, n);
    lookup_grouped_kernel<T><<<grouped_lookup_blocks(n), 256, 0, s>>>(
        L.qk, n, L.res, set);
    if (profile) prof_end(rec, s);
    GD_LAUNCH_CHECK();
    lookup_gather_kernel<T><<<grouped_lookup_blocks(n), 256, 0, s>>>(
        PlainLookup<T>{q, out_values, out_found}, L.dest, n, L.res);
  }

  struct FileCloser {
    std::FILE *p;
    explicit FileCloser(std::FILE *f) : p(f) {}
    // Flushes and closes, throwing when either fails (written files).
    void close_checked() {
      std::FILE *f = p;
      p = nullptr;
      const bool flushed = std::fflush(f) == 0;
      const bool closed = std::fclose(f) == 0;
      if (!flushed || !closed)
        throw Error(GD_INVALID_ARGUMENT,
                    "snapshot write failed at flush/close");
    }
    ~FileCloser() {
      if (p) std::fclose(p);
    }
  };

  static void write_all(std::FILE *f, const void *data, std::size_t bytes) {
    if (bytes && std::fwrite(data, 1, bytes, f) != bytes)
      throw Error(GD_INVALID_ARGUMENT, "snapshot write failed");
  }

  // Word-wise FNV-1a over 64-bit words (a zero-padded last word).
  static u64 digest_words(u64 h, const void *data, std::size_t bytes) {
    const unsigned char *b = static_cast<const unsigned char *>(data);
    std::size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
      u64 w;
      std::memcpy(&w, b + i, 8);
      h = (h ^ w) * 0x100000001b3ull;
    }
    if (i < bytes) {
      u64 w = 0;
      std::memcpy(&w, b + i, bytes - i);
      h = (h ^ w) * 0x100000001b3ull;
    }
    return h;
  }

  static constexpr std::size_t kChunkBytes = std::size_t(64) << 20;

  struct PinnedChunk {
    void *p = nullptr;
    PinnedChunk() { GD_HIP(hipHostMalloc(&p, kChunkBytes, hipHostMallocDefault)); }
    ~PinnedChunk() {
      if (p) (void)hipHostFree(p);
    }
  };

  // Device array -> file through a pinned chunk; returns the digest.
  template <class E>
  u64 stream_out(std::FILE *f, const E *dev, u64 n, hipStream_t s) {
    u64 h = 0xcbf29ce484222325ull;
    if (!n) return h;
    PinnedChunk buf;
    const u64 total = n * sizeof(E);
    for (u64 off = 0; off < total; off += kChunkBytes) {
      const std::size_t len = static_cast<std::size_t>(
          std::min<u64>(kChunkBytes, total - off));
      GD_HIP(hipMemcpyAsync(buf.p, reinterpret_cast<const char *>(dev) + off,
                            len, hipMemcpyDeviceToHost, s));
      GD_HIP(hipStreamSynchronize(s));
      h = digest_words(h, buf.p, len);
      write_all(f, buf.p, len);
    }
    return h;
  }

  // File -> device array through a pinned chunk; returns the digest.
  template <class E>
  u64 stream_in(std::FILE *f, E *dev, u64 n, hipStream_t s) {
    u64 h = 0xcbf29ce484222325ull;
    if (!n) return h;
    PinnedChunk buf;
    const u64 total = n * sizeof(E);
    for (u64 off = 0; off < total; off += kChunkBytes) {
      const std::size_t len = static_cast<std::size_t>(
          std::min<u64>(kChunkBytes, total - off));
      if (std::fread(buf.p, 1, len, f) != len)
        throw Error(GD_INVALID_ARGUMENT, "snapshot is truncated");
      h = digest_words(h, buf.p, len);
      GD_HIP(hipMemcpyAsync(reinterpret_cast<char *>(dev) + off, buf.p, len,
                            hipMemcpyHostToDevice, s));
      GD_HIP(hipStreamSynchronize(s));  // the chunk is reused
    }
    return h;
  }

  void reset_state() {
    pending_batches_ = 0;
    failed_ = false;
    status_ = 0;
  }

  // The pool for work on stream s (merges in serving mode have their own).
  DevicePool &pool_for(hipStream_t s) {
    return (ms_ && s == ms_) ? merge_pool_ : pool_;
  }

  Run alloc_run(u64 rows_ub, DevicePool &pool) {
    Run r;
    r.rows_ub = rows_ub;
    r.max_bits = dir_bits_for(rows_ub, kMaxDirBits, tuning_.dir_rows_log);
    // one spare row: pair loads of a window's last row stay in bounds
    r.rows = pool.get_n<Row>(rows_ub + 1);
    r.dir = static_cast<u32 *>(
        pool.get(kMetaBytes + ((u64(1) << r.max_bits) + 1) * sizeof(u32)));
    return r;
  }
  Run alloc_run(u64 rows_ub) { return alloc_run(rows_ub, pool_); }

  void free_run(Run &r) {
    if (defer_) {
      pending_free_.push_back(r.rows);
      pending_free_.push_back(r.dir);
    } else {
      pool_.put(r.rows);
      pool_.put(r.dir);
    }
    r.rows = nullptr;
    r.dir = nullptr;
  }

  void drop_all_runs() {
    for (auto &r : runs_) free_run(r);
    runs_.clear();
  }

  void launch_directory(const Run &r, hipStream_t s) {
    const u64 items = std::max<u64>(r.rows_ub, (u64(1) << r.max_bits) + 1);
    build_directory_kernel<T><<<grid_for(items, 256, 8192), 256, 0, s>>>(
        r.rows, r.meta(), r.entries(), r.max_bits, tuning_.dir_rows_log);
    GD_LAUNCH_CHECK();
  }

  // Sorts one tile of updates into a new run.
  Run sorted_run(const K *keys, const V *values, const u8 *flags, u32 n,
                 int mode, hipStream_t s) {
    Run r = alloc_run(n);
    if (n <= tuning_.small_batch_max) {
      // one workgroup sorts, deduplicates and indexes the whole batch
      small_batch_run_kernel<T><<<1, kSmallThreads, 0, s>>>(
          keys, n, values, flags, mode, r.rows, r.meta(), r.entries(),
          r.max_bits, tuning_.dir_rows_log);
      GD_LAUNCH_CHECK();
      return r;
    }
    if constexpr (sizeof(K) == 4) {
      if (tuning_.segmented_admission == 1 && n > 16384)
        return segmented_run(keys, values, flags, n, mode, s, r);
    }
    const u32 fin_tiles = (n + kFinTile - 1) / kFinTile;
    const u64 sort_words = rs::control_words(n, rs::passes_for<K>());
    const u64 fin_words = 64 + 2ull * fin_tiles;  // ticket pad + status u64
    Lease<K> ka(pool_, n), kb(pool_, n);
    Lease<u32> va(pool_, n), vb(pool_, n);
    Lease<u32> ctrl(pool_, sort_words + fin_words);
    u32 *fin_ctrl = ctrl.get() + sort_words;
    // the sort's clearing launch also clears the finalize state behind its
    // own and the run header
    GD_HIP(rs::sort_pairs<K>(keys, n, ka.get(), va.get(), kb.get(), vb.get(),
                             ctrl.get(), s, tuning_.sort_block,
                             sort_words + fin_words,
                             reinterpret_cast<u32 *>(r.meta()), kMetaWords));
    finalize_batch_kernel<T><<<fin_tiles, kFinThreads, 0, s>>>(
        ka.get(), va.get(), n, values, flags, mode, r.rows, r.meta(),
        r.entries(), r.max_bits, reinterpret_cast<u64 *>(fin_ctrl + 64),
        fin_ctrl);
    GD_LAUNCH_CHECK();
    return r;
  }

  // 32-bit batches: partition by the top varying key byte, then one
  // workgroup per segment sorts, dedupes, encodes and indexes it.
  Run segmented_run(const K *keys, const V *values, const u8 *flags, u32 n,
                    int mode, hipStream_t s, Run r) {
    // [spread partition control | status (256 x u64) | ticket, pad]
    const SpreadLayout<K> L(n);
    const u64 status_at = L.words + (L.words & 1);
    const u64 ticket_at = status_at + 2 * 256;
    const u64 words = ticket_at + 2;
    Lease<u32> ctrl(pool_, words);
    Lease<u32> ka(pool_, n), va(pool_, n), kb(pool_, n), vb(pool_, n);
    u32 *c = ctrl.get();
    const rs::Control pc = rs::control_layout(c, 1);
    // per-tile digit counts and spreads; the same launch clears the
    // partition's state, this kernel's state behind it (all but the
    // spans) and the run header.  Per-tile digit prefixes at the batch's
    // shift replace the partition's look-back.
    {
      K *span = reinterpret_cast<K *>(c + L.span_at);
      const u32 tiles = rs::tiles_for(n);
      Lease<u32> tile_hist(pool_, 2 * u64(tiles) * rs::kRadix);
      u32 *tile_prefix = tile_hist.get() + u64(tiles) * rs::kRadix;
      seg_tile_hist_kernel<K><<<tiles, rs::kThreads, 0, s>>>(
          keys, n, tile_hist.get(), span, c, L.span_at, c + L.shift_at,
          words - L.shift_at, reinterpret_cast<u32 *>(r.meta()), kMetaWords);
      seg_tile_scan_kernel<K><<<rs::kRadix, rs::kThreads, 0, s>>>(
          keys, n, tile_hist.get(), span, tiles, tile_prefix, pc.offsets,
          c + L.shift_at);
      rs::launch_pass<u32>(
          tuning_.sort_block ? tuning_.sort_block : rs::block_for(n), tiles,
          keys, SegPayload{values, flags, mode}, ka.get(), va.get(), n, 0,
          pc.offsets, pc.status, pc.tickets, s, nullptr, c + L.shift_at,
          tile_prefix);
    }
    seg_sort_kernel<T><<<256, kSegThreads, 0, s>>>(
        keys, ka.get(), va.get(), kb.get(), vb.get(), n, pc.offsets,
        c + L.shift_at, values, flags, mode, r.rows, r.meta(), r.entries(),
        r.max_bits, reinterpret_cast<u64 *>(c + status_at), c + ticket_at);
    GD_LAUNCH_CHECK();
    return r;
  }

  // Merges a (newer) with b (older) into a new run; inputs stay owned by
  // the caller.
  Run merge_two(const Run &a, const Run &b, bool drop_tombstones,
                hipStream_t s) {
    const u64 ub = a.rows_ub + b.rows_ub;
    DevicePool &pool = pool_for(s);
    Run out = alloc_run(ub, pool);
    const u32 tile_rows =
        (tuning_.merge_tile_log == 11 || T::kMergeItemsMax < 16) ? 2048u
                                                                  : 4096u;
    const u32 tiles =
        static_cast<u32>(std::max<u64>(1, (ub + tile_rows - 1) / tile_rows));
    const u64 ctrl_words = 64 + 2ull * tiles;
    Lease<u32> parts(pool, u64(tiles) + 1);
    Lease<u32> ctrl(pool, ctrl_words);
    u64 *status = reinterpret_cast<u64 *>(ctrl.get() + 64);
    const MergeSide<T> sa{a.rows, a.meta()};
    const MergeSide<T> sb{b.rows, b.meta()};
    merge_partition_kernel<T><<<(tiles + 4) / 4, 256, 0, s>>>(
        sa, sb, tiles, tile_rows, parts.get(), status, ctrl.get(),
        out.meta(), out.max_bits);
    GD_LAUNCH_CHECK();
    ProfRec rec{};
    if (profiling_) rec = prof_begin(0, s, 2);
    // tile = threads x items: 4096 rows as 256 x 16 or 512 x 8, 2048 rows
    // as 256 x 8 or 512 x 4
    const bool wide = tuning_.merge_block == 512;
    auto launch = [&](auto kernel, int threads) {
      kernel<<<tiles, threads, 0, s>>>(sa, sb, parts.get(), out.rows,
                                       out.meta(), out.entries(), out.max_bits,
                                       drop_tombstones ? 1 : 0, status,
                                       ctrl.get());
    };
    if (tile_rows == 4096) {
      if constexpr (T::kMergeItemsMax >= 16) {
        if (wide) launch(merge_kernel<T, 8, 512>, 512);
        else launch(merge_kernel<T, 16, 256>, 256);
      }
    } else {
      if (wide) launch(merge_kernel<T, 4, 512>, 512);
      else launch(merge_kernel<T, 8, 256>, 256);
    }
    GD_LAUNCH_CHECK();
    if (profiling_) {
      prof_end(rec, s);
      const Run in[2] = {a, b};
      prof_metas(rec.slot, in, 2, out, s);
    }
    return out;
  }

  // Copies the merge's input and output headers behind it (profiling).
  void prof_metas(u64 rec_slot, const Run *in, u32 ways, const Run &out,
                  hipStream_t s) {
    RunMeta *slot = prof_meta_ + kProfMetas * (rec_slot % kProfSlots);
    for (u32 i = 0; i < ways; ++i)
      GD_HIP(hipMemcpyAsync(slot + i, in[i].meta(), sizeof(RunMeta),
                            hipMemcpyDeviceToHost, s));
    GD_HIP(hipMemcpyAsync(slot + kMaxWay, out.meta(), sizeof(RunMeta),
                          hipMemcpyDeviceToHost, s));
  }

  // Merges in[0..ways) (newest first, 3 <= ways <= kMaxWay) into a new run
  // in one pass; inputs stay owned by the caller.
  Run merge_many(const Run *in, u32 ways, bool drop_tombstones,
                 hipStream_t s) {
    using Shape = KWayShape<T>;
    u64 ub = 0;
    KWayIn<T> kin{};
    for (u32 i = 0; i < ways; ++i) {
      ub += in[i].rows_ub;
      kin.run[i] = in[i].view();
    }
    kin.ways = ways;
    DevicePool &pool = pool_for(s);
    Run out = alloc_run(ub, pool);
    const u32 tiles = static_cast<u32>(
        std::max<u64>(1, (ub + Shape::kTarget - 1) / Shape::kTarget));
    Lease<u32> parts(pool, (u64(tiles) + 1) * kMaxWay);
    Lease<u32> bucket_of(pool, u64(tiles) + 1);
    Lease<u32> ctrl(pool, 64 + 2ull * tiles);
    u64 *status = reinterpret_cast<u64 *>(ctrl.get() + 64);
    // directory bits never exceed the allocation's: 2^coarse buckets cover
    // the coarsest directory of the inputs
    u32 coarse = kMaxDirBits;
    for (u32 i = 0; i < ways; ++i) coarse = std::min(coarse, in[i].max_bits);
    const u64 buckets = u64(1) << coarse;
    kway_bucket_kernel<T><<<static_cast<u32>((buckets + 255) / 256), 256, 0,
                            s>>>(kin, tiles, Shape::kTarget, bucket_of.get());
    kway_partition_kernel<T><<<(tiles + 4) / 4, 256, 0, s>>>(
        kin, tiles, Shape::kTarget, Shape::kLightMax, bucket_of.get(),
        parts.get(), status, ctrl.get(), out.meta(), out.max_bits);
    GD_LAUNCH_CHECK();
    ProfRec rec{};
    if (profiling_) rec = prof_begin(0, s, ways);
    kway_merge_kernel<T, Shape::kItems, Shape::kThreads>
        <<<tiles, Shape::kThreads, 0, s>>>(
            kin, parts.get(), out.rows, ub, out.meta(), out.entries(),
            out.max_bits, drop_tombstones ? 1 : 0, status, ctrl.get());
    GD_LAUNCH_CHECK();
    if (profiling_) {
      prof_end(rec, s);
      prof_metas(rec.slot, in, ways, out, s);
    }
    return out;
  }

  Run empty_run(hipStream_t s) {
    Run r = alloc_run(0, pool_for(s));
    zero_two_kernel<<<1, 256, 0, s>>>(reinterpret_cast<u32 *>(r.meta()),
                                      kMetaWords, r.entries(), 3);
    GD_LAUNCH_CHECK();
    launch_directory(r, s);
    return r;
  }

  // Merges every run (newest first) into one tombstone-free run.  With
  // consume=true the inputs are released.
  Run merge_all(std::vector<Run> &runs, bool consume, hipStream_t s) {
    Run acc;
    bool acc_owned = false;
    if (runs.size() == 1) {
      Run e = empty_run(s);
      acc = merge_two(runs[0], e, true, s);
      free_run(e);
      if (consume) free_run(runs[0]);
      return acc;
    }
    for (std::size_t i = 1; i < runs.size(); ++i) {
      const Run &a = i == 1 ? runs[0] : acc;
      Run next = merge_two(a, runs[i], i + 1 == runs.size(), s);
      if (acc_owned) free_run(acc);
      acc = next;
      acc_owned = true;
    }
    if (consume)
      for (auto &r : runs) free_run(r);
    return acc;
  }

  // Binary-counter policy: merge the two newest runs while the older one is
  // no larger (equal batches: log2(N/b) merges per row).  The runs such a
  // cascade would absorb one after another (each no larger than the runs
  // merged before it) are merged in one pass of up to merge_ways runs, which
  // reads and writes every row once instead of once per level.  Uneven
  // batches can stack runs of decreasing size that never trigger it, so the
  // number of runs is also capped: over the cap, the adjacent pair with the
  // smallest combined size merges (adjacent runs are consecutive in age, so
  // the newer-wins rule of the merge still holds).  Tombstones drop when the
  // output becomes the oldest run.
  void rebalance(hipStream_t s) {
    const std::size_t cap =
        std::max<std::size_t>(2, std::min<std::size_t>(tuning_.max_runs
                                                           ? tuning_.max_runs
                                                           : 10,
                                                       kMaxRuns - 1));
    const std::size_t max_ways =
        tuning_.merge_ways ? tuning_.merge_ways : std::size_t(kMaxWay);
    const std::size_t fan = std::min<std::size_t>(
        max_ways, tuning_.merge_fanout ? tuning_.merge_fanout : 2);
    while (runs_.size() >= 2) {
      std::size_t i, ways = 2;
      if (runs_.size() >= fan && runs_[fan - 1].rows_ub <= runs_[0].rows_ub) {
        i = 0;
        ways = fan;
        u64 acc = 0;
        for (std::size_t j = 0; j < fan; ++j) acc += runs_[j].rows_ub;
        while (ways < runs_.size() && ways < max_ways &&
               runs_[ways].rows_ub <= acc)
          acc += runs_[ways++].rows_ub;
      } else if (runs_.size() > cap) {
        i = 0;
        u64 best = ~0ull;
        for (std::size_t j = 0; j + 1 < runs_.size(); ++j) {
          const u64 c = runs_[j].rows_ub + runs_[j + 1].rows_ub;
          if (c < best) {
            best = c;
            i = j;
          }
        }
      } else {
        break;
      }
      const bool oldest = i + ways == runs_.size();
      Run out = ways == 2
                    ? merge_two(runs_[i], runs_[i + 1], oldest, s)
                    : merge_many(&runs_[i], static_cast<u32>(ways), oldest, s);
      if (ms_ && s == ms_) {
        PendingMerge pm{nullptr, std::vector<Run>(runs_.begin() + i,
                                                  runs_.begin() + i + ways)};
        GD_HIP(hipEventCreateWithFlags(&pm.done, hipEventDisableTiming));
        GD_HIP(hipEventRecord(pm.done, s));
        auto old = pending_.find(out.rows);  // a stale, completed record
        if (old != pending_.end()) {
          (void)hipEventDestroy(old->second.done);
          pending_.erase(old);
        }
        pending_.emplace(out.rows, pm);
      }
      for (std::size_t j = 0; j < ways; ++j) free_run(runs_[i + j]);
      runs_.erase(runs_.begin() + i, runs_.begin() + i + ways);
      runs_.insert(runs_.begin() + i, out);
    }
  }

  u32 read_count(const Run &r, hipStream_t s) {
    RunMeta m{};
    GD_HIP(hipMemcpyAsync(&m, r.meta(), sizeof(m), hipMemcpyDeviceToHost, s));
    GD_HIP(hipStreamSynchronize(s));
    return m.count;
  }

  // Epoch boundary: survivors must fit the capacity (GPULSMOpt.cuh:7350).
  void epoch_capacity_check(hipStream_t s) {
    u64 ub = 0;
    for (const auto &r : runs_) ub += r.rows_ub;
    if (ub <= capacity_) return;
    // exact row counts
    u64 exact = 0;
    for (auto &r : runs_) {
      r.rows_ub = read_count(r, s);
      exact += r.rows_ub;
    }
    if (exact <= capacity_) return;
    // only a full merge tells the live count
    Run merged = merge_all(runs_, /*consume=*/true, s);
    runs_.clear();
    merged.rows_ub = read_count(merged, s);
    runs_.push_back(merged);
    if (merged.rows_ub > capacity_) {
      failed_ = true;
      status_ = GD_PUBLICATION_OUTPUT_OVERFLOW;
    }
  }

  RunSet<T> run_set() const {
    RunSet<T> set{};
    set.count = static_cast<u32>(runs_.size());
    for (std::size_t i = 0; i < runs_.size(); ++i) set.run[i] = runs_[i].view();
    return set;
  }

  struct Deferred {
    hipEvent_t merges_done = nullptr;   // phase 1 (merge stream)
    hipEvent_t queries_done = nullptr;  // phase 2 (query stream)
    std::vector<void *> buffers;
  };
  struct PendingMerge {
    hipEvent_t done;
    std::vector<Run> in;  // inputs, newest first
  };

  u64 batch_capacity_ = 1;
  u64 capacity_ = 0;
  // segmented admission (1) for 32-bit batches; 64-bit batches always take
  // the radix passes
  gd_lsm_tuning tuning_{3, 12, 0, 18, static_cast<u32>(T::kSmallTile), 10,
                        1, 512, kMaxWay, 2};

  struct ProfRec {
    int kind;  // 0 merge, 1 lookup, 2 admission of one tile, 3 range
    hipEvent_t start, end;
    u64 slot, n;  // n: merge inputs, lookup or range queries, tile keys
    u64 runs;     // range: runs searched
  };
  static constexpr u64 kProfSlots = 4096;
  static constexpr u64 kProfMetas = kMaxWay + 1;  // inputs, then the output
  bool profiling_ = false;
  std::vector<ProfRec> prof_;
  RunMeta *prof_meta_ = nullptr;  // pinned: [slot][inputs..., out]
  u64 prof_next_ = 0;

  ProfRec prof_begin(int kind, hipStream_t s, u64 n) {
    ProfRec r{kind, nullptr, nullptr, prof_next_++, n, 0};
    GD_HIP(hipEventCreate(&r.start));
    GD_HIP(hipEventCreate(&r.end));
    GD_HIP(hipEventRecord(r.start, s));
    return r;
  }
  void prof_end(ProfRec r, hipStream_t s) {
    GD_HIP(hipEventRecord(r.end, s));
    prof_.push_back(r);
  }
  void clear_profile() {
    for (auto &r : prof_) {
      (void)hipEventDestroy(r.start);
      (void)hipEventDestroy(r.end);
    }
    prof_.clear();
  }
  bool defer_ = false;
  std::vector<void *> pending_free_;
  std::deque<Deferred> deferred_;
  hipStream_t ms_ = nullptr, qs_ = nullptr;  // serving mode streams
  std::unordered_map<const void *, PendingMerge> pending_;
  PoolUsage usage_;
  DevicePool pool_{&usage_};
  DevicePool merge_pool_{&usage_};
  OperationOrder order_;
  std::vector<Run> runs_;  // newest first
  u32 pending_batches_ = 0;
  bool failed_ = false;
  u32 status_ = 0;
};

using Lsm = LsmT<KV<u32>>;
using Lsm64 = LsmT<KV<u64>>;

}  // namespace lsm
}  // namespace gdict
