// Serving engine over the LSM dictionary (gd_serve_*): continuous batching
// of insert / erase / lookup / range requests of any size.
//
// A request is admitted into a *step* when it arrives, as long as the
// step's staging memory fits the HBM budget next to the dictionary.  A step
// coalesces every admitted update (in arrival order, so the newest update
// still wins) into ONE mixed batch, and every admitted lookup and range
// query into one launch each.  Updates of a step are applied before its
// queries, so every query sees every update admitted up to its step: the
// result of each request is the one a sequential replay in step order
// gives (linearizable: requests of one step are concurrent).
//
// Overlap: updates run on the update stream, queries on the query stream.
// The queries of step k read the runs that existed after step k's updates
// while step k+1's updates already sort and merge; runs those updates
// release stay allocated until the query stream has passed them
// (Lsm::retire_pending / reclaim).  Up to `max_inflight_steps` steps are in
// flight; the engine thread completes requests as their step's query event
// fires, stamping completion times for p50/p99/p999.
//
// Gathering scattered request buffers into one batch (and scattering results
// back) is one kernel each, driven by a piece table copied once per step,
// so a step costs the same number of launches for 1 or 10,000 requests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "lsm_runtime.hpp"

namespace gdict {
namespace serve {

using lsm::Lsm;
using lsm::RunSet;

struct Piece {
  const void *a;   // gather: source keys / lo;  scatter: destination values
  const void *b;   // gather: source values / hi; scatter: destination found
  u64 count;
  u64 offset;      // position of the piece in the coalesced batch
  u32 flag;        // gather: erase flag of an update piece
  u32 pad;
};

__device__ __forceinline__ u32 piece_of(const Piece *pieces, u32 np, u64 i) {
  u32 lo = 0, hi = np;  // last piece with offset <= i
  while (hi - lo > 1) {
    const u32 mid = (lo + hi) >> 1;
    if (pieces[mid].offset <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

static __global__ void gather_kernel(const Piece *pieces, u32 np, u64 total,
                                     u32 *out_a, u32 *out_b, u8 *out_flag) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += stride) {
    const Piece &p = pieces[piece_of(pieces, np, i)];
    const u64 j = i - p.offset;
    out_a[i] = static_cast<const u32 *>(p.a)[j];
    if (out_b) out_b[i] = p.b ? static_cast<const u32 *>(p.b)[j] : 0u;
    if (out_flag) out_flag[i] = static_cast<u8>(p.flag);
  }
}

static __global__ void scatter_kernel(const Piece *pieces, u32 np, u64 total,
                                      const u32 *src_a, const u8 *src_b) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += stride) {
    const Piece &p = pieces[piece_of(pieces, np, i)];
    const u64 j = i - p.offset;
    static_cast<u32 *>(const_cast<void *>(p.a))[j] = src_a[i];
    if (p.b && src_b) static_cast<u8 *>(const_cast<void *>(p.b))[j] = src_b[i];
  }
}

// A mixed range step computes counts and sums; keep the one each query
// asked for.
static __global__ void select_kernel(const u8 *is_count, u32 *out,
                                     const u32 *sums, u64 n) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    if (!is_count[i]) out[i] = sums[i];
}

// Lookup results of requests without a found array get the reference's
// miss sentinel 0xFFFFFFFF (GPULSMOpt.cuh:4729-4744); with one, 0 + found 0.
static __global__ void lookup_fix_kernel(const Piece *pieces, u32 np, u64 n,
                                         const u32 *outv, const u8 *outf,
                                         u32 *fixed) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const Piece &p = pieces[piece_of(pieces, np, i)];
    fixed[i] = (outf[i] || p.b) ? outv[i] : 0xffffffffu;
  }
}

// Pinned host buffers for the per-step piece tables, recycled by size.
class PinnedCache {
 public:
  ~PinnedCache() {
    for (auto &kv : free_) (void)hipHostFree(kv.second);
  }
  void *get(std::size_t bytes) {
    const std::size_t cls = DevicePool::size_class(bytes);
    auto it = free_.find(cls);
    if (it != free_.end()) {
      void *p = it->second;
      free_.erase(it);
      return p;
    }
    void *p = nullptr;
    GD_HIP(hipHostMalloc(&p, cls, hipHostMallocDefault));
    size_[p] = cls;
    return p;
  }
  void put(void *p) {
    if (p) free_.emplace(size_[p], p);
  }

 private:
  std::multimap<std::size_t, void *> free_;
  std::unordered_map<void *, std::size_t> size_;
};

using Clock = std::chrono::steady_clock;

struct Request {
  u64 ticket = 0;
  gd_op op = GD_OP_LOOKUP;
  const void *a = nullptr;
  const void *b = nullptr;
  u64 count = 0;
  void *out = nullptr;
  u8 *out_found = nullptr;
  Clock::time_point submitted;
  Clock::time_point completed;
  gd_status status = GD_OK;
  u64 step = 0;
  bool done = false;
};

struct Step {
  u64 id = 0;
  std::vector<std::shared_ptr<Request>> reqs;
  hipEvent_t queries_done = nullptr;
  std::vector<void *> scratch;  // engine-pool buffers released at completion
  Piece *host_pieces = nullptr;  // pinned piece tables
  u64 staged_bytes = 0;
  gd_status update_status = GD_OK;
};

class Server : public HandleBase {
 public:
  explicit Server(const gd_serve_config &c)
      : cfg_(c), lsm_(c.lsm) {
    lsm_.set_deferred_reclaim(true);
    if (c.tuning) lsm_.set_tuning(*c.tuning);
    GD_HIP(hipStreamCreateWithFlags(&us_, hipStreamNonBlocking));
    GD_HIP(hipStreamCreateWithFlags(&qs_, hipStreamNonBlocking));
    max_inflight_ = c.max_inflight_steps ? c.max_inflight_steps : 2;
    max_step_updates_ = c.max_step_keys ? c.max_step_keys : (u64(1) << 22);
    max_step_queries_ = c.max_step_keys ? c.max_step_keys : (u64(1) << 24);
    budget_ = c.hbm_budget_bytes;
    worker_ = std::thread([this] { run(); });
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    try {
      GD_HIP(hipStreamSynchronize(us_));
      GD_HIP(hipStreamSynchronize(qs_));
    } catch (...) {
    }
    for (auto &st : inflight_) release_step(st);
    inflight_.clear();
    (void)hipStreamDestroy(us_);
    (void)hipStreamDestroy(qs_);
  }

  u64 submit(gd_op op, const void *a, const void *b, u64 count, void *out,
             u8 *out_found) {
    if (op > GD_OP_RANGE_SUM)
      throw Error(GD_INVALID_ARGUMENT, "unknown request operation");
    if (count && !a) throw Error(GD_INVALID_ARGUMENT, "null request keys");
    if (count && op == GD_OP_INSERT && !b)
      throw Error(GD_INVALID_ARGUMENT, "null insert values");
    if (count && (op == GD_OP_RANGE_COUNT || op == GD_OP_RANGE_SUM) && !b)
      throw Error(GD_INVALID_ARGUMENT, "null range upper bounds");
    if (count && op >= GD_OP_LOOKUP && !out)
      throw Error(GD_INVALID_ARGUMENT, "null request output");
    if (budget_ && request_bytes(op, count) > budget_)
      throw Error(GD_CAPACITY, "request of " + std::to_string(count) +
                                   " keys exceeds the HBM budget");
    auto r = std::make_shared<Request>();
    r->op = op;
    r->a = a;
    r->b = b;
    r->count = count;
    r->out = out;
    r->out_found = out_found;
    r->submitted = Clock::now();
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (worker_error_ != GD_OK)
        throw Error(worker_error_, "serving engine stopped: " + worker_msg_);
      r->ticket = ++next_ticket_;
      queue_.push_back(r);
      table_[r->ticket] = r;
    }
    cv_.notify_all();
    return r->ticket;
  }

  // Waits for one request; returns its status.
  gd_status wait(u64 ticket, u64 *step_out) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = table_.find(ticket);
    if (it == table_.end())
      throw Error(GD_INVALID_ARGUMENT, "unknown ticket");
    auto r = it->second;
    done_cv_.wait(lk, [&] { return r->done || worker_error_ != GD_OK; });
    if (!r->done) throw Error(worker_error_, worker_msg_);
    if (step_out) *step_out = r->step;
    table_.erase(ticket);
    return r->status;
  }

  void drain() {
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] {
      return (queue_.empty() && busy_ == 0) || worker_error_ != GD_OK;
    });
    if (worker_error_ != GD_OK) throw Error(worker_error_, worker_msg_);
  }

  void stats(gd_serve_stats *out) {
    std::lock_guard<std::mutex> lk(mu_);
    out->requests = completed_;
    out->steps = steps_done_;
    out->coalesced_updates = coalesced_updates_;
    out->coalesced_queries = coalesced_queries_;
    std::vector<double> l = latencies_;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {
      if (l.empty()) return 0.0;
      const std::size_t i = std::min(
          l.size() - 1, static_cast<std::size_t>(p * (l.size() - 1) + 0.5));
      return l[i];
    };
    out->latency_p50_us = pct(0.50);
    out->latency_p99_us = pct(0.99);
    out->latency_p999_us = pct(0.999);
    out->latency_max_us = l.empty() ? 0.0 : l.back();
    out->peak_staged_bytes = peak_staged_;
    out->resident_bytes = resident_bytes_.load();
  }

  void reset_stats() {
    std::lock_guard<std::mutex> lk(mu_);
    latencies_.clear();
    completed_ = 0;
    steps_done_ = 0;
    coalesced_updates_ = 0;
    coalesced_queries_ = 0;
    peak_staged_ = 0;
  }

 private:
  static u64 request_bytes(gd_op op, u64 count) {
    // staging of the coalesced batch: keys + values + flags, or queries +
    // results (+ found), plus the grouped-lookup scratch (2 x 4 B)
    switch (op) {
      case GD_OP_INSERT:
      case GD_OP_ERASE: return count * 9 + 4 * count * 8;  // + sort scratch
      case GD_OP_LOOKUP: return count * (4 + 4 + 1 + 8);
      default: return count * (4 + 4 + 4);
    }
  }

  void run() {
    while (true) {
      std::vector<std::shared_ptr<Request>> take;
      u64 staged = 0;
      {
        std::unique_lock<std::mutex> lk(mu_);
        // with work in flight, do not block: complete it first
        if (queue_.empty() && inflight_.empty())
          cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
        if (stop_ && queue_.empty() && inflight_.empty()) return;
        // admission: arrival order, up to the step limits and the budget
        u64 upd = 0, qry = 0;
        const u64 resident = lsm_.resident_bytes() + in_flight_bytes_;
        while (!queue_.empty()) {
          auto &r = queue_.front();
          const bool is_upd = r->op <= GD_OP_ERASE;
          const u64 bytes = request_bytes(r->op, r->count);
          if (!take.empty()) {
            if (is_upd && upd + r->count > max_step_updates_) break;
            if (!is_upd && qry + r->count > max_step_queries_) break;
            if (budget_ && resident + staged + bytes > budget_) break;
          }
          (is_upd ? upd : qry) += r->count;
          staged += bytes;
          take.push_back(r);
          queue_.pop_front();
        }
        busy_ += take.size();
      }
      try {
        if (!take.empty()) launch_step(std::move(take), staged);
        // complete steps: all of them when idle, else down to the limit
        complete_steps();
      } catch (const Error &e) {
        fail_all(e.status, e.what());
        return;
      } catch (const std::exception &e) {
        fail_all(GD_HIP_ERROR, e.what());
        return;
      }
    }
  }

  void fail_all(gd_status st, const std::string &msg) {
    std::lock_guard<std::mutex> lk(mu_);
    worker_error_ = st;
    worker_msg_ = msg;
    done_cv_.notify_all();
  }

  template <class T>
  T *scratch(Step &st, u64 n) {
    T *p = pool_.get_n<T>(std::max<u64>(n, 1));
    st.scratch.push_back(p);
    return p;
  }

  void launch_step(std::vector<std::shared_ptr<Request>> reqs, u64 staged) {
    Step st;
    st.id = ++step_counter_;
    st.staged_bytes = staged;
    std::vector<Piece> upd, look, rng;
    u64 n_upd = 0, n_look = 0, n_rng = 0;
    bool any_erase = false, any_insert = false;
    for (auto &r : reqs) {
      r->step = st.id;
      if (!r->count) continue;
      if (r->op <= GD_OP_ERASE) {
        upd.push_back({r->a, r->op == GD_OP_INSERT ? r->b : nullptr, r->count,
                       n_upd, r->op == GD_OP_ERASE ? 1u : 0u, 0});
        n_upd += r->count;
        (r->op == GD_OP_ERASE ? any_erase : any_insert) = true;
      } else if (r->op == GD_OP_LOOKUP) {
        look.push_back({r->a, nullptr, r->count, n_look, 0, 0});
        n_look += r->count;
      } else {
        rng.push_back({r->a, r->b, r->count, n_rng,
                       r->op == GD_OP_RANGE_COUNT ? 1u : 0u, 0});
        n_rng += r->count;
      }
    }
    // one pinned piece table per step: [upd | look | look-out | rng | rng-out]
    const std::size_t total_pieces =
        upd.size() + 2 * look.size() + 2 * rng.size();
    if (total_pieces)
      st.host_pieces =
          static_cast<Piece *>(pinned_.get(total_pieces * sizeof(Piece)));
    Piece *dev_pieces =
        total_pieces ? scratch<Piece>(st, total_pieces) : nullptr;
    std::size_t at = 0;
    auto put = [&](const std::vector<Piece> &v) {
      std::copy(v.begin(), v.end(), st.host_pieces + at);
      const std::size_t first = at;
      at += v.size();
      return first;
    };
    const std::size_t upd_at = put(upd);
    const std::size_t look_at = put(look);
    std::vector<Piece> look_out, rng_out;
    {
      std::size_t k = 0;
      for (auto &r : reqs)
        if (r->count && r->op == GD_OP_LOOKUP)
          look_out.push_back({r->out, r->out_found, r->count,
                              look[k++].offset, 0, 0});
      k = 0;
      for (auto &r : reqs)
        if (r->count && r->op >= GD_OP_RANGE_COUNT)
          rng_out.push_back({r->out, nullptr, r->count, rng[k++].offset, 0, 0});
    }
    const std::size_t look_out_at = put(look_out);
    const std::size_t rng_at = put(rng);
    const std::size_t rng_out_at = put(rng_out);
    if (total_pieces)
      GD_HIP(hipMemcpyAsync(dev_pieces, st.host_pieces,
                            total_pieces * sizeof(Piece),
                            hipMemcpyHostToDevice, us_));

    // ---------------------------------------------------------- updates
    if (n_upd) {
      u32 *keys = scratch<u32>(st, n_upd);
      u32 *vals = scratch<u32>(st, n_upd);
      u8 *flags = scratch<u8>(st, n_upd);
      gather_kernel<<<grid_for(n_upd, 256, 1u << 14), 256, 0, us_>>>(
          dev_pieces + upd_at, static_cast<u32>(upd.size()), n_upd, keys,
          vals, flags);
      GD_LAUNCH_CHECK();
      const int mode = any_erase ? (any_insert ? lsm::kModeMixed
                                               : lsm::kModeErase)
                                 : lsm::kModeInsert;
      try {
        lsm_.admit_on(keys, vals, flags, n_upd, mode, us_);
      } catch (const Error &e) {
        if (e.status == GD_HIP_ERROR) throw;
        st.update_status = e.status;  // reported on this step's updates
      }
      // runs released by these updates may still be read by queries
      // already enqueued on the query stream
      lsm_.retire_pending(qs_);
    }
    hipEvent_t updates_done;
    GD_HIP(hipEventCreateWithFlags(&updates_done, hipEventDisableTiming));
    GD_HIP(hipEventRecord(updates_done, us_));
    GD_HIP(hipStreamWaitEvent(qs_, updates_done, 0));
    (void)hipEventDestroy(updates_done);

    // ---------------------------------------------------------- queries
    const RunSet snap = lsm_.snapshot();
    if (n_look) {
      u32 *q = scratch<u32>(st, n_look);
      u32 *outv = scratch<u32>(st, n_look);
      u8 *outf = scratch<u8>(st, n_look);
      gather_kernel<<<grid_for(n_look, 256, 1u << 14), 256, 0, qs_>>>(
          dev_pieces + look_at, static_cast<u32>(look.size()), n_look, q,
          nullptr, nullptr);
      GD_LAUNCH_CHECK();
      if (n_look >= lsm_.grouped_lookup_min() && !lsm_.empty()) {
        u32 *qk = scratch<u32>(st, n_look);
        u32 *qp = scratch<u32>(st, n_look);
        u32 *ctrl = scratch<u32>(st, rs::control_words(n_look, 1));
        GD_HIP(rs::partition_pairs<u32>(q, static_cast<u32>(n_look), 24, qk,
                                        qp, ctrl, qs_,
                                        lsm_.tuning().sort_block));
        const u32 blocks =
            (std::min<u32>(grid_for(n_look, 1024, 1u << 16), 1u << 16) + 7u) &
            ~7u;
        lsm::lookup_grouped_kernel<<<blocks, 256, 0, qs_>>>(
            qk, qp, n_look, outv, outf, snap);
      } else {
        lsm::lookup_kernel<<<grid_for(n_look, 256, 1u << 16), 256, 0, qs_>>>(
            q, n_look, outv, outf, snap);
      }
      GD_LAUNCH_CHECK();
      // requests without a found array get the reference's miss sentinel
      scatter_lookup(st, dev_pieces + look_out_at,
                     static_cast<u32>(look_out.size()), n_look, outv, outf);
    }
    if (n_rng) {
      u32 *lo = scratch<u32>(st, n_rng);
      u32 *hi = scratch<u32>(st, n_rng);
      u8 *is_count = scratch<u8>(st, n_rng);
      u32 *out = scratch<u32>(st, n_rng);
      gather_kernel<<<grid_for(n_rng, 256, 1u << 14), 256, 0, qs_>>>(
          dev_pieces + rng_at, static_cast<u32>(rng.size()), n_rng, lo, hi,
          is_count);
      GD_LAUNCH_CHECK();
      // counts and sums share the kernel: a count query sums 1 per row
      bool all_count = true, all_sum = true;
      for (auto &p : rng) (p.flag ? all_sum : all_count) = false;
      if (all_count || all_sum) {
        lsm::range_kernel<<<grid_for(n_rng, lsm::kRangeWaves, 1u << 16),
                            lsm::kRangeThreads, 0, qs_>>>(
            lo, hi, n_rng, out, all_count ? 1 : 0, snap);
      } else {
        u32 *out2 = scratch<u32>(st, n_rng);
        lsm::range_kernel<<<grid_for(n_rng, lsm::kRangeWaves, 1u << 16),
                            lsm::kRangeThreads, 0, qs_>>>(lo, hi, n_rng, out,
                                                          1, snap);
        lsm::range_kernel<<<grid_for(n_rng, lsm::kRangeWaves, 1u << 16),
                            lsm::kRangeThreads, 0, qs_>>>(lo, hi, n_rng, out2,
                                                          0, snap);
        select_kernel<<<grid_for(n_rng, 256, 1u << 14), 256, 0, qs_>>>(
            is_count, out, out2, n_rng);
      }
      GD_LAUNCH_CHECK();
      scatter_kernel<<<grid_for(n_rng, 256, 1u << 14), 256, 0, qs_>>>(
          dev_pieces + rng_out_at, static_cast<u32>(rng_out.size()), n_rng,
          out, nullptr);
      GD_LAUNCH_CHECK();
    }
    GD_HIP(hipEventCreateWithFlags(&st.queries_done, hipEventDisableTiming));
    GD_HIP(hipEventRecord(st.queries_done, qs_));
    st.reqs = std::move(reqs);
    {
      std::lock_guard<std::mutex> lk(mu_);
      coalesced_updates_ += upd.size();
      coalesced_queries_ += look.size() + rng.size();
      in_flight_bytes_ += st.staged_bytes;
      peak_staged_ = std::max(peak_staged_, in_flight_bytes_);
    }
    inflight_.push_back(std::move(st));
  }

  void scatter_lookup(Step &st, const Piece *pieces, u32 np, u64 n,
                      const u32 *outv, const u8 *outf) {
    // miss value: 0 with a found array, 0xFFFFFFFF without
    // (GPULSMOpt.cuh:4729-4744); outv already holds 0 for misses
    u32 *fixed = scratch<u32>(st, n);
    lookup_fix_kernel<<<grid_for(n, 256, 1u << 14), 256, 0, qs_>>>(
        pieces, np, n, outv, outf, fixed);
    scatter_kernel<<<grid_for(n, 256, 1u << 14), 256, 0, qs_>>>(
        pieces, np, n, fixed, outf);
    GD_LAUNCH_CHECK();
  }

  void complete_steps() {
    while (!inflight_.empty()) {
      Step &st = inflight_.front();
      const bool idle = [&] {
        std::lock_guard<std::mutex> lk(mu_);
        return queue_.empty();
      }();
      if (inflight_.size() < max_inflight_ && !idle) {
        // room for another step: only harvest what already finished
        const hipError_t q = hipEventQuery(st.queries_done);
        if (q == hipErrorNotReady) break;
        GD_HIP(q);
      } else {
        GD_HIP(hipEventSynchronize(st.queries_done));
      }
      const auto now = Clock::now();
      {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto &r : st.reqs) {
          r->completed = now;
          r->status = r->op <= GD_OP_ERASE ? st.update_status : GD_OK;
          r->done = true;
          latencies_.push_back(
              std::chrono::duration<double, std::micro>(now - r->submitted)
                  .count());
          ++completed_;
        }
        busy_ -= st.reqs.size();
        in_flight_bytes_ -= st.staged_bytes;
        ++steps_done_;
      }
      done_cv_.notify_all();
      release_step(st);
      inflight_.pop_front();
      lsm_.reclaim();
      resident_bytes_.store(lsm_.resident_bytes() + pool_.held_bytes());
    }
  }

  void release_step(Step &st) {
    for (void *p : st.scratch) pool_.put(p);
    st.scratch.clear();
    pinned_.put(st.host_pieces);
    st.host_pieces = nullptr;
    if (st.queries_done) (void)hipEventDestroy(st.queries_done);
    st.queries_done = nullptr;
  }

  gd_serve_config cfg_;
  Lsm lsm_;
  DevicePool pool_;  // step staging, released when the step completes
  PinnedCache pinned_;
  std::atomic<u64> resident_bytes_{0};
  hipStream_t us_ = nullptr, qs_ = nullptr;
  u32 max_inflight_ = 2;
  u64 max_step_updates_ = 0, max_step_queries_ = 0;
  u64 budget_ = 0;

  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<std::shared_ptr<Request>> queue_;
  std::unordered_map<u64, std::shared_ptr<Request>> table_;
  std::deque<Step> inflight_;
  u64 next_ticket_ = 0, step_counter_ = 0;
  u64 busy_ = 0;
  bool stop_ = false;
  gd_status worker_error_ = GD_OK;
  std::string worker_msg_;
  std::vector<double> latencies_;
  u64 completed_ = 0, steps_done_ = 0;
  u64 coalesced_updates_ = 0, coalesced_queries_ = 0;
  u64 in_flight_bytes_ = 0, peak_staged_ = 0;
  std::thread worker_;
};

}  // namespace serve
}  // namespace gdict

using gdict::guarded;
using gdict::serve::Server;

struct gd_server : Server {
  using Server::Server;
};

extern "C" {

gd_status gd_serve_create(const gd_serve_config *config, gd_server **out) {
  if (!config || !out) {
    gdict::set_create_error("null config or output handle");
    return GD_INVALID_ARGUMENT;
  }
  *out = nullptr;
  return guarded(nullptr, [&] {
    (void)gdict::lsm::Lsm::capacity_for(config->lsm);
    *out = new gd_server(*config);
  });
}

gd_status gd_serve_destroy(gd_server *h) {
  delete h;
  return GD_OK;
}

gd_status gd_serve_submit(gd_server *h, gd_op op, const void *a,
                          const void *b, uint64_t count, void *out,
                          uint8_t *out_found, uint64_t *out_ticket) {
  if (!h || !out_ticket) return GD_INVALID_ARGUMENT;
  return guarded(h, [&] {
    *out_ticket = h->submit(op, a, b, count, out, out_found);
  });
}

gd_status gd_serve_wait(gd_server *h, uint64_t ticket, uint64_t *out_step) {
  if (!h) return GD_INVALID_ARGUMENT;
  gd_status result = GD_OK;
  gd_status st = guarded(h, [&] { result = h->wait(ticket, out_step); });
  if (st != GD_OK) return st;
  if (result != GD_OK)
    h->last_error = "the request's update step failed (status " +
                    std::to_string(result) + ")";
  return result;
}

gd_status gd_serve_drain(gd_server *h) {
  if (!h) return GD_INVALID_ARGUMENT;
  return guarded(h, [&] { h->drain(); });
}

gd_status gd_serve_get_stats(gd_server *h, gd_serve_stats *out, int reset) {
  if (!h || !out) return GD_INVALID_ARGUMENT;
  return guarded(h, [&] {
    h->stats(out);
    if (reset) h->reset_stats();
  });
}

}  // extern "C"
