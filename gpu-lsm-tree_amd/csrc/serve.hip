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
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "lsm_runtime.hpp"

namespace gdict {
namespace serve {

using lsm::KV;
using lsm::LsmT;
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

template <class K, class V>
__global__ void gather_kernel(const Piece *pieces_arg, u32 np, u64 total,
                              K *out_a, V *out_b, u8 *out_flag) {
  // tables of up to 256 pieces are searched in LDS
  __shared__ Piece s_pieces[256];
  const Piece *pieces = pieces_arg;
  if (np <= 256) {
    for (u32 i = threadIdx.x; i < np; i += blockDim.x)
      s_pieces[i] = pieces_arg[i];
    pieces = s_pieces;
  }
  __syncthreads();
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += stride) {
    const Piece &p = pieces[piece_of(pieces, np, i)];
    const u64 j = i - p.offset;
    out_a[i] = static_cast<const K *>(p.a)[j];
    if (out_b) out_b[i] = p.b ? static_cast<const V *>(p.b)[j] : V(0);
    if (out_flag) out_flag[i] = static_cast<u8>(p.flag);
  }
}

// A lookup or range request inside a step's coalesced query batch.
struct QueryPiece {
  const void *a;   // queries / lo (keys)
  const void *b;   // hi (ranges)
  void *out;       // values / counts / sums (value width)
  u8 *found;       // lookups: nullable found flags
  u64 offset;      // position of the request in the batch
  u32 count_only;  // ranges: count (1) or sum (0)
  u32 pad;
};

// Piece tables of up to kStagePieces entries are copied into LDS by the
// kernels that search them once per query.
constexpr u32 kStagePieces = 256;

__device__ __forceinline__ const QueryPiece *stage_pieces(
    const QueryPiece *pieces, u32 np, void *lds) {
  if (np > kStagePieces) return pieces;
  QueryPiece *dst = static_cast<QueryPiece *>(lds);
  for (u32 i = threadIdx.x; i < np; i += blockDim.x) dst[i] = pieces[i];
  return dst;
}

__device__ __forceinline__ u32 qpiece_of(const QueryPiece *pieces, u32 np,
                                         u64 i) {
  u32 lo = 0, hi = np;  // last piece with offset <= i
  while (hi - lo > 1) {
    const u32 mid = (lo + hi) >> 1;
    if (pieces[mid].offset <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// IO policies of the query kernels: read each query from its request and
// write its result into the request's own output (and found) buffer, with
// the miss convention of a plain lookup (GPULSMOpt.cuh:4714-4806).
template <class T>
struct PieceLookup {
  using K = typename T::K;
  using V = typename T::V;
  const QueryPiece *pieces;
  u32 np;
  static constexpr u32 kStageBytes = kStagePieces * sizeof(QueryPiece);
  __device__ __forceinline__ PieceLookup stage(void *lds) const {
    PieceLookup r = *this;
    r.pieces = stage_pieces(pieces, np, lds);
    return r;
  }
  struct Slot {
    V *out;
    u8 *found;
    u64 j;
  };
  __device__ __forceinline__ K load(u64 i, Slot &s) const {
    const QueryPiece &p = pieces[qpiece_of(pieces, np, i)];
    s.out = static_cast<V *>(p.out);
    s.found = p.found;
    s.j = i - p.offset;
    return static_cast<const K *>(p.a)[s.j];
  }
  __device__ __forceinline__ void put(const Slot &s, bool live,
                                      V value) const {
    if (s.found) {
      s.out[s.j] = live ? value : V(0);
      s.found[s.j] = live ? 1 : 0;
    } else {
      s.out[s.j] = live ? value : T::kAllOnes;
    }
  }
};

template <class T>
struct PieceRange {
  using K = typename T::K;
  using V = typename T::V;
  const QueryPiece *pieces;
  u32 np;
  static constexpr u32 kStageBytes = kStagePieces * sizeof(QueryPiece);
  __device__ __forceinline__ PieceRange stage(void *lds) const {
    PieceRange r = *this;
    r.pieces = stage_pieces(pieces, np, lds);
    return r;
  }
  struct Slot {
    V *out;
    u64 j;
  };
  __device__ __forceinline__ void load(u64 q, Slot &s, K &a, K &z,
                                       bool &count) const {
    const QueryPiece &p = pieces[qpiece_of(pieces, np, q)];
    s.out = static_cast<V *>(p.out);
    s.j = q - p.offset;
    a = static_cast<const K *>(p.a)[s.j];
    z = static_cast<const K *>(p.b)[s.j];
    count = p.count_only != 0;
  }
  __device__ __forceinline__ void put(const Slot &s, V v) const {
    s.out[s.j] = v;
  }
};

// Queries of a step's lookups as one contiguous array (grouped path).
template <class K>
__global__ void query_gather_kernel(const QueryPiece *pieces_arg, u32 np,
                                    u64 total, K *out) {
  __shared__ __align__(16) char s_io[kStagePieces * sizeof(QueryPiece)];
  const QueryPiece *pieces = stage_pieces(pieces_arg, np, s_io);
  __syncthreads();
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += stride) {
    const QueryPiece &p = pieces[qpiece_of(pieces, np, i)];
    out[i] = static_cast<const K *>(p.a)[i - p.offset];
  }
}

// ------------------------------------------------------- step graphs
// A small step (updates that fit one fused small-batch admission, lookups
// below the grouped threshold, ranges that need no compaction, at most
// kGraphPieces requests of each kind) runs as ONE graph launch: a copy of
// the step's descriptor (pinned host -> device), the gather of the update
// pieces, the fused admission, and one persistent query kernel whose
// workgroups take lookup and range chunks from a work queue (an atomic
// ticket in the descriptor, zero in every copy).  The graph is captured
// once per descriptor slot with every kernel reading its sizes and buffers
// from the slot, so it is replayed as is; no node clears memory.
constexpr u32 kGraphPieces = 64;
constexpr u32 kGraphSlots = 8;
constexpr u32 kGraphQueryBlocks = 256;
constexpr u32 kStepLookupChunk = 256;   // lookups per work item
constexpr u32 kStepRangeChunk = 32;     // ranges per work item

template <class T>
struct StepDesc {
  using K = typename T::K;
  using V = typename T::V;
  using Row = typename T::Row;
  // updates: pieces -> (keys, vals, flags) -> one run
  u64 n_upd;
  u32 np_upd;
  int mode;
  Row *run_rows;
  lsm::RunMeta *run_meta;
  u32 *run_dir;
  u32 run_max_bits;
  u32 rows_log;
  // queries over `snap`
  u64 n_look;
  u64 n_rng;
  u32 np_look;
  u32 np_rng;
  unsigned int ticket;
  u32 pad;
  RunSet<T> snap;
  Piece upd[kGraphPieces];
  QueryPiece look[kGraphPieces];
  QueryPiece rng[kGraphPieces];
};

// A step's updates read through its piece table (entry i of the step's
// batch is entry i - offset of the piece holding i).
template <class T>
struct PieceBatch {
  const Piece *pieces;
  u32 np;
  __device__ __forceinline__ const Piece &at(u32 i, u32 &j) const {
    const Piece &p = pieces[piece_of(pieces, np, i)];
    j = static_cast<u32>(i - p.offset);
    return p;
  }
  __device__ __forceinline__ typename T::K key(u32 i) const {
    u32 j;
    const Piece &p = at(i, j);
    return static_cast<const typename T::K *>(p.a)[j];
  }
  __device__ __forceinline__ typename T::V value(u32 i) const {
    u32 j;
    const Piece &p = at(i, j);
    return p.b ? static_cast<const typename T::V *>(p.b)[j]
               : typename T::V(0);
  }
  __device__ __forceinline__ bool erased(u32 i) const {
    u32 j;
    return at(i, j).flag != 0;
  }
};

// The step's fused admission, reading the update requests in place.
template <class T>
__global__ __launch_bounds__(lsm::kSmallThreads) void step_admit_kernel(
    const StepDesc<T> *d) {
  const u32 n = static_cast<u32>(d->n_upd);
  if (!n) return;
  __shared__ Piece s_p[kGraphPieces];
  const u32 np = d->np_upd;
  for (u32 i = threadIdx.x; i < np; i += blockDim.x) s_p[i] = d->upd[i];
  __syncthreads();
  lsm::small_batch_run<T>(PieceBatch<T>{s_p, np}, n, d->mode, d->run_rows,
                          d->run_meta, d->run_dir, d->run_max_bits,
                          d->rows_log);
}

// The step's lookups and ranges in one launch: a workgroup takes work items
// (kStepLookupChunk lookups, a thread per lookup, or kStepRangeChunk ranges
// for the range kernel's wave rounds) from the descriptor's ticket until
// none is left; workgroups beyond the item count leave at once.
template <class T>
__global__ __launch_bounds__(lsm::kRangeThreads) void step_query_kernel(
    StepDesc<T> *d) {
  const u64 n_look = d->n_look, n_rng = d->n_rng;
  const u32 lc = static_cast<u32>((n_look + kStepLookupChunk - 1) /
                                  kStepLookupChunk);
  const u32 rc =
      static_cast<u32>((n_rng + kStepRangeChunk - 1) / kStepRangeChunk);
  if (blockIdx.x >= lc + rc) return;
  __shared__ QueryPiece s_look[kGraphPieces];
  __shared__ QueryPiece s_rng[kGraphPieces];
  __shared__ u32 s_item;
  const u32 npl = d->np_look, npr = d->np_rng;
  for (u32 i = threadIdx.x; i < npl; i += blockDim.x) s_look[i] = d->look[i];
  for (u32 i = threadIdx.x; i < npr; i += blockDim.x) s_rng[i] = d->rng[i];
  const RunSet<T> &runs = d->snap;
  const PieceLookup<T> lio{s_look, npl};
  const PieceRange<T> rio{s_rng, npr};
  while (true) {
    __syncthreads();  // the pieces are staged; s_item is free again
    if (threadIdx.x == 0) s_item = atomicAdd(&d->ticket, 1u);
    __syncthreads();
    const u32 item = s_item;
    if (item >= lc + rc) break;
    if (item < lc) {
      const u64 end = std::min<u64>(n_look, u64(item + 1) * kStepLookupChunk);
      for (u64 i = u64(item) * kStepLookupChunk + threadIdx.x; i < end;
           i += blockDim.x) {
        typename PieceLookup<T>::Slot slot;
        const typename T::K key = lio.load(i, slot);
        typename T::V value = 0;
        const bool live = lsm::lookup_one(runs, key, value);
        lio.put(slot, live, value);
      }
    } else {
      const u64 r0 = u64(item - lc) * kStepRangeChunk;
      lsm::range_block<T, PieceRange<T>>(
          rio, r0, std::min<u64>(n_rng, r0 + kStepRangeChunk), runs,
          threadIdx.x / kWave, lsm::kRangeWaves);
    }
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
  hipEvent_t updates_done = nullptr;  // admissions of the step (us_)
  hipEvent_t queries_done = nullptr;  // its queries (qs_)
  bool updates_stamped = false;
  std::vector<void *> scratch;  // engine-pool buffers released at completion
  char *host_pieces = nullptr;  // pinned piece tables
  u64 staged_bytes = 0;
  gd_status update_status = GD_OK;
};

// Type-erased engine behind the C ABI (32- or 64-bit keys and values).
class ServerBase {
 public:
  virtual ~ServerBase() = default;
  virtual u64 submit(gd_op op, const void *a, const void *b, u64 count,
                     void *out, u8 *out_found) = 0;
  virtual gd_status wait(u64 ticket, u64 *step_out) = 0;
  virtual void drain() = 0;
  virtual void stats(gd_serve_stats *out) = 0;
  virtual void reset_stats() = 0;
};

template <class T>
class Server : public ServerBase {
 public:
  using K = typename T::K;
  using V = typename T::V;

  explicit Server(const gd_serve_config &c)
      : cfg_(c), lsm_(c.lsm) {
    if (c.tuning) lsm_.set_tuning(*c.tuning);
    GD_HIP(hipStreamCreateWithFlags(&us_, hipStreamNonBlocking));
    GD_HIP(hipStreamCreateWithFlags(&qs_, hipStreamNonBlocking));
    GD_HIP(hipStreamCreateWithFlags(&ms_, hipStreamNonBlocking));
    lsm_.set_async(ms_, qs_);
    max_inflight_ = c.max_inflight_steps ? c.max_inflight_steps : 2;
    max_step_updates_ = c.max_step_keys ? c.max_step_keys : (u64(1) << 22);
    max_step_queries_ = c.max_step_keys ? c.max_step_keys : (u64(1) << 24);
    budget_ = c.hbm_budget_bytes;
    if (c.step_launch > 2)
      throw Error(GD_INVALID_ARGUMENT, "step_launch must be 0, 1 or 2");
    if (c.step_launch != 0) init_slots(/*capture=*/c.step_launch == 2);
    worker_ = std::thread([this] { run(); });
  }

  ~Server() override {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    try {
      GD_HIP(hipStreamSynchronize(us_));
      GD_HIP(hipStreamSynchronize(ms_));
      GD_HIP(hipStreamSynchronize(qs_));
      lsm_.detach_streams();
    } catch (...) {
    }
    for (auto &st : inflight_) release_step(st);
    inflight_.clear();
    release_graphs();
    (void)hipStreamDestroy(us_);
    (void)hipStreamDestroy(ms_);
    (void)hipStreamDestroy(qs_);
  }

  u64 submit(gd_op op, const void *a, const void *b, u64 count, void *out,
             u8 *out_found) override {
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
  gd_status wait(u64 ticket, u64 *step_out) override {
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

  void drain() override {
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] {
      return (queue_.empty() && busy_ == 0) || worker_error_ != GD_OK;
    });
    if (worker_error_ != GD_OK) throw Error(worker_error_, worker_msg_);
  }

  void stats(gd_serve_stats *out) override {
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
    out->live_bytes = live_bytes_.load();
    out->peak_resident_bytes = peak_resident_.load();
    out->graph_steps = graph_steps_;
  }

  void reset_stats() override {
    std::lock_guard<std::mutex> lk(mu_);
    latencies_.clear();
    completed_ = 0;
    steps_done_ = 0;
    coalesced_updates_ = 0;
    coalesced_queries_ = 0;
    peak_staged_ = 0;
    graph_steps_ = 0;
  }

 private:
  static u64 request_bytes(gd_op op, u64 count) {
    // staging of the coalesced batch: keys + values + flags, or queries +
    // results (+ found), plus the grouped-lookup scratch (queries, grouped
    // keys, destinations, grouped results with their found flags)
    constexpr u64 kb = sizeof(K), vb = sizeof(V);
    switch (op) {
      case GD_OP_INSERT:
      case GD_OP_ERASE:  // + sort scratch (two key and two position arrays)
        return count * (kb + vb + 1) + count * (2 * kb + 8) + count * 8;
      case GD_OP_LOOKUP: return count * (kb + vb + 1 + 2 * kb + 4 + 2 * vb);
      default: return count * (2 * kb + vb);
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

  template <class E>
  E *scratch(Step &st, u64 n) {
    E *p = pool_.get_n<E>(std::max<u64>(n, 1));
    st.scratch.push_back(p);
    return p;
  }

  // ------------------------------------------------------- step graphs
  struct GraphSlot {
    StepDesc<T> *host = nullptr;  // pinned: the step's descriptor
    StepDesc<T> *dev = nullptr;   // the kernels' copy of it
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;    // the last step using the slot
    bool used = false;
  };

  // The three operations of a fused step, on `s`: the descriptor copy, the
  // admission and the query kernel.
  void enqueue_fused(const GraphSlot &sl, hipStream_t s) {
    GD_HIP(hipMemcpyAsync(sl.dev, sl.host, sizeof(StepDesc<T>),
                          hipMemcpyHostToDevice, s));
    step_admit_kernel<T><<<1, lsm::kSmallThreads, 0, s>>>(sl.dev);
    step_query_kernel<T><<<kGraphQueryBlocks, lsm::kRangeThreads, 0, s>>>(
        sl.dev);
  }

  // Descriptor slots; in graph mode one captured graph per slot.
  void init_slots(bool capture) {
    hipStream_t cs = nullptr;
    GD_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    try {
      slots_.resize(kGraphSlots);
      for (GraphSlot &sl : slots_) {
        GD_HIP(hipHostMalloc(reinterpret_cast<void **>(&sl.host),
                             sizeof(StepDesc<T>), hipHostMallocDefault));
        std::memset(static_cast<void *>(sl.host), 0, sizeof(StepDesc<T>));
        GD_HIP(hipMalloc(reinterpret_cast<void **>(&sl.dev),
                         sizeof(StepDesc<T>)));
        GD_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        if (!capture) continue;
        GD_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        try {
          enqueue_fused(sl, cs);
          e = hipGetLastError();
        } catch (const Error &) {
          e = hipErrorUnknown;
        }
        hipGraph_t g = nullptr;
        const hipError_t end = hipStreamEndCapture(cs, &g);
        GD_HIP(e);
        GD_HIP(end);
        e = hipGraphInstantiate(&sl.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        GD_HIP(e);
      }
    } catch (...) {
      (void)hipStreamDestroy(cs);
      release_graphs();
      throw;
    }
    (void)hipStreamDestroy(cs);
  }

  void release_graphs() {
    for (GraphSlot &sl : slots_) {
      if (sl.done) {
        (void)hipEventSynchronize(sl.done);
        (void)hipEventDestroy(sl.done);
      }
      if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
      if (sl.host) (void)hipHostFree(sl.host);
      (void)hipFree(sl.dev);
    }
    slots_.clear();
  }

  // Runs a small step as a fused step (descriptor copy, admission, query
  // kernel; one graph launch in graph mode); false (nothing enqueued) when
  // the step does not fit a slot or the next slot is still in use.
  bool launch_graph_step(Step &st, const std::vector<Piece> &upd,
                         const std::vector<QueryPiece> &look,
                         const std::vector<QueryPiece> &rng, u64 n_upd,
                         u64 n_look, u64 n_rng, int mode) {
    if (slots_.empty() || upd.size() > kGraphPieces ||
        look.size() > kGraphPieces || rng.size() > kGraphPieces)
      return false;
    if (n_upd > std::min<u64>(lsm_.small_batch_limit(), T::kSmallTile))
      return false;
    if (n_look >= lsm_.grouped_lookup_min()) return false;
    if (n_upd && lsm_.failed()) return false;  // the regular path reports
    GraphSlot &sl = slots_[next_slot_];
    if (sl.used) {
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) return false;
      GD_HIP(q);
    }
    StepDesc<T> &d = *sl.host;
    typename LsmT<T>::Run run{};
    if (n_upd) run = lsm_.alloc_small_run(static_cast<u32>(n_upd));
    if (!lsm_.snapshot_nowait(d.snap, n_upd ? &run : nullptr)) {
      if (n_upd) lsm_.release_run(run);
      return false;
    }
    d.n_upd = n_upd;
    d.np_upd = static_cast<u32>(upd.size());
    d.mode = mode;
    d.run_rows = run.rows;
    d.run_meta = n_upd ? run.meta() : nullptr;
    d.run_dir = n_upd ? run.entries() : nullptr;
    d.run_max_bits = run.max_bits;
    d.rows_log = lsm_.dir_rows_log();
    d.n_look = n_look;
    d.n_rng = n_rng;
    d.np_look = static_cast<u32>(look.size());
    d.np_rng = static_cast<u32>(rng.size());
    d.ticket = 0;
    std::copy(upd.begin(), upd.end(), d.upd);
    std::copy(look.begin(), look.end(), d.look);
    std::copy(rng.begin(), rng.end(), d.rng);
    // the step's admission waits for the previous step's queries, as in
    // the regular path
    if (!inflight_.empty())
      GD_HIP(hipStreamWaitEvent(us_, inflight_.back().queries_done, 0));
    if (sl.exec) GD_HIP(hipGraphLaunch(sl.exec, us_));
    else enqueue_fused(sl, us_);
    GD_LAUNCH_CHECK();
    GD_HIP(hipEventRecord(sl.done, us_));
    sl.used = true;
    next_slot_ = (next_slot_ + 1) % kGraphSlots;
    if (n_upd) lsm_.adopt_run(run, us_);
    GD_HIP(hipEventCreateWithFlags(&st.updates_done, hipEventDisableTiming));
    GD_HIP(hipEventRecord(st.updates_done, us_));
    // query-stream order covers the graph's reads (released runs are kept
    // until the query stream passes them)
    GD_HIP(hipStreamWaitEvent(qs_, st.updates_done, 0));
    GD_HIP(hipEventCreateWithFlags(&st.queries_done, hipEventDisableTiming));
    GD_HIP(hipEventRecord(st.queries_done, qs_));
    lsm_.retire_pending();
    ++graph_steps_;
    return true;
  }

  void launch_step(std::vector<std::shared_ptr<Request>> reqs, u64 staged) {
    Step st;
    st.id = ++step_counter_;
    st.staged_bytes = staged;
    std::vector<Piece> upd;
    std::vector<QueryPiece> look, rng;
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
        look.push_back({r->a, nullptr, r->out, r->out_found, n_look, 0, 0});
        n_look += r->count;
      } else {
        rng.push_back({r->a, r->b, r->out, nullptr, n_rng,
                       r->op == GD_OP_RANGE_COUNT ? 1u : 0u, 0});
        n_rng += r->count;
      }
    }
    const int step_mode = any_erase ? (any_insert ? lsm::kModeMixed
                                                  : lsm::kModeErase)
                                    : lsm::kModeInsert;
    if (!(n_rng && lsm_.range_compaction_pays(n_rng)) &&
        launch_graph_step(st, upd, look, rng, n_upd, n_look, n_rng,
                          step_mode)) {
      st.reqs = std::move(reqs);
      std::lock_guard<std::mutex> lk(mu_);
      coalesced_updates_ += upd.size();
      coalesced_queries_ += look.size() + rng.size();
      in_flight_bytes_ += st.staged_bytes;
      peak_staged_ = std::max(peak_staged_, in_flight_bytes_);
      inflight_.push_back(std::move(st));
      return;
    }
    // one pinned table per step: [update pieces | lookups | ranges]
    static_assert(sizeof(Piece) % 8 == 0 && sizeof(QueryPiece) % 8 == 0,
                  "piece tables are packed back to back");
    const std::size_t upd_bytes = upd.size() * sizeof(Piece);
    const std::size_t look_bytes = look.size() * sizeof(QueryPiece);
    const std::size_t table_bytes =
        upd_bytes + look_bytes + rng.size() * sizeof(QueryPiece);
    char *dev_table = nullptr;
    if (table_bytes) {
      st.host_pieces = static_cast<char *>(pinned_.get(table_bytes));
      std::memcpy(st.host_pieces, upd.data(), upd_bytes);
      std::memcpy(st.host_pieces + upd_bytes, look.data(), look_bytes);
      std::memcpy(st.host_pieces + upd_bytes + look_bytes, rng.data(),
                  rng.size() * sizeof(QueryPiece));
      dev_table = scratch<char>(st, table_bytes);
      GD_HIP(hipMemcpyAsync(dev_table, st.host_pieces, table_bytes,
                            hipMemcpyHostToDevice, us_));
    }
    const Piece *dev_upd = reinterpret_cast<const Piece *>(dev_table);
    const QueryPiece *dev_look =
        reinterpret_cast<const QueryPiece *>(dev_table + upd_bytes);
    const QueryPiece *dev_rng = reinterpret_cast<const QueryPiece *>(
        dev_table + upd_bytes + look_bytes);

    // ------------------------------------------------------------ updates
    // Admission only: the merges it triggers run on the merge stream and
    // neither this step's queries nor later admissions wait for them.
    // Admissions do wait for the previous step's queries: a sorting launch
    // whose workgroups chain their offsets (and take a whole CU's LDS)
    // next to a query launch of tens of thousands of workgroups gets CUs
    // only as the query workgroups drain, and both then run many times
    // slower than back to back.
    if (n_upd && !inflight_.empty())
      GD_HIP(hipStreamWaitEvent(us_, inflight_.back().queries_done, 0));
    if (n_upd) {
      K *keys = scratch<K>(st, n_upd);
      V *vals = scratch<V>(st, n_upd);
      u8 *flags = scratch<u8>(st, n_upd);
      gather_kernel<K, V><<<grid_for(n_upd, 256, 1u << 14), 256, 0, us_>>>(
          dev_upd, static_cast<u32>(upd.size()), n_upd, keys, vals, flags);
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
    }
    GD_HIP(hipEventCreateWithFlags(&st.updates_done, hipEventDisableTiming));
    GD_HIP(hipEventRecord(st.updates_done, us_));
    GD_HIP(hipStreamWaitEvent(qs_, st.updates_done, 0));

    // ------------------------------------------------------------ queries
    // the runs after this step's admissions; merges still in flight read
    // as their inputs.  A large range batch over a stack with a large share
    // of newer rows first compacts the stack (and waits for it), as the
    // plain range call does.
    const bool compact = n_rng && lsm_.range_compaction_pays(n_rng);
    if (compact) lsm_.compact_async(us_);
    const RunSet<T> snap = lsm_.snapshot_for(qs_, compact);
    if (n_look) {
      const PieceLookup<T> io{dev_look, static_cast<u32>(look.size())};
      if (n_look >= lsm_.grouped_lookup_min() && snap.count) {
        K *q = scratch<K>(st, n_look);
        K *qk = scratch<K>(st, n_look);
        u32 *dest = scratch<u32>(st, n_look);
        auto *res = scratch<lsm::GroupedResult<T>>(st, n_look);
        u32 *ctrl = scratch<u32>(st, lsm::SpreadLayout<K>(n_look).words);
        query_gather_kernel<K><<<grid_for(n_look, 256, 1u << 14), 256, 0,
                                  qs_>>>(
            dev_look, static_cast<u32>(look.size()), n_look, q);
        lsm::partition_by_spread<K>(q, static_cast<u32>(n_look), qk, nullptr,
                                    dest, rs::PayloadPos{}, ctrl, qs_,
                                    lsm_.tuning().sort_block);
        lsm::lookup_grouped_kernel<T><<<lsm::grouped_lookup_blocks(n_look), 256,
                                     0, qs_>>>(qk, n_look, res, snap);
        GD_LAUNCH_CHECK();
        lsm::lookup_gather_kernel<T><<<lsm::grouped_lookup_blocks(n_look), 256, 0,
                                    qs_>>>(io, dest, n_look, res);
      } else {
        lsm::lookup_kernel<T><<<grid_for(n_look, 256, 1u << 16), 256, 0, qs_>>>(
            io, n_look, snap);
      }
      GD_LAUNCH_CHECK();
    }
    if (n_rng) {
      // counts and sums share the kernel: each request chooses its own
      const u64 waves = (n_rng + lsm::range_queries_per_wave(snap.count) - 1) /
                        lsm::range_queries_per_wave(snap.count);
      lsm::range_kernel<T><<<grid_for(waves, lsm::kRangeWaves, 1u << 16),
                             lsm::kRangeThreads, 0, qs_>>>(
          PieceRange<T>{dev_rng, static_cast<u32>(rng.size())}, n_rng, snap);
      GD_LAUNCH_CHECK();
    }
    GD_HIP(hipEventCreateWithFlags(&st.queries_done, hipEventDisableTiming));
    GD_HIP(hipEventRecord(st.queries_done, qs_));
    // runs released by this step's merges stay until the queries that may
    // read them (this step's and earlier) are done
    lsm_.retire_pending();
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

  // Stamps the requests of `st` that `updates` selects as complete.
  void finish_requests(Step &st, bool updates, Clock::time_point now) {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto &r : st.reqs) {
      if ((r->op <= GD_OP_ERASE) != updates) continue;
      r->completed = now;
      r->status = updates ? st.update_status : GD_OK;
      r->done = true;
      latencies_.push_back(
          std::chrono::duration<double, std::micro>(now - r->submitted)
              .count());
      ++completed_;
      --busy_;
    }
  }

  void complete_steps() {
    // updates complete once admitted (later queries see them from then on,
    // whether or not the merges they triggered have finished)
    for (auto &st : inflight_) {
      if (st.updates_stamped) continue;
      const hipError_t q = hipEventQuery(st.updates_done);
      if (q == hipErrorNotReady) break;
      GD_HIP(q);
      finish_requests(st, true, Clock::now());
      st.updates_stamped = true;
    }
    bool any = false;
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
      if (!st.updates_stamped) {
        finish_requests(st, true, now);
        st.updates_stamped = true;
      }
      finish_requests(st, false, now);
      {
        std::lock_guard<std::mutex> lk(mu_);
        in_flight_bytes_ -= st.staged_bytes;
        ++steps_done_;
      }
      any = true;
      release_step(st);
      inflight_.pop_front();
    }
    done_cv_.notify_all();
    if (any) {
      lsm_.reclaim();
      gd_memory_stats m{};
      lsm_.memory(&m);
      const u64 held = m.held_bytes + pool_.held_bytes();
      resident_bytes_.store(held);
      live_bytes_.store(m.live_bytes + pool_.live_bytes());
      if (held > peak_resident_.load()) peak_resident_.store(held);
    }
  }

  void release_step(Step &st) {
    for (void *p : st.scratch) pool_.put(p);
    st.scratch.clear();
    pinned_.put(st.host_pieces);
    st.host_pieces = nullptr;
    if (st.queries_done) (void)hipEventDestroy(st.queries_done);
    if (st.updates_done) (void)hipEventDestroy(st.updates_done);
    st.queries_done = nullptr;
    st.updates_done = nullptr;
  }

  gd_serve_config cfg_;
  LsmT<T> lsm_;
  DevicePool pool_;  // step staging, released when the step completes
  PinnedCache pinned_;
  std::atomic<u64> resident_bytes_{0};
  std::atomic<u64> live_bytes_{0};
  std::atomic<u64> peak_resident_{0};  // written by the completion thread only
  hipStream_t us_ = nullptr, ms_ = nullptr, qs_ = nullptr;
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
  std::vector<GraphSlot> slots_;
  u32 next_slot_ = 0;
  u64 graph_steps_ = 0;
  std::thread worker_;
};

}  // namespace serve
}  // namespace gdict

using gdict::guarded;

struct gd_server : gdict::HandleBase {
  std::unique_ptr<gdict::serve::ServerBase> impl;
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
    auto h = std::make_unique<gd_server>();
    if (config->key_bytes == 0 || config->key_bytes == 4) {
      h->impl = std::make_unique<
          gdict::serve::Server<gdict::lsm::KV<gdict::u32>>>(*config);
    } else if (config->key_bytes == 8) {
      h->impl = std::make_unique<
          gdict::serve::Server<gdict::lsm::KV<gdict::u64>>>(*config);
    } else {
      throw gdict::Error(GD_INVALID_ARGUMENT, "key_bytes must be 4 or 8");
    }
    *out = h.release();
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
    *out_ticket = h->impl->submit(op, a, b, count, out, out_found);
  });
}

gd_status gd_serve_wait(gd_server *h, uint64_t ticket, uint64_t *out_step) {
  if (!h) return GD_INVALID_ARGUMENT;
  gd_status result = GD_OK;
  gd_status st =
      guarded(h, [&] { result = h->impl->wait(ticket, out_step); });
  if (st != GD_OK) return st;
  if (result != GD_OK)
    h->last_error = "the request's update step failed (status " +
                    std::to_string(result) + ")";
  return result;
}

gd_status gd_serve_drain(gd_server *h) {
  if (!h) return GD_INVALID_ARGUMENT;
  return guarded(h, [&] { h->impl->drain(); });
}

gd_status gd_serve_get_stats(gd_server *h, gd_serve_stats *out, int reset) {
  if (!h || !out) return GD_INVALID_ARGUMENT;
  return guarded(h, [&] {
    h->impl->stats(out);
    if (reset) h->impl->reset_stats();
  });
}

}  // extern "C"
