// Host-side plumbing shared by the LSM and FliX runtimes: error transport
// across the C ABI, a caching device pool, and per-handle operation order.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <iterator>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "gdict.h"

namespace gdict {

// Internal exception carrying a gd_status; converted at the ABI boundary.
struct Error : std::runtime_error {
  gd_status status;
  Error(gd_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

inline void hip_check(hipError_t e, const char *what, const char *file,
                      int line) {
  if (e != hipSuccess) {
    (void)hipGetLastError();  // clear a sticky launch error
    throw Error(GD_HIP_ERROR, std::string(hipGetErrorString(e)) + " in " +
                                  what + " at " + file + ":" +
                                  std::to_string(line));
  }
}
#define GD_HIP(call) ::gdict::hip_check((call), #call, __FILE__, __LINE__)
#define GD_LAUNCH_CHECK() GD_HIP(hipGetLastError())

// Caching device allocator.  Buffers are recycled by size class.  All work
// of a handle is totally ordered on the device (event chaining), so a buffer
// released by host bookkeeping may be handed to work enqueued later without
// further synchronisation.
//
// The cache is bounded: when the cached (released, unused) bytes exceed
// max(kCacheFloor, live / 2), the largest cached blocks go back to HIP until
// the cache is at most max(kCacheFloor, live / 4).  hipFree synchronises the
// device before it frees, so a trim never frees memory that enqueued work
// still reads; it only costs that wait, and it happens when a large
// structure has been replaced (a merge of the biggest runs), not per batch.
// Bytes held by a group of pools of one handle, with their common peak.
struct PoolUsage {
  std::size_t held = 0, peak = 0;
};

class DevicePool {
 public:
  static constexpr std::size_t kCacheFloor = std::size_t(4) << 30;

  explicit DevicePool(PoolUsage *group = nullptr) : group_(group) {}
  DevicePool(const DevicePool &) = delete;
  DevicePool &operator=(const DevicePool &) = delete;
  ~DevicePool() { release_all(); }

  // Size classes: p/8 steps between p/2 and p (p = the next power of two)
  // up to 1 GiB, waste below 25%; p/128 steps above, waste below 1.6% (a
  // 2^31-row run of 16-byte rows plus one spare row takes 32.5 GiB, not
  // 36).  A request takes the
  // smallest cached block of its class up to 1/8 above it.
  static std::size_t size_class(std::size_t bytes) {
    if (bytes <= 4096) return 4096;
    std::size_t p = 4096;
    while (p < bytes) p <<= 1;
    const std::size_t q = p > (std::size_t(1) << 30) ? p >> 7 : p >> 3;
    std::size_t c = p >> 1;
    while (c < bytes) c += q;
    return c;
  }

  void *get(std::size_t bytes) {
    const std::size_t cls = size_class(bytes);
    auto it = free_.lower_bound(cls);
    if (it != free_.end() && it->first <= cls + cls / 8) {
      void *p = it->second;
      cached_ -= it->first;
      free_.erase(it);
      return p;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      // give cached blocks back and retry once
      trim();
      e = hipMalloc(&p, cls);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(GD_OUT_OF_MEMORY,
                    "device allocation of " + std::to_string(cls) +
                        " bytes failed (" + hipGetErrorString(e) + ")");
      }
    }
    size_of_[p] = cls;
    held_ += cls;
    if (held_ > peak_) peak_ = held_;
    if (group_) {
      group_->held += cls;
      group_->peak = std::max(group_->peak, group_->held);
    }
    return p;
  }

  template <class T>
  T *get_n(std::size_t n) {
    return static_cast<T *>(get(n * sizeof(T) ? n * sizeof(T) : 1));
  }

  void put(void *p) {
    if (!p) return;
    auto it = size_of_.find(p);
    if (it == size_of_.end()) return;
    free_.emplace(it->second, p);
    cached_ += it->second;
    const std::size_t live = held_ - cached_;
    if (cached_ > std::max(kCacheFloor, live / 2))
      trim_to(std::max(kCacheFloor, live / 4));
  }

  // Frees cached blocks, largest first, until at most `limit` bytes stay
  // cached.  hipFree waits for the device first.
  void trim_to(std::size_t limit) {
    while (cached_ > limit && !free_.empty()) {
      auto last = std::prev(free_.end());
      (void)hipFree(last->second);
      held_ -= last->first;
      cached_ -= last->first;
      if (group_) group_->held -= last->first;
      size_of_.erase(last->second);
      free_.erase(last);
      ++trims_;
    }
  }
  void trim() { trim_to(0); }

  void release_all() {
    for (auto &kv : size_of_) (void)hipFree(kv.first);
    size_of_.clear();
    free_.clear();
    if (group_) group_->held -= held_;
    held_ = 0;
    cached_ = 0;
  }

  std::size_t held_bytes() const { return held_; }
  std::size_t cached_bytes() const { return cached_; }
  std::size_t live_bytes() const { return held_ - cached_; }
  std::size_t peak_bytes() const { return peak_; }
  std::size_t trimmed_blocks() const { return trims_; }

 private:
  std::multimap<std::size_t, void *> free_;
  std::unordered_map<void *, std::size_t> size_of_;
  PoolUsage *group_ = nullptr;
  std::size_t held_ = 0, cached_ = 0, peak_ = 0, trims_ = 0;
};

// RAII lease of a pool buffer.
template <class T>
class Lease {
 public:
  Lease() = default;
  Lease(DevicePool &pool, std::size_t n) : pool_(&pool), p_(pool.get_n<T>(n)) {}
  Lease(const Lease &) = delete;
  Lease &operator=(const Lease &) = delete;
  Lease(Lease &&o) noexcept : pool_(o.pool_), p_(o.p_) { o.p_ = nullptr; }
  Lease &operator=(Lease &&o) noexcept {
    reset();
    pool_ = o.pool_;
    p_ = o.p_;
    o.p_ = nullptr;
    return *this;
  }
  ~Lease() { reset(); }
  void reset() {
    if (p_ && pool_) pool_->put(p_);
    p_ = nullptr;
  }
  T *get() const { return p_; }
  T *release() {
    T *p = p_;
    p_ = nullptr;
    return p;
  }

 private:
  DevicePool *pool_ = nullptr;
  T *p_ = nullptr;
};

// Event chaining that totally orders the work of one handle across streams.
class OperationOrder {
 public:
  OperationOrder() {
    GD_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
  }
  ~OperationOrder() {
    if (done_) {
      (void)hipEventSynchronize(done_);
      (void)hipEventDestroy(done_);
    }
  }
  // An operation on the stream the previous one ended on is ordered after
  // it by the stream itself; one on another stream waits for its event.
  void begin(hipStream_t s) {
    if (!ended_ || s != last_) GD_HIP(hipStreamWaitEvent(s, done_, 0));
  }
  void end(hipStream_t s) {
    GD_HIP(hipEventRecord(done_, s));
    last_ = s;
    ended_ = true;
  }
  void wait_host() { GD_HIP(hipEventSynchronize(done_)); }

 private:
  hipEvent_t done_ = nullptr;
  hipStream_t last_ = nullptr;
  bool ended_ = false;
};

// Base of every handle: last error message.
struct HandleBase {
  std::string last_error;
  std::mutex mutex;
};

void set_create_error(const std::string &m);
const char *create_error();

// Runs `body` and converts exceptions into a gd_status + message.
template <class F>
gd_status guarded(HandleBase *h, F &&body) {
  try {
    body();
    return GD_OK;
  } catch (const Error &e) {
    if (h) h->last_error = e.what(); else set_create_error(e.what());
    return e.status;
  } catch (const std::bad_alloc &e) {
    if (h) h->last_error = "bad_alloc"; else set_create_error("bad_alloc");
    return GD_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    if (h) h->last_error = e.what(); else set_create_error(e.what());
    return GD_HIP_ERROR;
  }
}

inline unsigned grid_for(std::uint64_t items, unsigned threads,
                         unsigned cap = 1u << 20) {
  std::uint64_t g = (items + threads - 1) / threads;
  if (g == 0) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

}  // namespace gdict
