// C ABI of the LSM dictionary (gd_lsm_*); the runtime is lsm_runtime.hpp.
#include <array>
#include <string>

#include "lsm_runtime.hpp"

namespace gdict {

namespace {
thread_local std::string g_create_error;
}
void set_create_error(const std::string &m) { g_create_error = m; }
const char *create_error() { return g_create_error.c_str(); }

}  // namespace gdict

using gdict::guarded;
using gdict::lsm::Lsm;
using gdict::lsm::Lsm64;

struct gd_lsm : Lsm {
  using Lsm::Lsm;
  // This handle's device copy of `scratch` (gd_lsm_lookup_noalloc), looked
  // up at creation; null until a lookup succeeds.
  void *noalloc_scratch = nullptr;
};

struct gd_lsm64 : Lsm64 {
  using Lsm64::Lsm64;
};

namespace {

template <class H>
gd_status dump_sorted(H *h, typename H::K *out_keys, typename H::V *out_values,
                      uint64_t capacity, uint64_t *out_count, hipStream_t s) {
  if (!h || !out_count) return GD_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(h->mutex);
  uint64_t n = 0;
  gd_status st =
      guarded(h, [&] { n = h->dump(out_keys, out_values, capacity, s); });
  if (st != GD_OK) return st;
  *out_count = n;
  if (n > capacity) {
    h->last_error = "dump capacity " + std::to_string(capacity) +
                    " is smaller than the " + std::to_string(n) + " live keys";
    return GD_LENGTH;
  }
  return GD_OK;
}

// Scratch of gd_lsm_lookup_noalloc: the grouped lookup path's partition
// and result buffers for up to kNoAllocMaxQueries queries (16.6 MiB at
// 2^20).  Every GPU the library runs on holds one copy for the life of the
// process, at its own address; every allocation-free call on that GPU
// shares it, so such calls that overlap on one GPU race.
constexpr uint64_t kNoAllocMaxQueries = uint64_t(1) << 20;
constexpr std::size_t kNoAllocScratchBytes = std::size_t(17) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Device address of `scratch` per device id.  In host code the array's name
// is only a stand-in for the symbol; kernels receive this address.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
std::array<void *, kMaxDevices> g_scratch_ptr{};

// The address of the current device's copy of `scratch`, looked up once per
// device (hipGetSymbolAddress with that device current).
hipError_t current_device_scratch(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace

extern "C" {

const char *gd_last_error_message(const void *handle) {
  if (!handle) return gdict::create_error();
  return static_cast<const gdict::HandleBase *>(handle)->last_error.c_str();
}

gd_status gd_lsm_create(const gd_config *config, gd_lsm **out) {
  if (!config || !out) {
    gdict::set_create_error("null config or output handle");
    return GD_INVALID_ARGUMENT;
  }
  *out = nullptr;
  gd_status st = guarded(nullptr, [&] {
    (void)Lsm::capacity_for(*config);  // validate before touching the GPU
    *out = new gd_lsm(*config);
  });
  if (st == GD_OK) {
    // set-up of the allocation-free lookup on the creating (current)
    // device; a failure here is reported by that lookup, which retries
    void *p = nullptr;
    if (current_device_scratch(&p) == hipSuccess) (*out)->noalloc_scratch = p;
    else (void)hipGetLastError();
  }
  return st;
}

gd_status gd_lsm_destroy(gd_lsm *h) {
  delete h;
  return GD_OK;
}

#define GD_LSM_CALL(h, body)                                  \
  do {                                                        \
    if (!(h)) return GD_INVALID_ARGUMENT;                     \
    std::lock_guard<std::mutex> lock__((h)->mutex);           \
    return guarded((h), [&] { body; });                       \
  } while (0)

gd_status gd_lsm_clear(gd_lsm *h, hipStream_t s) { GD_LSM_CALL(h, h->clear(s)); }

gd_status gd_lsm_bulk_build(gd_lsm *h, const uint32_t *keys,
                            const uint32_t *values, uint64_t count,
                            hipStream_t s) {
  GD_LSM_CALL(h, h->bulk_build(keys, values, count, s));
}

gd_status gd_lsm_insert(gd_lsm *h, const uint32_t *keys,
                        const uint32_t *values, uint64_t count,
                        hipStream_t s) {
  GD_LSM_CALL(h, h->admit(keys, values, nullptr, count,
                          gdict::lsm::kModeInsert, s));
}

gd_status gd_lsm_erase(gd_lsm *h, const uint32_t *keys, uint64_t count,
                       hipStream_t s) {
  GD_LSM_CALL(h, h->admit(keys, nullptr, nullptr, count,
                          gdict::lsm::kModeErase, s));
}

gd_status gd_lsm_apply_mixed(gd_lsm *h, const uint32_t *keys,
                             const uint32_t *values, const uint8_t *is_erase,
                             uint64_t count, hipStream_t s) {
  GD_LSM_CALL(h, h->admit(keys, values, is_erase, count,
                          gdict::lsm::kModeMixed, s));
}

gd_status gd_lsm_lookup(gd_lsm *h, const uint32_t *queries, uint64_t count,
                        uint32_t *out_values, uint8_t *out_found,
                        int quotients_grouped, hipStream_t s) {
  (void)quotients_grouped;
  GD_LSM_CALL(h, h->lookup(queries, count, out_values, out_found, s));
}

gd_status gd_lsm_lookup_noalloc(gd_lsm *h, const uint32_t *queries,
                                uint64_t count, uint32_t *out_values,
                                uint8_t *out_found, hipStream_t s) {
  GD_LSM_CALL(h, {
    if (count > kNoAllocMaxQueries)
      throw gdict::Error(
          GD_INVALID_ARGUMENT,
          "gd_lsm_lookup_noalloc takes at most " +
              std::to_string(kNoAllocMaxQueries) + " queries, got " +
              std::to_string(count));
    if (!h->noalloc_scratch) {
      void *p = nullptr;
      GD_HIP(current_device_scratch(&p));
      h->noalloc_scratch = p;
    }
    h->lookup_in_scratch(queries, count, out_values, out_found,
                         static_cast<char *>(h->noalloc_scratch),
                         kNoAllocScratchBytes, s);
  });
}

gd_status gd_lsm_range_sum(gd_lsm *h, const uint32_t *lo, const uint32_t *hi,
                           uint64_t count, uint32_t *out_sums,
                           hipStream_t s) {
  GD_LSM_CALL(h, h->range(lo, hi, count, out_sums, false, s));
}

gd_status gd_lsm_range_count(gd_lsm *h, const uint32_t *lo,
                             const uint32_t *hi, uint64_t count,
                             uint32_t *out_counts, hipStream_t s) {
  GD_LSM_CALL(h, h->range(lo, hi, count, out_counts, true, s));
}

gd_status gd_lsm_successor(gd_lsm *h, const uint32_t *queries, uint64_t count,
                           uint32_t *out_keys, hipStream_t s) {
  GD_LSM_CALL(h, h->successor(queries, count, out_keys, s));
}

gd_status gd_lsm_dump_sorted(gd_lsm *h, uint32_t *out_keys,
                             uint32_t *out_values, uint64_t capacity,
                             uint64_t *out_count, hipStream_t s) {
  return dump_sorted(h, out_keys, out_values, capacity, out_count, s);
}

gd_status gd_lsm_compact(gd_lsm *h, hipStream_t s) { GD_LSM_CALL(h, h->compact(s)); }

gd_status gd_lsm_carry_status(gd_lsm *h, uint32_t *out_status) {
  if (!out_status) return GD_INVALID_ARGUMENT;
  GD_LSM_CALL(h, *out_status = h->carry_status());
}

gd_status gd_lsm_resident_bytes(gd_lsm *h, uint64_t *out_bytes) {
  if (!out_bytes) return GD_INVALID_ARGUMENT;
  GD_LSM_CALL(h, *out_bytes = h->resident_bytes());
}

gd_status gd_lsm_memory(gd_lsm *h, gd_memory_stats *out) {
  if (!out) return GD_INVALID_ARGUMENT;
  GD_LSM_CALL(h, h->memory(out));
}

gd_status gd_lsm_trim(gd_lsm *h) { GD_LSM_CALL(h, h->trim()); }

gd_status gd_lsm_set_tuning(gd_lsm *h, const gd_lsm_tuning *tuning) {
  if (!tuning) return GD_INVALID_ARGUMENT;
  GD_LSM_CALL(h, h->set_tuning(*tuning));
}

gd_status gd_lsm_get_tuning(gd_lsm *h, gd_lsm_tuning *out) {
  if (!out) return GD_INVALID_ARGUMENT;
  GD_LSM_CALL(h, *out = h->tuning());
}

gd_status gd_lsm_set_profiling(gd_lsm *h, int enable) {
  GD_LSM_CALL(h, h->set_profiling(enable != 0));
}

gd_status gd_lsm_read_profile(gd_lsm *h, gd_lsm_profile *out, int reset) {
  if (!out) return GD_INVALID_ARGUMENT;
  GD_LSM_CALL(h, h->read_profile(out, reset != 0));
}

gd_status gd_lsm_shape(gd_lsm *h, uint32_t *out_runs, uint64_t *out_rows_ub) {
  if (!out_runs || !out_rows_ub) return GD_INVALID_ARGUMENT;
  GD_LSM_CALL(h, h->shape(out_runs, out_rows_ub));
}

gd_status gd_lsm_snapshot_save(gd_lsm *h, const char *path,
                               uint64_t *out_count, hipStream_t s) {
  GD_LSM_CALL(h, {
    const uint64_t n = h->save(path, s);
    if (out_count) *out_count = n;
  });
}

gd_status gd_lsm_snapshot_restore(gd_lsm *h, const char *path,
                                  uint64_t *out_count, hipStream_t s) {
  GD_LSM_CALL(h, {
    const uint64_t n = h->restore(path, s);
    if (out_count) *out_count = n;
  });
}

// ------------------------------------------------------------ 64-bit LSM
gd_status gd_lsm64_create(const gd_config *config, gd_lsm64 **out) {
  if (!config || !out) {
    gdict::set_create_error("null config or output handle");
    return GD_INVALID_ARGUMENT;
  }
  *out = nullptr;
  return guarded(nullptr, [&] {
    (void)Lsm64::capacity_for(*config);
    *out = new gd_lsm64(*config);
  });
}

gd_status gd_lsm64_destroy(gd_lsm64 *h) {
  delete h;
  return GD_OK;
}

gd_status gd_lsm64_clear(gd_lsm64 *h, hipStream_t s) {
  GD_LSM_CALL(h, h->clear(s));
}

gd_status gd_lsm64_bulk_build(gd_lsm64 *h, const uint64_t *keys,
                              const uint64_t *values, uint64_t count,
                              hipStream_t s) {
  GD_LSM_CALL(h, h->bulk_build(keys, values, count, s));
}

gd_status gd_lsm64_insert(gd_lsm64 *h, const uint64_t *keys,
                          const uint64_t *values, uint64_t count,
                          hipStream_t s) {
  GD_LSM_CALL(h, h->admit(keys, values, nullptr, count,
                          gdict::lsm::kModeInsert, s));
}

gd_status gd_lsm64_erase(gd_lsm64 *h, const uint64_t *keys, uint64_t count,
                         hipStream_t s) {
  GD_LSM_CALL(h, h->admit(keys, nullptr, nullptr, count,
                          gdict::lsm::kModeErase, s));
}

gd_status gd_lsm64_apply_mixed(gd_lsm64 *h, const uint64_t *keys,
                               const uint64_t *values,
                               const uint8_t *is_erase, uint64_t count,
                               hipStream_t s) {
  GD_LSM_CALL(h, h->admit(keys, values, is_erase, count,
                          gdict::lsm::kModeMixed, s));
}

gd_status gd_lsm64_lookup(gd_lsm64 *h, const uint64_t *queries,
                          uint64_t count, uint64_t *out_values,
                          uint8_t *out_found, hipStream_t s) {
  GD_LSM_CALL(h, h->lookup(queries, count, out_values, out_found, s));
}

gd_status gd_lsm64_range_sum(gd_lsm64 *h, const uint64_t *lo,
                             const uint64_t *hi, uint64_t count,
                             uint64_t *out_sums, hipStream_t s) {
  GD_LSM_CALL(h, h->range(lo, hi, count, out_sums, false, s));
}

gd_status gd_lsm64_range_count(gd_lsm64 *h, const uint64_t *lo,
                               const uint64_t *hi, uint64_t count,
                               uint64_t *out_counts, hipStream_t s) {
  GD_LSM_CALL(h, h->range(lo, hi, count, out_counts, true, s));
}

gd_status gd_lsm64_successor(gd_lsm64 *h, const uint64_t *queries,
                             uint64_t count, uint64_t *out_keys,
                             hipStream_t s) {
  GD_LSM_CALL(h, h->successor(queries, count, out_keys, s));
}

gd_status gd_lsm64_dump_sorted(gd_lsm64 *h, uint64_t *out_keys,
                               uint64_t *out_values, uint64_t capacity,
                               uint64_t *out_count, hipStream_t s) {
  return dump_sorted(h, out_keys, out_values, capacity, out_count, s);
}

gd_status gd_lsm64_compact(gd_lsm64 *h, hipStream_t s) {
  GD_LSM_CALL(h, h->compact(s));
}

gd_status gd_lsm64_carry_status(gd_lsm64 *h, uint32_t *out_status) {
  if (!out_status) return GD_INVALID_ARGUMENT;
  GD_LSM_CALL(h, *out_status = h->carry_status());
}

gd_status gd_lsm64_resident_bytes(gd_lsm64 *h, uint64_t *out_bytes) {
  if (!out_bytes) return GD_INVALID_ARGUMENT;
  GD_LSM_CALL(h, *out_bytes = h->resident_bytes());
}

gd_status gd_lsm64_memory(gd_lsm64 *h, gd_memory_stats *out) {
  if (!out) return GD_INVALID_ARGUMENT;
  GD_LSM_CALL(h, h->memory(out));
}

gd_status gd_lsm64_trim(gd_lsm64 *h) { GD_LSM_CALL(h, h->trim()); }

gd_status gd_lsm64_set_tuning(gd_lsm64 *h, const gd_lsm_tuning *tuning) {
  if (!tuning) return GD_INVALID_ARGUMENT;
  GD_LSM_CALL(h, h->set_tuning(*tuning));
}

gd_status gd_lsm64_get_tuning(gd_lsm64 *h, gd_lsm_tuning *out) {
  if (!out) return GD_INVALID_ARGUMENT;
  GD_LSM_CALL(h, *out = h->tuning());
}

gd_status gd_lsm64_snapshot_save(gd_lsm64 *h, const char *path,
                                 uint64_t *out_count, hipStream_t s) {
  GD_LSM_CALL(h, {
    const uint64_t n = h->save(path, s);
    if (out_count) *out_count = n;
  });
}

gd_status gd_lsm64_snapshot_restore(gd_lsm64 *h, const char *path,
                                    uint64_t *out_count, hipStream_t s) {
  GD_LSM_CALL(h, {
    const uint64_t n = h->restore(path, s);
    if (out_count) *out_count = n;
  });
}

gd_status gd_lsm64_shape(gd_lsm64 *h, uint32_t *out_runs,
                         uint64_t *out_rows_ub) {
  if (!out_runs || !out_rows_ub) return GD_INVALID_ARGUMENT;
  GD_LSM_CALL(h, h->shape(out_runs, out_rows_ub));
}

}  // extern "C"
