// C++ mirror of the reference's host interfaces over the gdict C ABI.
//
// A caller of the reference switches by including this header instead of
// GPULSMOpt.cuh / gpu_dictionary_adapter.cuh: the batch structs, the
// DictionaryConfig fields, the class name, the member functions and the
// exception types are the reference's (GPULSMOpt.cuh:5599-5872,
// gpu_dictionary_adapter.cuh:1-49), with cudaStream_t spelled hipStream_t.
//
//   std::invalid_argument  <- GD_INVALID_ARGUMENT
//   std::bad_alloc         <- GD_OUT_OF_MEMORY
//   std::length_error      <- GD_LENGTH
//   std::runtime_error     <- GD_PUBLICATION_FAILED, GD_HIP_ERROR, GD_CAPACITY
#pragma once

#include <cstddef>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>

#include "gdict.h"

enum class DeleteOrderPolicy : std::uint8_t {
  adaptive,
  uniform_live,
};

enum class BaseDeleteValuePolicy : std::uint8_t {
  eager,
  mark_only,
};

struct DictionaryConfig {
  std::size_t max_elements = 0;
  std::size_t batch_capacity = 0;
  std::size_t level_zero_capacity = 0;
  DeleteOrderPolicy delete_order = DeleteOrderPolicy::adaptive;
  BaseDeleteValuePolicy base_delete_values = BaseDeleteValuePolicy::eager;
};

struct DeviceKeyValueBatch {
  const std::uint32_t *keys = nullptr;
  const std::uint32_t *values = nullptr;
  std::size_t count = 0;
};

struct DeviceLookupBatch {
  const std::uint32_t *queries = nullptr;
  std::size_t count = 0;
  std::uint32_t *out_values = nullptr;
  std::uint8_t *out_found = nullptr;
};

struct DeviceSuccessorBatch {
  const std::uint32_t *queries = nullptr;
  std::size_t count = 0;
  std::uint32_t *out_keys = nullptr;
};

struct DeviceRangeOutputBatch {
  const std::uint32_t *lo = nullptr;
  const std::uint32_t *hi = nullptr;
  std::size_t query_count = 0;
  std::uint32_t *out_sums = nullptr;
};

namespace gdict_cpp {

inline void throw_status(gd_status s, const void *handle) {
  if (s == GD_OK) return;
  const std::string msg = gd_last_error_message(handle);
  switch (s) {
    case GD_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case GD_OUT_OF_MEMORY: throw std::bad_alloc();
    case GD_LENGTH: throw std::length_error(msg);
    default: throw std::runtime_error(msg);
  }
}

}  // namespace gdict_cpp

class GPULSMOpt {
 public:
  struct DeviceKeyBatch {
    const std::uint32_t *keys = nullptr;
    std::size_t count = 0u;
  };

  explicit GPULSMOpt(const DictionaryConfig &config) {
    gd_config c{config.max_elements, config.batch_capacity,
                config.level_zero_capacity,
                static_cast<std::uint32_t>(config.delete_order),
                static_cast<std::uint32_t>(config.base_delete_values)};
    gdict_cpp::throw_status(gd_lsm_create(&c, &h_), nullptr);
  }
  GPULSMOpt(const GPULSMOpt &) = delete;
  GPULSMOpt &operator=(const GPULSMOpt &) = delete;
  ~GPULSMOpt() { gd_lsm_destroy(h_); }

  void clear(hipStream_t stream) { check(gd_lsm_clear(h_, stream)); }

  void bulk_build(const std::uint32_t *keys, const std::uint32_t *values,
                  std::size_t count, hipStream_t stream) {
    check(gd_lsm_bulk_build(h_, keys, values, count, stream));
  }

  void insert(const DeviceKeyValueBatch &b, hipStream_t stream) {
    check(gd_lsm_insert(h_, b.keys, b.values, b.count, stream));
  }

  void erase(const DeviceKeyBatch &b, hipStream_t stream) {
    check(gd_lsm_erase(h_, b.keys, b.count, stream));
  }

  void lookup(const DeviceLookupBatch &b, hipStream_t stream,
              bool quotients_grouped = false) {
    check(gd_lsm_lookup(h_, b.queries, b.count, b.out_values, b.out_found,
                        quotients_grouped ? 1 : 0, stream));
  }

  void range(const DeviceRangeOutputBatch &b, hipStream_t stream) {
    check(gd_lsm_range_sum(h_, b.lo, b.hi, b.query_count, b.out_sums, stream));
  }

  void successor(const DeviceSuccessorBatch &b, hipStream_t stream) {
    check(gd_lsm_successor(h_, b.queries, b.count, b.out_keys, stream));
  }

  std::uint32_t canonical_carry_status() const {
    std::uint32_t s = 0;
    check(gd_lsm_carry_status(h_, &s));
    return s;
  }

  std::size_t gpu_resident_bytes() const {
    std::uint64_t b = 0;
    check(gd_lsm_resident_bytes(h_, &b));
    return static_cast<std::size_t>(b);
  }

  gd_lsm *handle() const { return h_; }

 private:
  void check(gd_status s) const { gdict_cpp::throw_status(s, h_); }
  gd_lsm *h_ = nullptr;
};

// ----------------------------------------------------------------- FliX
// Mirror of the FliX updatable index, cg_rtx_index_updates
// (FliX/impl_cg_rtx_index_updates.cuh:1422-2693, the non-OptiX members a
// caller uses: build_bucket_layer_only :1711, rebuild_bucket_layer_only
// :1821, insert :2195, remove :2253, insert_and_remove :2321,
// lookups_ordered :2485, lookups_successor :2647, compute_total_size
// :1657, gpu_resident_bytes :1600, destroy :2679).  The template
// parameters are the reference's: exactly one of cache_line (5, 10, 20:
// half, one or two cache lines per node) and node_size_log (nodes of
// 2^node_size_log pairs) is non-zero; the node must hold 8 to 64 pairs
// here.  Values are smallsize (uint32) offsets as in the reference.
// Errors are thrown like GPULSMOpt's (the reference only prints them).
using smallsize = std::uint32_t;

template <typename key_type_, std::uint8_t cache_line = 0,
          std::uint8_t node_size_log = 0,
          std::uint8_t percentage_initial_fill = 50>
class cg_rtx_index_updates {
  static_assert(sizeof(key_type_) == 4 || sizeof(key_type_) == 8,
                "key_type must be 4 or 8 bytes");
  static_assert((cache_line == 0) + (node_size_log == 0) == 1,
                "Exactly one of cache_line or node_size_log must be 0");
  static_assert(percentage_initial_fill <= 100,
                "percentage_initial_fill cannot exceed 100");

 public:
  using key_type = key_type_;
  // pairs per node (FliX/definitions_updates.cuh:275-298)
  static constexpr smallsize node_size =
      cache_line == 0 ? (smallsize(1) << node_size_log)
      : sizeof(key_type) == 4
          ? (cache_line == 10 ? 14 : cache_line == 20 ? 30 : 6)
          : (cache_line == 10 ? 9 : cache_line == 20 ? 20 : 4);
  static_assert(node_size >= 8 && node_size <= 64,
                "nodes of 8 to 64 pairs are supported");

  cg_rtx_index_updates() = default;
  cg_rtx_index_updates(const cg_rtx_index_updates &) = delete;
  cg_rtx_index_updates &operator=(const cg_rtx_index_updates &) = delete;
  ~cg_rtx_index_updates() { destroy(); }

  // Values are the input positions; max_size bounds the keys the index
  // is sized for (the allocation region); available_memory_bytes is not
  // needed (device memory comes from the library's pool).
  void build_bucket_layer_only(const key_type *keys, std::size_t size,
                               std::size_t max_size,
                               std::size_t available_memory_bytes,
                               double *build_time_ms,
                               std::size_t *build_bytes) {
    (void)available_memory_bytes;
    destroy();
    gd_flix_config c{sizeof(key_type), sizeof(smallsize), node_size,
                     percentage_initial_fill,
                     0u};
    gdict_cpp::throw_status(gd_flix_create(&c, &h_), nullptr);
    timed(build_time_ms,
          [&] { check(gd_flix_build(h_, keys, size, max_size, nullptr)); });
    if (build_bytes) *build_bytes += gpu_resident_bytes();
  }

  void rebuild_bucket_layer_only(double *build_time_ms,
                                 std::size_t *build_bytes) {
    timed(build_time_ms, [&] { check(gd_flix_rebuild(h_, nullptr)); });
    if (build_bytes) *build_bytes += gpu_resident_bytes();
  }

  void insert(const key_type *update_list, const smallsize *offsets,
              std::size_t size, hipStream_t stream) {
    check(gd_flix_insert(h_, update_list, offsets, size, stream));
  }

  void remove(const key_type *update_list, std::size_t size,
              hipStream_t stream) {
    check(gd_flix_remove(h_, update_list, size, stream));
  }

  // Deletions first, then insertions (impl_cg_rtx_index_updates.cuh:2321).
  void insert_and_remove(const key_type *update_list,
                         const smallsize *offsets, std::size_t size,
                         const key_type *delete_list, hipStream_t stream) {
    remove(delete_list, size, stream);
    insert(update_list, offsets, size, stream);
  }

  // Ascending keys; result[j] belongs to the j-th key, all-ones on a miss.
  void lookups_ordered(const key_type *keys, smallsize *result,
                       std::size_t size, hipStream_t stream) {
    check(gd_flix_lookup_sorted(h_, keys, size, result, stream));
  }

  // Ascending keys; smallest stored key >= each, key_type(0xFFFFFFFF)
  // when none.
  void lookups_successor(const key_type *keys, key_type *result,
                         std::size_t size, hipStream_t stream) {
    check(gd_flix_successor_sorted(h_, keys, size, result, stream));
  }

  std::size_t compute_total_size(hipStream_t stream) {
    std::uint64_t n = 0;
    check(gd_flix_total_keys(h_, &n, stream));
    return static_cast<std::size_t>(n);
  }

  std::size_t gpu_resident_bytes() {
    std::uint64_t b = 0;
    check(gd_flix_resident_bytes(h_, &b));
    return static_cast<std::size_t>(b);
  }

  void destroy() {
    if (h_) gd_flix_destroy(h_);
    h_ = nullptr;
  }

  gd_flix *handle() const { return h_; }

 private:
  void check(gd_status s) const {
    if (!h_ && s == GD_INVALID_ARGUMENT)
      throw std::invalid_argument("cg_rtx_index_updates used before build");
    gdict_cpp::throw_status(s, h_);
  }

  template <class F>
  void timed(double *ms, F &&body) {
    if (!ms) {
      body();
      return;
    }
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
      throw std::runtime_error("hipEventCreate failed");
    (void)hipEventRecord(a, nullptr);
    try {
      body();
    } catch (...) {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
      throw;
    }
    (void)hipEventRecord(b, nullptr);
    (void)hipEventSynchronize(b);
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, a, b);
    *ms = t;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }

  gd_flix *h_ = nullptr;
};
