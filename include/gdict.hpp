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
