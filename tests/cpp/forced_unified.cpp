// C++ caller of the reference-shaped interface (include/gdict.hpp): the
// forced-unified validation scenarios of FliX/paper_lsmu_sweep.cu:441-700,
// written against GPULSMOpt / DictionaryConfig / Device*Batch exactly as a
// user of the reference would, with a host unordered_map as the expectation.
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "gdict.hpp"

#define HIPOK(x)                                                       \
  do {                                                                 \
    if ((x) != hipSuccess) throw std::runtime_error(#x);               \
  } while (0)

namespace {

constexpr std::uint32_t kNotFound = 0xffffffffu;

std::uint32_t sparse_validation_key(std::uint32_t q) {
  return (q << 16) | ((q * 40503u + 17u) & 0xffffu);
}

template <class T>
struct DevBuf {
  T *p = nullptr;
  explicit DevBuf(std::size_t n) { HIPOK(hipMalloc(&p, n * sizeof(T) + 4)); }
  ~DevBuf() { (void)hipFree(p); }
  void up(const T *h, std::size_t n) {
    HIPOK(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice));
  }
  std::vector<T> down(std::size_t n) {
    std::vector<T> h(n);
    HIPOK(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
  }
};

int failures = 0;

void report(const char *name, std::size_t live, std::size_t queries,
            std::uint64_t errors) {
  std::printf("FORCED_UNIFIED_VALIDATION scenario=%s live_keys=%zu "
              "queries=%zu errors=%llu\n",
              name, live, queries, (unsigned long long)errors);
  failures += errors != 0;
}

}  // namespace

int main() {
  DictionaryConfig cfg;
  cfg.max_elements = 1u << 20;
  cfg.batch_capacity = 1u << 16;
  cfg.level_zero_capacity = (1u << 16) * 16u;
  GPULSMOpt index(cfg);
  std::unordered_map<std::uint32_t, std::uint32_t> expected;
  std::vector<std::uint32_t> removed;
  DevBuf<std::uint32_t> keys(4096), values(4096);
  std::vector<std::uint32_t> hk(4096), hv(4096);

  auto validate = [&](const char *name) {
    std::vector<std::uint32_t> q, e;
    for (const auto &kv : expected) {
      q.push_back(kv.first);
      e.push_back(kv.second);
    }
    for (std::uint32_t k : removed) {
      q.push_back(k);
      e.push_back(kNotFound);
    }
    for (std::uint32_t s = 0; q.size() < expected.size() + removed.size() +
                                          4096u; ++s) {
      const std::uint32_t k = (25000u << 16) | (s & 0xffffu);
      if (expected.count(k)) continue;
      q.push_back(k);
      e.push_back(kNotFound);
    }
    DevBuf<std::uint32_t> dq(q.size()), dr(q.size());
    dq.up(q.data(), q.size());
    index.lookup({dq.p, q.size(), dr.p, nullptr}, nullptr);
    HIPOK(hipDeviceSynchronize());
    const auto got = dr.down(q.size());
    std::uint64_t errors = 0;
    for (std::size_t i = 0; i < q.size(); ++i) errors += got[i] != e[i];
    report(name, expected.size(), q.size(), errors);
  };

  for (std::uint32_t b = 0; b < 16; ++b) {
    for (std::uint32_t i = 0; i < 2048; ++i) {
      const std::uint32_t quotient = b * 2048 + i;
      hk[i] = sparse_validation_key(quotient);
      hv[i] = 0x01000000u + quotient;
      expected[hk[i]] = hv[i];
    }
    keys.up(hk.data(), 2048);
    values.up(hv.data(), 2048);
    index.insert({keys.p, values.p, 2048}, nullptr);
  }
  validate("sparse_unique");
  for (std::uint32_t b = 0; b < 16; ++b) {
    for (std::uint32_t i = 0; i < 4096; ++i) {
      const std::uint32_t suffix = b * 4096 + i;
      hk[i] = (30000u << 16) | suffix;
      hv[i] = 0x02000000u + suffix;
      expected[hk[i]] = hv[i];
    }
    keys.up(hk.data(), 4096);
    values.up(hv.data(), 4096);
    index.insert({keys.p, values.p, 4096}, nullptr);
  }
  validate("dense_unique");
  for (std::uint32_t b = 0; b < 16; ++b) {
    for (std::uint32_t i = 0; i < 4096; ++i) {
      hk[i] = sparse_validation_key(i % 1024);
      hv[i] = 0x10000000u + b * 4096 + i;
      expected[hk[i]] = hv[i];
    }
    keys.up(hk.data(), 4096);
    values.up(hv.data(), 4096);
    index.insert({keys.p, values.p, 4096}, nullptr);
  }
  validate("duplicate_newest_wins");
  for (std::uint32_t i = 0; i < 1024; ++i) {
    hk[i] = sparse_validation_key(i);
    removed.push_back(hk[i]);
    expected.erase(hk[i]);
  }
  keys.up(hk.data(), 1024);
  for (int b = 0; b < 16; ++b) index.erase({keys.p, 1024}, nullptr);
  validate("tombstone_visibility");

  {  // multi-level rollover
    DictionaryConfig rc;
    rc.max_elements = 512;
    rc.batch_capacity = 4;
    rc.level_zero_capacity = 64;
    GPULSMOpt r(rc);
    DevBuf<std::uint32_t> rk(256), rv(256);
    std::array<std::uint32_t, 256> exp{};
    std::array<std::uint32_t, 4> ck{}, cv{};
    for (std::uint32_t b = 0; b < 512; ++b) {
      for (std::uint32_t i = 0; i < 4; ++i) {
        const std::uint32_t ord = b * 4 + i;
        ck[i] = ord % 256;
        cv[i] = 0x28000000u + ord;
        exp[ck[i]] = cv[i];
      }
      rk.up(ck.data(), 4);
      rv.up(cv.data(), 4);
      r.insert({rk.p, rv.p, 4}, nullptr);
    }
    std::array<std::uint32_t, 256> q{};
    for (std::uint32_t i = 0; i < 256; ++i) q[i] = i;
    rk.up(q.data(), 256);
    r.lookup({rk.p, 256, rv.p, nullptr}, nullptr);
    HIPOK(hipDeviceSynchronize());
    const auto got = rv.down(256);
    std::uint64_t errors = 0;
    for (int i = 0; i < 256; ++i) errors += got[i] != exp[i];
    const std::uint32_t status = r.canonical_carry_status();
    std::printf("CANONICAL_MULTILEVEL_ROLLOVER_VALIDATION status=%u "
                "errors=%llu\n", status, (unsigned long long)errors);
    failures += status != 0 || errors != 0;
  }
  {  // controlled capacity exhaustion
    DictionaryConfig cc;
    cc.max_elements = 1024;
    cc.batch_capacity = 4;
    cc.level_zero_capacity = 1024;
    GPULSMOpt c(cc);
    DevBuf<std::uint32_t> ck(4), cv(4);
    std::array<std::uint32_t, 4> hk4{}, hv4{};
    for (std::uint32_t b = 0; b < 17 * 16; ++b) {
      for (std::uint32_t i = 0; i < 4; ++i) {
        hk4[i] = b * 4 + i;
        hv4[i] = 0x30000000u + hk4[i];
      }
      ck.up(hk4.data(), 4);
      cv.up(hv4.data(), 4);
      c.insert({ck.p, cv.p, 4}, nullptr);
    }
    HIPOK(hipDeviceSynchronize());
    const std::uint32_t status = c.canonical_carry_status();
    const bool controlled_status =
        (status & GD_PUBLICATION_OUTPUT_OVERFLOW) != 0;
    bool controlled_exception = false;
    try {
      c.insert({ck.p, cv.p, 4}, nullptr);
    } catch (const std::runtime_error &e) {
      controlled_exception =
          std::string(e.what()).find("capacity") != std::string::npos;
    }
    std::printf("CANONICAL_CAPACITY_VALIDATION status=%u controlled_status=%d "
                "controlled_exception=%d\n", status, controlled_status,
                controlled_exception);
    failures += !controlled_status || !controlled_exception;
  }
  {  // invalid arguments surface as std::invalid_argument
    bool thrown = false;
    try {
      index.insert({nullptr, nullptr, 4}, nullptr);
    } catch (const std::invalid_argument &) {
      thrown = true;
    }
    std::printf("INVALID_ARGUMENT_VALIDATION thrown=%d\n", thrown);
    failures += !thrown;
  }
  std::printf("FAILURES %d\n", failures);
  return failures ? 1 : 0;
}
