// C++ caller of the FliX mirror (include/gdict.hpp: cg_rtx_index_updates),
// written as a user of the reference's index would call it:
//   1. the interface test of FliX/main.cu:98-156 (build keys 100..10 with
//      values = input positions, inserts 5..105 step 10 with offsets
//      10..20, deletes {30, 35, 40, 45, 75, 80}, probes 5..105 step 5),
//      expectations derived by hand;
//   2. the reference driver's update loop shape
//      (FliX/benchmarks_updates.cuh:1534, 1698, 1829, 2049, 2129, 1929,
//      2487): build, insert_and_remove rounds, lookups_ordered,
//      lookups_successor, rebuild_bucket_layer_only and compute_total_size,
//      checked against a host std::map after every round, for 32- and
//      64-bit keys;
//   3. misuse: a call before the build throws std::invalid_argument.
// Prints one line per check and "FAILURES n" at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <map>
#include <random>
#include <stdexcept>
#include <vector>

#include "gdict.hpp"

#define HIPOK(x)                                                       \
  do {                                                                 \
    if ((x) != hipSuccess) throw std::runtime_error(#x);               \
  } while (0)

namespace {

int failures = 0;

void expect(bool ok, const char *what) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) ++failures;
}

template <class T>
struct DevBuf {
  T *p = nullptr;
  std::size_t n = 0;
  explicit DevBuf(const std::vector<T> &h) : n(h.size()) {
    HIPOK(hipMalloc(&p, std::max<std::size_t>(n, 1) * sizeof(T)));
    if (n) HIPOK(hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice));
  }
  explicit DevBuf(std::size_t count) : n(count) {
    HIPOK(hipMalloc(&p, std::max<std::size_t>(n, 1) * sizeof(T)));
  }
  ~DevBuf() { (void)hipFree(p); }
  std::vector<T> down() const {
    std::vector<T> h(n);
    HIPOK(hipDeviceSynchronize());
    if (n) HIPOK(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
  }
};

void interface_known_answer() {
  constexpr smallsize kNotFound = 0xffffffffu;
  cg_rtx_index_updates<std::uint32_t, 0, 5> index;
  const std::vector<std::uint32_t> build = {100, 90, 80, 70, 60,
                                            50,  40, 30, 20, 10};
  std::vector<std::uint32_t> ins, probes;
  std::vector<smallsize> offs;
  for (std::uint32_t k = 5, o = 10; k <= 105; k += 10, ++o) {
    ins.push_back(k);
    offs.push_back(o);
  }
  for (std::uint32_t k = 5; k <= 105; k += 5) probes.push_back(k);
  const std::vector<std::uint32_t> dels = {30, 35, 40, 45, 75, 80};
  DevBuf<std::uint32_t> d_build(build), d_ins(ins), d_probes(probes),
      d_dels(dels);
  DevBuf<smallsize> d_offs(offs), d_out(probes.size());
  double ms = -1.0;
  std::size_t bytes = 0;
  index.build_bucket_layer_only(d_build.p, build.size(), 21, 0, &ms, &bytes);
  expect(ms >= 0.0 && bytes > 0, "build reports its time and bytes");
  auto show = [&] {
    index.lookups_ordered(d_probes.p, d_out.p, probes.size(), nullptr);
    return d_out.down();
  };
  std::vector<smallsize> exp(probes.size());
  for (std::size_t i = 0; i < probes.size(); ++i)
    exp[i] = probes[i] % 10 ? kNotFound : (100 - probes[i]) / 10;
  expect(show() == exp, "known answer after the build");
  index.insert(d_ins.p, d_offs.p, ins.size(), nullptr);
  for (std::size_t i = 0; i < probes.size(); ++i)
    exp[i] = probes[i] % 10 == 0 ? (100 - probes[i]) / 10
                                 : 10 + (probes[i] - 5) / 10;
  expect(show() == exp, "known answer after the inserts");
  index.remove(d_dels.p, dels.size(), nullptr);
  for (std::size_t i = 0; i < probes.size(); ++i)
    if (std::find(dels.begin(), dels.end(), probes[i]) != dels.end())
      exp[i] = kNotFound;
  expect(show() == exp, "known answer after the deletes");
  expect(index.compute_total_size(nullptr) == 15, "total size 21 - 6");
}

template <class K>
void update_rounds(const char *name) {
  std::mt19937_64 rng(1234 + sizeof(K));
  const K top = sizeof(K) == 4 ? K(0xfffffff0u) : K(1) << 60;
  auto draw = [&] { return K(2 + rng() % (top - 2)); };
  cg_rtx_index_updates<K, 0, 5> index;
  std::map<K, smallsize> live;
  std::vector<K> build;
  while (build.size() < 50000) {
    const K k = draw();
    if (live.emplace(k, smallsize(build.size())).second) build.push_back(k);
  }
  DevBuf<K> d_build(build);
  index.build_bucket_layer_only(d_build.p, build.size(), 200000, 0, nullptr,
                                nullptr);
  bool ok = index.compute_total_size(nullptr) == live.size();
  for (int round = 0; round < 6; ++round) {
    // equal-sized lists: deletes of live keys first, then new keys
    const std::size_t n = 8000;
    std::vector<K> keys_live;
    for (auto &kv : live) keys_live.push_back(kv.first);
    std::shuffle(keys_live.begin(), keys_live.end(), rng);
    std::vector<K> del(keys_live.begin(), keys_live.begin() + n);
    std::sort(del.begin(), del.end());
    for (K k : del) live.erase(k);
    std::vector<K> ins;
    std::vector<smallsize> offs;
    while (ins.size() < n) {
      const K k = draw();
      if (live.count(k) ||
          std::find(ins.begin(), ins.end(), k) != ins.end())
        continue;
      ins.push_back(k);
      offs.push_back(smallsize(rng()));
    }
    for (std::size_t i = 0; i < n; ++i) live[ins[i]] = offs[i];
    DevBuf<K> d_ins(ins), d_del(del);
    DevBuf<smallsize> d_offs(offs);
    index.insert_and_remove(d_ins.p, d_offs.p, n, d_del.p, nullptr);
    if (round == 3) index.rebuild_bucket_layer_only(nullptr, nullptr);
    // sorted probes: half live keys, half random
    std::vector<K> q;
    std::vector<K> now;
    for (auto &kv : live) now.push_back(kv.first);
    for (int i = 0; i < 20000; ++i)
      q.push_back(i % 2 ? now[rng() % now.size()] : draw());
    std::sort(q.begin(), q.end());
    DevBuf<K> d_q(q), d_succ(q.size());
    DevBuf<smallsize> d_res(q.size());
    index.lookups_ordered(d_q.p, d_res.p, q.size(), nullptr);
    index.lookups_successor(d_q.p, d_succ.p, q.size(), nullptr);
    const std::vector<smallsize> res = d_res.down();
    const std::vector<K> succ = d_succ.down();
    for (std::size_t i = 0; i < q.size(); ++i) {
      auto it = live.find(q[i]);
      const smallsize ev = it == live.end() ? 0xffffffffu : it->second;
      auto lb = live.lower_bound(q[i]);
      const K es = lb == live.end() ? K(0xffffffffu) : lb->first;
      if (res[i] != ev || succ[i] != es) ok = false;
    }
    if (index.compute_total_size(nullptr) != live.size()) ok = false;
  }
  std::printf("FLIX_UPDATE_ROUNDS keys=%s live=%zu\n", name, live.size());
  expect(ok, name);
}

void misuse() {
  cg_rtx_index_updates<std::uint32_t, 0, 5> index;
  bool threw = false;
  try {
    (void)index.compute_total_size(nullptr);
  } catch (const std::invalid_argument &) {
    threw = true;
  }
  expect(threw, "use before build throws std::invalid_argument");
}

}  // namespace

int main() {
  try {
    interface_known_answer();
    update_rounds<std::uint32_t>("key32");
    update_rounds<std::uint64_t>("key64");
    misuse();
  } catch (const std::exception &e) {
    std::printf("EXCEPTION %s\n", e.what());
    ++failures;
  }
  std::printf("FAILURES %d\n", failures);
  return failures ? 1 : 0;
}
