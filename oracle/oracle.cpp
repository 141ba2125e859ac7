// Oracle: CPU restatement of the reference's observable dictionary semantics.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this file.  Only the tests, the smoke check and the CPU-baseline leg
// of the benchmark use it, and only as the checker / the CPU baseline.
//
// Contents
//   1. Counter-based generators (splitmix64, MurmurHash3_x86_32, log-uniform,
//      Zipf by rejection-inversion, the paper-sweep key bijection).
//   2. LSM oracle: a sorted-vector dictionary that replays GPULSMOpt's
//      observable contract (GPULSMOpt.cuh):
//        * newest update wins: batch order, then position inside the batch
//          (GPULSMOpt.cuh:398-404 raw_age, :2594-2596 metadata = slot<<20|i,
//          :4755-4770 newest age wins inside a pending batch);
//        * erase writes a tombstone that masks every older version
//          (:2594-2596 kRawTombstone, :4727-4735 tombstone -> miss);
//        * lookup miss = 0xFFFFFFFF without out_found, else value 0 and
//          found 0 (:4729-4744, kInvalid :53);
//        * range = inclusive [lo,hi] sum of live values, accumulated in
//          uint64 and truncated to uint32, lo > hi -> 0 (:1231-1238,
//          :1424-1427, :1566-1568);
//        * successor = smallest live key >= query, none -> 0xFFFFFFFF
//          (:5566-5594);
//        * bulk_build = stable sort, last occurrence of a key wins
//          (:5755-5774, mark_last_key :2633-2638);
//        * update tiling into batch_capacity tiles <= 2^20, 16 batches per
//          epoch, capacity = max(max_elements, 16*batch) (:5606-5616,
//          :255-270, :6895-6973);
//        * a publication whose survivors exceed capacity fails with status
//          kPublicationOutputOverflow (8); reads keep seeing every accepted
//          update and every later update throws (:6750-6849, :7350-7363).
//          Restated as "live keys > capacity at an epoch boundary", which is
//          the reference's survivor rule when the destination holds the
//          whole dictionary and no tombstones are retained (its capacity
//          scenario, FliX/paper_lsmu_sweep.cu:643-695).
//   3. FliX oracle: sorted-vector set with insert-if-absent semantics
//      (FliX/coarse_granular_inserts_tiles_bulk.cuh:1423-1429 skips keys that
//      are already stored; deletes remove present keys and ignore absent
//      ones, FliX/coarse_granular_deletes_tiles.cuh:1159-1299; lookups return
//      the stored offset or not_found, FliX/coarse_granular_lookups_tile_bulk.cuh:
//      312-481; successor returns the ceiling key or key_type(not_found),
//      :630-969; build values are input positions,
//      FliX/impl_cg_rtx_index_updates.cuh:665-814 init_offsets).
//
// Everything is exported with C linkage for ctypes.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <utility>
#include <vector>

namespace {

// ---------------------------------------------------------------- generators
inline std::uint64_t splitmix64(std::uint64_t x) {
  // SplitMix64 (Steele, Lea, Flood 2014), the same finalizer constants as
  // digest_mix in FliX/paper_lsmu_sweep.cu:188-194.
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

inline std::uint32_t rotl32(std::uint32_t x, int r) {
  return (x << r) | (x >> (32 - r));
}

// MurmurHash3_x86_32 (Austin Appleby, public domain) over the four
// little-endian bytes of `key`.
inline std::uint32_t murmur3_u32(std::uint32_t key, std::uint32_t seed) {
  const std::uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
  std::uint32_t h = seed;
  std::uint32_t k = key;
  k *= c1;
  k = rotl32(k, 15);
  k *= c2;
  h ^= k;
  h = rotl32(h, 13);
  h = h * 5u + 0xe6546b64u;
  h ^= 4u;  // length in bytes
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

// Generic MurmurHash3_x86_32 over a byte string (used by the known-answer
// test against the algorithm's published vectors).
std::uint32_t murmur3_bytes(const std::uint8_t *data, std::size_t len,
                            std::uint32_t seed) {
  const std::uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
  std::uint32_t h = seed;
  const std::size_t blocks = len / 4;
  for (std::size_t i = 0; i < blocks; ++i) {
    std::uint32_t k = std::uint32_t(data[4 * i]) |
                      (std::uint32_t(data[4 * i + 1]) << 8) |
                      (std::uint32_t(data[4 * i + 2]) << 16) |
                      (std::uint32_t(data[4 * i + 3]) << 24);
    k *= c1;
    k = rotl32(k, 15);
    k *= c2;
    h ^= k;
    h = rotl32(h, 13);
    h = h * 5u + 0xe6546b64u;
  }
  const std::uint8_t *tail = data + blocks * 4;
  std::uint32_t k1 = 0;
  switch (len & 3u) {
    case 3: k1 ^= std::uint32_t(tail[2]) << 16; [[fallthrough]];
    case 2: k1 ^= std::uint32_t(tail[1]) << 8; [[fallthrough]];
    case 1:
      k1 ^= tail[0];
      k1 *= c1;
      k1 = rotl32(k1, 15);
      k1 *= c2;
      h ^= k1;
  }
  h ^= std::uint32_t(len);
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

// floor(2^(log2 a + u*(log2 b - log2 a))) with u = (x >> 11) * 2^-53,
// clamped to [a, b].
inline std::uint64_t log_uniform(std::uint64_t x, std::uint64_t a,
                                 std::uint64_t b) {
  const double u = double(x >> 11) * 0x1.0p-53;
  const double la = std::log2(double(a)), lb = std::log2(double(b));
  double v = std::floor(std::exp2(la + u * (lb - la)));
  std::uint64_t r = v < double(a) ? a : std::uint64_t(v);
  return r > b ? b : r;
}

// Zipf(s) over ranks 1..n by rejection-inversion (W. Hormann and
// G. Derflinger, "Rejection-inversion to generate variates from monotone
// discrete distributions", ACM TOMACS 6(3), 1996).
struct Zipf {
  double s, n, h_x1, h_n, sv;
  static double helper1(double x) {  // log1p(x)/x
    return std::fabs(x) > 1e-8 ? std::log1p(x) / x : 1.0 - x * (0.5 - x / 3.0);
  }
  static double helper2(double x) {  // expm1(x)/x
    return std::fabs(x) > 1e-8 ? std::expm1(x) / x
                               : 1.0 + x * 0.5 * (1.0 + x / 3.0);
  }
  double h(double x) const { return std::exp(-s * std::log(x)); }
  double H(double x) const {
    const double lx = std::log(x);
    return helper2((1.0 - s) * lx) * lx;
  }
  double H_inv(double x) const {
    double t = x * (1.0 - s);
    if (t < -1.0) t = -1.0;
    return std::exp(helper1(t) * x);
  }
  Zipf(double exponent, std::uint64_t count) : s(exponent), n(double(count)) {
    h_x1 = H(1.5) - 1.0;
    h_n = H(n + 0.5);
    sv = 2.0 - H_inv(H(2.5) - h(2.0));
  }
  // Draws with the counter-based stream x_j = splitmix64(base + j).
  std::uint64_t sample(std::uint64_t base) const {
    for (std::uint64_t j = 0;; ++j) {
      const double u01 = double(splitmix64(base + j) >> 11) * 0x1.0p-53;
      const double u = h_n + u01 * (h_x1 - h_n);
      const double x = H_inv(u);
      double k = std::floor(x + 0.5);
      if (k < 1.0) k = 1.0;
      if (k > n) k = n;
      if (k - x <= sv || u >= H(k + 0.5) - h(k)) return std::uint64_t(k);
    }
  }
};

// ----------------------------------------------------------------- LSM oracle
constexpr std::uint32_t kBatchesPerEpoch = 16u;
constexpr std::uint64_t kMaximumOperationTile = 1ull << 20;
constexpr std::uint32_t kStatusOutputOverflow = 1u << 3;

enum OracleCode : int {
  ORC_OK = 0,
  ORC_INVALID_ARGUMENT = 1,
  ORC_PUBLICATION_FAILED = 4,
};

// K is the key and value type (uint32 for GPULSMOpt; uint64 for the 64-bit
// extension, where sums wrap modulo 2^64 and the sentinel is all-ones).
template <class K>
struct Update {
  K key;
  K value;
  bool tombstone;
};

template <class K>
struct LsmOracleT {
  using Entry = std::pair<K, K>;
  std::uint64_t batch_capacity = 1;
  std::uint64_t capacity = 0;
  // Live contents, sorted by key.
  std::vector<Entry> live;
  std::uint32_t pending_batches = 0;
  bool failed = false;
  std::uint32_t status = 0;
  std::string message;

  // Apply one admitted batch: later positions override earlier ones.
  void apply_tile(const Update<K> *u, std::size_t n) {
    std::vector<std::size_t> order(n);
    std::iota(order.begin(), order.end(), std::size_t{0});
    std::stable_sort(order.begin(), order.end(),
                     [&](std::size_t a, std::size_t b) {
                       return u[a].key < u[b].key;
                     });
    // keep the last occurrence of every key (the newest age)
    std::vector<Update<K>> winners;
    winners.reserve(n);
    for (std::size_t i = 0; i < n; ++i) {
      if (i + 1 < n && u[order[i]].key == u[order[i + 1]].key) continue;
      winners.push_back(u[order[i]]);
    }
    std::vector<Entry> merged;
    merged.reserve(live.size() + winners.size());
    std::size_t a = 0, b = 0;
    while (a < live.size() || b < winners.size()) {
      if (b == winners.size() ||
          (a < live.size() && live[a].first < winners[b].key)) {
        merged.push_back(live[a++]);
      } else {
        if (a < live.size() && live[a].first == winners[b].key) ++a;
        if (!winners[b].tombstone)
          merged.emplace_back(winners[b].key, winners[b].value);
        ++b;
      }
    }
    live.swap(merged);
  }

  int admit(const Update<K> *u, std::size_t n) {
    if (failed) {
      message =
          "GPULSMOpt publication failed with status " +
          std::to_string(status) + "; capacity=" + std::to_string(capacity) +
          "; pending updates were preserved";
      return ORC_PUBLICATION_FAILED;
    }
    std::size_t consumed = 0;
    while (consumed < n) {
      const std::size_t tile =
          std::min<std::size_t>(n - consumed, batch_capacity);
      apply_tile(u + consumed, tile);
      consumed += tile;
      if (++pending_batches == kBatchesPerEpoch) {
        pending_batches = 0;
        if (live.size() > capacity) {
          failed = true;
          status = kStatusOutputOverflow;
        }
      }
      if (consumed < n && failed) {
        message =
            "GPULSMOpt publication failed while tiling an update (capacity=" +
            std::to_string(capacity) + "); accepted records were preserved";
        return ORC_PUBLICATION_FAILED;
      }
    }
    return ORC_OK;
  }

  const Entry *find(K k) const {
    auto it = std::lower_bound(
        live.begin(), live.end(), k,
        [](const Entry &e, K x) { return e.first < x; });
    if (it == live.end() || it->first != k) return nullptr;
    return &*it;
  }
};

using LsmOracle = LsmOracleT<std::uint32_t>;
using LsmOracle64 = LsmOracleT<std::uint64_t>;

// --------------------------------------------------- LSM oracle operations
template <class K>
LsmOracleT<K> *lsm_create(std::uint64_t max_elements,
                          std::uint64_t batch_capacity) {
  auto *o = new LsmOracleT<K>();
  o->batch_capacity = std::min<std::uint64_t>(
      kMaximumOperationTile, std::max<std::uint64_t>(1u, batch_capacity));
  o->capacity = std::max<std::uint64_t>(max_elements,
                                        o->batch_capacity * kBatchesPerEpoch);
  if (o->capacity > 0xffffffffull) {
    delete o;
    return nullptr;
  }
  return o;
}

template <class K>
void lsm_clear(LsmOracleT<K> *o) {
  o->live.clear();
  o->pending_batches = 0;
  o->failed = false;
  o->status = 0;
}

template <class K>
int lsm_bulk_build(LsmOracleT<K> *o, const K *keys, const K *values,
                   std::uint64_t n) {
  if (n > 0xffffffffull || (n && (!keys || !values)))
    return ORC_INVALID_ARGUMENT;
  lsm_clear(o);
  std::vector<Update<K>> u(n);
  for (std::uint64_t i = 0; i < n; ++i) u[i] = {keys[i], values[i], false};
  o->apply_tile(u.data(), n);  // one stable pass, last occurrence wins
  return ORC_OK;
}

// Mixed batch: entry i is an erase when is_erase[i] != 0 (all erases when
// values is null and is_erase is null); age = position.
template <class K>
int lsm_admit(LsmOracleT<K> *o, const K *keys, const K *values,
              const std::uint8_t *is_erase, bool all_erase, std::uint64_t n) {
  if (!n) return ORC_OK;
  if (!keys || (!all_erase && !values)) return ORC_INVALID_ARGUMENT;
  std::vector<Update<K>> u(n);
  for (std::uint64_t i = 0; i < n; ++i) {
    const bool e = all_erase || (is_erase && is_erase[i] != 0);
    u[i] = {keys[i], e ? K(0) : values[i], e};
  }
  return o->admit(u.data(), n);
}

template <class K>
void lsm_lookup(LsmOracleT<K> *o, const K *queries, std::uint64_t n,
                K *out_values, std::uint8_t *out_found) {
  for (std::uint64_t i = 0; i < n; ++i) {
    const auto *e = o->find(queries[i]);
    if (e) {
      out_values[i] = e->second;
      if (out_found) out_found[i] = 1;
    } else {
      out_values[i] = out_found ? K(0) : static_cast<K>(~K(0));
      if (out_found) out_found[i] = 0;
    }
  }
}

// Sums accumulate in uint64 (wrapping for 64-bit values) and are truncated
// to K.
template <class K>
void lsm_range(LsmOracleT<K> *o, const K *lo, const K *hi, std::uint64_t n,
               K *out, bool count_only) {
  // prefix sums over the live snapshot give O(log n) queries
  std::vector<std::uint64_t> prefix(o->live.size() + 1, 0);
  for (std::size_t i = 0; i < o->live.size(); ++i)
    prefix[i + 1] = prefix[i] + (count_only ? 1u : o->live[i].second);
  using Entry = typename LsmOracleT<K>::Entry;
  for (std::uint64_t i = 0; i < n; ++i) {
    if (lo[i] > hi[i]) {
      out[i] = 0;
      continue;
    }
    const std::size_t a = std::size_t(
        std::lower_bound(o->live.begin(), o->live.end(), lo[i],
                         [](const Entry &e, K x) { return e.first < x; }) -
        o->live.begin());
    const std::size_t b = std::size_t(
        std::upper_bound(o->live.begin(), o->live.end(), hi[i],
                         [](K x, const Entry &e) { return x < e.first; }) -
        o->live.begin());
    out[i] = static_cast<K>(prefix[b] - prefix[a]);
  }
}

template <class K>
void lsm_successor(LsmOracleT<K> *o, const K *queries, std::uint64_t n,
                   K *out) {
  using Entry = typename LsmOracleT<K>::Entry;
  for (std::uint64_t i = 0; i < n; ++i) {
    auto it = std::lower_bound(
        o->live.begin(), o->live.end(), queries[i],
        [](const Entry &e, K x) { return e.first < x; });
    out[i] = it == o->live.end() ? static_cast<K>(~K(0)) : it->first;
  }
}

template <class K>
void lsm_dump(LsmOracleT<K> *o, K *keys, K *values) {
  for (std::size_t i = 0; i < o->live.size(); ++i) {
    keys[i] = o->live[i].first;
    values[i] = o->live[i].second;
  }
}

// ---------------------------------------------------------------- FliX oracle
template <class K, class V>
struct FlixOracle {
  std::vector<std::pair<K, V>> data;  // sorted by key, unique keys
  static K key_not_found() { return static_cast<K>(0xffffffffu); }
  static V value_not_found() { return static_cast<V>(~V{0}); }

  typename std::vector<std::pair<K, V>>::const_iterator lb(K k) const {
    return std::lower_bound(
        data.begin(), data.end(), k,
        [](const std::pair<K, V> &e, K x) { return e.first < x; });
  }

  // Sort by key (stable), keep the first occurrence of each key.
  static std::vector<std::pair<K, V>> sorted_unique_first(
      const K *keys, const V *values, std::size_t n) {
    std::vector<std::pair<K, V>> v(n);
    for (std::size_t i = 0; i < n; ++i) v[i] = {keys[i], values[i]};
    std::stable_sort(v.begin(), v.end(),
                     [](const std::pair<K, V> &a, const std::pair<K, V> &b) {
                       return a.first < b.first;
                     });
    std::vector<std::pair<K, V>> out;
    out.reserve(n);
    for (std::size_t i = 0; i < n; ++i)
      if (out.empty() || out.back().first != v[i].first) out.push_back(v[i]);
    return out;
  }

  void build(const K *keys, std::size_t n) {
    std::vector<V> idx(n);
    for (std::size_t i = 0; i < n; ++i) idx[i] = static_cast<V>(i);
    data = sorted_unique_first(keys, idx.data(), n);
  }

  void insert(const K *keys, const V *values, std::size_t n) {
    auto batch = sorted_unique_first(keys, values, n);
    std::vector<std::pair<K, V>> merged;
    merged.reserve(data.size() + batch.size());
    std::size_t a = 0, b = 0;
    while (a < data.size() || b < batch.size()) {
      if (b == batch.size() ||
          (a < data.size() && data[a].first <= batch[b].first)) {
        if (b < batch.size() && data[a].first == batch[b].first) ++b;
        merged.push_back(data[a++]);
      } else {
        merged.push_back(batch[b++]);
      }
    }
    data.swap(merged);
  }

  void remove(const K *keys, std::size_t n) {
    std::vector<K> del(keys, keys + n);
    std::sort(del.begin(), del.end());
    std::vector<std::pair<K, V>> kept;
    kept.reserve(data.size());
    std::size_t d = 0;
    for (const auto &e : data) {
      while (d < del.size() && del[d] < e.first) ++d;
      if (d < del.size() && del[d] == e.first) continue;
      kept.push_back(e);
    }
    data.swap(kept);
  }
};

}  // namespace

extern "C" {

// ------------------------------------------------------------ generator ABI
std::uint64_t orc_splitmix64(std::uint64_t x) { return splitmix64(x); }

std::uint32_t orc_murmur3_u32(std::uint32_t key, std::uint32_t seed) {
  return murmur3_u32(key, seed);
}

std::uint32_t orc_murmur3_bytes(const std::uint8_t *data, std::uint64_t len,
                                std::uint32_t seed) {
  return murmur3_bytes(data, len, seed);
}

std::uint64_t orc_log_uniform(std::uint64_t x, std::uint64_t a,
                              std::uint64_t b) {
  return log_uniform(x, a, b);
}

void orc_zipf_fill(double s, std::uint64_t n, std::uint64_t seed,
                   std::uint64_t count, std::uint64_t *out_ranks) {
  const Zipf z(s, n);
  for (std::uint64_t i = 0; i < count; ++i)
    out_ranks[i] = z.sample(splitmix64(seed + i));
}

void orc_fill_splitmix64(std::uint64_t seed, std::uint64_t n,
                         std::uint64_t *out) {
  for (std::uint64_t i = 0; i < n; ++i) out[i] = splitmix64(seed + i);
}

void orc_fill_murmur3(const std::uint32_t *keys, std::uint64_t n,
                      std::uint32_t seed, std::uint32_t *out) {
  for (std::uint64_t i = 0; i < n; ++i) out[i] = murmur3_u32(keys[i], seed);
}

void orc_fill_log_uniform(std::uint64_t seed, std::uint64_t n, std::uint64_t a,
                          std::uint64_t b, std::uint64_t *out) {
  for (std::uint64_t i = 0; i < n; ++i)
    out[i] = log_uniform(splitmix64(seed + i), a, b);
}

// FliX/paper_lsmu_sweep.cu:64-67 (key_mask = 2^30 - 1 for the 30-bit sweep).
std::uint32_t orc_paper_key_for_index(std::uint32_t index) {
  return (index * 747796405u + 289133645u) & ((1u << 30) - 1u);
}

// FliX/paper_lsmu_sweep.cu:69-77.
std::uint32_t orc_paper_mix32(std::uint32_t value) {
  value ^= value >> 16;
  value *= 0x7feb352du;
  value ^= value >> 15;
  value *= 0x846ca68bu;
  value ^= value >> 16;
  return value;
}

// Range checksum of FliX/paper_lsmu_sweep.cu:196-227:
// (sum of results, xor of digest_mix(query << 32 | result)).
void orc_range_digest(const std::uint32_t *results, std::uint64_t n,
                      std::uint64_t query_offset, std::uint64_t *out_sum,
                      std::uint64_t *out_xor) {
  std::uint64_t s = 0, x = 0;
  for (std::uint64_t i = 0; i < n; ++i) {
    const std::uint64_t q = query_offset + i;
    s += results[i];
    x ^= splitmix64((q << 32) | results[i]);
  }
  *out_sum = s;
  *out_xor = x;
}

// ------------------------------------------------------------------ LSM ABI
void *orc_lsm_create(std::uint64_t max_elements, std::uint64_t batch_capacity,
                     std::uint64_t level_zero_capacity) {
  (void)level_zero_capacity;  // layout-only in the reference
  return lsm_create<std::uint32_t>(max_elements, batch_capacity);
}

void orc_lsm_destroy(void *h) { delete static_cast<LsmOracle *>(h); }

void orc_lsm_clear(void *h) { lsm_clear(static_cast<LsmOracle *>(h)); }

int orc_lsm_bulk_build(void *h, const std::uint32_t *keys,
                       const std::uint32_t *values, std::uint64_t n) {
  return lsm_bulk_build(static_cast<LsmOracle *>(h), keys, values, n);
}

int orc_lsm_insert(void *h, const std::uint32_t *keys,
                   const std::uint32_t *values, std::uint64_t n) {
  return lsm_admit(static_cast<LsmOracle *>(h), keys, values, nullptr, false,
                   n);
}

int orc_lsm_erase(void *h, const std::uint32_t *keys, std::uint64_t n) {
  return lsm_admit<std::uint32_t>(static_cast<LsmOracle *>(h), keys, nullptr,
                                  nullptr, true, n);
}

int orc_lsm_apply_mixed(void *h, const std::uint32_t *keys,
                        const std::uint32_t *values,
                        const std::uint8_t *is_erase, std::uint64_t n) {
  if (n && !is_erase) return ORC_INVALID_ARGUMENT;
  return lsm_admit(static_cast<LsmOracle *>(h), keys, values, is_erase, false,
                   n);
}

void orc_lsm_lookup(void *h, const std::uint32_t *queries, std::uint64_t n,
                    std::uint32_t *out_values, std::uint8_t *out_found) {
  lsm_lookup(static_cast<LsmOracle *>(h), queries, n, out_values, out_found);
}

void orc_lsm_range_sum(void *h, const std::uint32_t *lo,
                       const std::uint32_t *hi, std::uint64_t n,
                       std::uint32_t *out) {
  lsm_range(static_cast<LsmOracle *>(h), lo, hi, n, out, false);
}

void orc_lsm_range_count(void *h, const std::uint32_t *lo,
                         const std::uint32_t *hi, std::uint64_t n,
                         std::uint32_t *out) {
  lsm_range(static_cast<LsmOracle *>(h), lo, hi, n, out, true);
}

void orc_lsm_successor(void *h, const std::uint32_t *queries, std::uint64_t n,
                       std::uint32_t *out) {
  lsm_successor(static_cast<LsmOracle *>(h), queries, n, out);
}

std::uint64_t orc_lsm_size(void *h) {
  return static_cast<LsmOracle *>(h)->live.size();
}

void orc_lsm_dump(void *h, std::uint32_t *keys, std::uint32_t *values) {
  lsm_dump(static_cast<LsmOracle *>(h), keys, values);
}

std::uint32_t orc_lsm_status(void *h) {
  return static_cast<LsmOracle *>(h)->status;
}

std::uint64_t orc_lsm_capacity(void *h) {
  return static_cast<LsmOracle *>(h)->capacity;
}

const char *orc_lsm_message(void *h) {
  return static_cast<LsmOracle *>(h)->message.c_str();
}

// ----------------------------------------------- LSM ABI, 64-bit extension
void *orc_lsm64_create(std::uint64_t max_elements,
                       std::uint64_t batch_capacity) {
  return lsm_create<std::uint64_t>(max_elements, batch_capacity);
}

void orc_lsm64_destroy(void *h) { delete static_cast<LsmOracle64 *>(h); }

void orc_lsm64_clear(void *h) { lsm_clear(static_cast<LsmOracle64 *>(h)); }

int orc_lsm64_bulk_build(void *h, const std::uint64_t *keys,
                         const std::uint64_t *values, std::uint64_t n) {
  return lsm_bulk_build(static_cast<LsmOracle64 *>(h), keys, values, n);
}

int orc_lsm64_insert(void *h, const std::uint64_t *keys,
                     const std::uint64_t *values, std::uint64_t n) {
  return lsm_admit(static_cast<LsmOracle64 *>(h), keys, values, nullptr,
                   false, n);
}

int orc_lsm64_erase(void *h, const std::uint64_t *keys, std::uint64_t n) {
  return lsm_admit<std::uint64_t>(static_cast<LsmOracle64 *>(h), keys,
                                  nullptr, nullptr, true, n);
}

int orc_lsm64_apply_mixed(void *h, const std::uint64_t *keys,
                          const std::uint64_t *values,
                          const std::uint8_t *is_erase, std::uint64_t n) {
  if (n && !is_erase) return ORC_INVALID_ARGUMENT;
  return lsm_admit(static_cast<LsmOracle64 *>(h), keys, values, is_erase,
                   false, n);
}

void orc_lsm64_lookup(void *h, const std::uint64_t *queries, std::uint64_t n,
                      std::uint64_t *out_values, std::uint8_t *out_found) {
  lsm_lookup(static_cast<LsmOracle64 *>(h), queries, n, out_values,
             out_found);
}

void orc_lsm64_range_sum(void *h, const std::uint64_t *lo,
                         const std::uint64_t *hi, std::uint64_t n,
                         std::uint64_t *out) {
  lsm_range(static_cast<LsmOracle64 *>(h), lo, hi, n, out, false);
}

void orc_lsm64_range_count(void *h, const std::uint64_t *lo,
                           const std::uint64_t *hi, std::uint64_t n,
                           std::uint64_t *out) {
  lsm_range(static_cast<LsmOracle64 *>(h), lo, hi, n, out, true);
}

void orc_lsm64_successor(void *h, const std::uint64_t *queries,
                         std::uint64_t n, std::uint64_t *out) {
  lsm_successor(static_cast<LsmOracle64 *>(h), queries, n, out);
}

std::uint64_t orc_lsm64_size(void *h) {
  return static_cast<LsmOracle64 *>(h)->live.size();
}

void orc_lsm64_dump(void *h, std::uint64_t *keys, std::uint64_t *values) {
  lsm_dump(static_cast<LsmOracle64 *>(h), keys, values);
}

std::uint32_t orc_lsm64_status(void *h) {
  return static_cast<LsmOracle64 *>(h)->status;
}

// ----------------------------------------------------------------- FliX ABI
// key_bytes and value_bytes are 4 or 8.
struct FlixAny {
  int kb, vb;
  void *impl;
};

#define FLIX_DISPATCH(h, ...)                                              \
  do {                                                                      \
    FlixAny *f__ = static_cast<FlixAny *>(h);                               \
    if (f__->kb == 4 && f__->vb == 4) {                                     \
      using K = std::uint32_t; using V = std::uint32_t;                     \
      auto *o = static_cast<FlixOracle<K, V> *>(f__->impl); __VA_ARGS__;           \
    } else if (f__->kb == 4) {                                              \
      using K = std::uint32_t; using V = std::uint64_t;                     \
      auto *o = static_cast<FlixOracle<K, V> *>(f__->impl); __VA_ARGS__;           \
    } else if (f__->vb == 4) {                                              \
      using K = std::uint64_t; using V = std::uint32_t;                     \
      auto *o = static_cast<FlixOracle<K, V> *>(f__->impl); __VA_ARGS__;           \
    } else {                                                                \
      using K = std::uint64_t; using V = std::uint64_t;                     \
      auto *o = static_cast<FlixOracle<K, V> *>(f__->impl); __VA_ARGS__;           \
    }                                                                       \
  } while (0)

void *orc_flix_create(int key_bytes, int value_bytes) {
  if ((key_bytes != 4 && key_bytes != 8) ||
      (value_bytes != 4 && value_bytes != 8))
    return nullptr;
  auto *f = new FlixAny{key_bytes, value_bytes, nullptr};
  if (key_bytes == 4 && value_bytes == 4)
    f->impl = new FlixOracle<std::uint32_t, std::uint32_t>();
  else if (key_bytes == 4)
    f->impl = new FlixOracle<std::uint32_t, std::uint64_t>();
  else if (value_bytes == 4)
    f->impl = new FlixOracle<std::uint64_t, std::uint32_t>();
  else
    f->impl = new FlixOracle<std::uint64_t, std::uint64_t>();
  return f;
}

void orc_flix_destroy(void *h) {
  if (!h) return;
  FLIX_DISPATCH(h, delete o);
  delete static_cast<FlixAny *>(h);
}

void orc_flix_build(void *h, const void *keys, std::uint64_t n) {
  FLIX_DISPATCH(h, o->build(static_cast<const K *>(keys), n));
}

void orc_flix_insert(void *h, const void *keys, const void *values,
                     std::uint64_t n) {
  FLIX_DISPATCH(h, o->insert(static_cast<const K *>(keys),
                             static_cast<const V *>(values), n));
}

void orc_flix_remove(void *h, const void *keys, std::uint64_t n) {
  FLIX_DISPATCH(h, o->remove(static_cast<const K *>(keys), n));
}

void orc_flix_lookup(void *h, const void *queries, std::uint64_t n,
                     void *out_values) {
  FLIX_DISPATCH(h, {
    const K *q = static_cast<const K *>(queries);
    V *out = static_cast<V *>(out_values);
    for (std::uint64_t i = 0; i < n; ++i) {
      auto it = o->lb(q[i]);
      out[i] = (it != o->data.end() && it->first == q[i])
                   ? it->second
                   : FlixOracle<K, V>::value_not_found();
    }
  });
}

void orc_flix_successor(void *h, const void *queries, std::uint64_t n,
                        void *out_keys) {
  FLIX_DISPATCH(h, {
    const K *q = static_cast<const K *>(queries);
    K *out = static_cast<K *>(out_keys);
    for (std::uint64_t i = 0; i < n; ++i) {
      auto it = o->lb(q[i]);
      out[i] = it != o->data.end() ? it->first
                                   : FlixOracle<K, V>::key_not_found();
    }
  });
}

// Counts per query: number of stored keys in [lo, hi] (0 when lo > hi).
void orc_flix_range_counts(void *h, const void *lo, const void *hi,
                           std::uint64_t n, std::uint64_t *out_counts) {
  FLIX_DISPATCH(h, {
    const K *l = static_cast<const K *>(lo);
    const K *r = static_cast<const K *>(hi);
    for (std::uint64_t i = 0; i < n; ++i) {
      if (l[i] > r[i]) {
        out_counts[i] = 0;
        continue;
      }
      auto a = o->lb(l[i]);
      auto b = std::upper_bound(
          o->data.begin(), o->data.end(), r[i],
          [](K x, const std::pair<K, V> &e) { return x < e.first; });
      out_counts[i] = std::uint64_t(b - a);
    }
  });
}

// Enumerates the pairs of query i at out_keys/out_values[offsets[i] ...].
void orc_flix_range_enumerate(void *h, const void *lo, const void *hi,
                              std::uint64_t n, const std::uint64_t *offsets,
                              void *out_keys, void *out_values) {
  FLIX_DISPATCH(h, {
    const K *l = static_cast<const K *>(lo);
    const K *r = static_cast<const K *>(hi);
    K *ok = static_cast<K *>(out_keys);
    V *ov = static_cast<V *>(out_values);
    for (std::uint64_t i = 0; i < n; ++i) {
      if (l[i] > r[i]) continue;
      std::uint64_t w = offsets[i];
      for (auto it = o->lb(l[i]); it != o->data.end() && it->first <= r[i];
           ++it, ++w) {
        ok[w] = it->first;
        ov[w] = it->second;
      }
    }
  });
}

std::uint64_t orc_flix_size(void *h) {
  std::uint64_t s = 0;
  FLIX_DISPATCH(h, s = o->data.size());
  return s;
}

void orc_flix_dump(void *h, void *keys, void *values) {
  FLIX_DISPATCH(h, {
    K *ok = static_cast<K *>(keys);
    V *ov = static_cast<V *>(values);
    for (std::size_t i = 0; i < o->data.size(); ++i) {
      ok[i] = o->data[i].first;
      ov[i] = o->data[i].second;
    }
  });
}

}  // extern "C"
