// CPU baseline: a multithreaded sorted-array LSM with the reference's
// semantics (newest update wins, tombstones mask older runs and are dropped
// when a merge produces the oldest run).  TEST/BENCH INFRASTRUCTURE ONLY:
// it is timed beside the GPU as a baseline and is never part of the product.
//
// Structure: every batch is sorted (std::sort of (key, position) chunks in
// parallel, then a parallel merge tree), deduplicated keeping the newest
// occurrence, and pushed as a run; the two newest runs merge while the older
// is no larger in batch rows (a binary counter for equal batches).  Merges
// split both runs at key boundaries so each thread merges an independent
// slice.  Lookups and range queries run in parallel over queries and
// binary-search each run; range counts and sums over a tombstone-free
// oldest run are rank and prefix-sum differences.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <functional>
#include <thread>
#include <vector>

namespace {

using u8 = std::uint8_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;

struct Run {
  std::vector<u32> keys;
  std::vector<u32> vals;
  std::vector<u8> tomb;
  // rows of the batches merged into the run (an upper bound on its size):
  // the merge policy compares these, as the GPU library's does, so equal
  // batches make an exact binary counter
  u64 ub = 0;
  std::size_t size() const { return keys.size(); }
};

bool any_tombstone(const Run &r) {
  return std::find(r.tomb.begin(), r.tomb.end(), u8(1)) != r.tomb.end();
}

void parallel_for(unsigned threads, std::size_t n,
                  const std::function<void(std::size_t, std::size_t)> &body) {
  if (threads <= 1 || n < 4096) {
    body(0, n);
    return;
  }
  std::vector<std::thread> pool;
  const std::size_t chunk = (n + threads - 1) / threads;
  for (unsigned t = 0; t < threads; ++t) {
    const std::size_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    pool.emplace_back(body, a, b);
  }
  for (auto &th : pool) th.join();
}

struct CpuLsm {
  unsigned threads = 1;
  std::vector<Run> runs;  // newest first
  // Of the oldest run, when it holds no tombstone: prefix sums of its
  // values (prefix[i] = sum of vals[0..i)), so a range over it is a rank
  // difference.  Rebuilt by prepare_ranges() after the oldest run changes.
  std::vector<u64> prefix;
  bool prefix_valid = false;
  bool oldest_plain = false;

  // Sort one batch into a run.
  Run make_run(const u32 *keys, const u32 *vals, const u8 *flags, u64 n,
               int mode) {
    std::vector<u64> items(n);  // key << 32 | position
    parallel_for(threads, n, [&](std::size_t a, std::size_t b) {
      for (std::size_t i = a; i < b; ++i)
        items[i] = (u64(keys[i]) << 32) | u64(i);
    });
    // sort chunks in parallel, then merge pairwise in parallel rounds
    const unsigned parts = std::max(1u, std::min<unsigned>(threads, n / 4096 + 1));
    std::vector<std::size_t> bounds(parts + 1);
    for (unsigned p = 0; p <= parts; ++p) bounds[p] = n * p / parts;
    {
      std::vector<std::thread> pool;
      for (unsigned p = 0; p < parts; ++p)
        pool.emplace_back([&, p] {
          std::sort(items.begin() + bounds[p], items.begin() + bounds[p + 1]);
        });
      for (auto &t : pool) t.join();
    }
    std::vector<u64> tmp(n);
    for (unsigned width = 1; width < parts; width *= 2) {
      std::vector<std::thread> pool;
      for (unsigned p = 0; p < parts; p += 2 * width) {
        const std::size_t a = bounds[p];
        const std::size_t m = bounds[std::min(parts, p + width)];
        const std::size_t b = bounds[std::min(parts, p + 2 * width)];
        pool.emplace_back([&, a, m, b] {
          std::merge(items.begin() + a, items.begin() + m,
                     items.begin() + m, items.begin() + b, tmp.begin() + a);
        });
      }
      for (auto &t : pool) t.join();
      items.swap(tmp);
    }
    Run r;
    r.ub = n;
    r.keys.reserve(n);
    r.vals.reserve(n);
    r.tomb.reserve(n);
    for (u64 i = 0; i < n; ++i) {
      const u32 k = u32(items[i] >> 32);
      if (i + 1 < n && u32(items[i + 1] >> 32) == k) continue;  // newer exists
      const u32 pos = u32(items[i]);
      const bool t = mode == 1 || (mode == 2 && flags[pos]);
      r.keys.push_back(k);
      r.vals.push_back(t ? 0u : vals[pos]);
      r.tomb.push_back(t ? 1 : 0);
    }
    return r;
  }

  // a newer than b; equal keys keep a's row.
  Run merge(const Run &a, const Run &b, bool drop) {
    const unsigned parts = std::max(1u, std::min<unsigned>(
        threads, (a.size() + b.size()) / 65536 + 1));
    std::vector<std::size_t> sa(parts + 1), sb(parts + 1);
    sa[0] = sb[0] = 0;
    sa[parts] = a.size();
    sb[parts] = b.size();
    for (unsigned p = 1; p < parts; ++p) {
      sa[p] = a.size() * p / parts;
      sb[p] = sa[p] < a.size()
                  ? std::size_t(std::lower_bound(b.keys.begin(), b.keys.end(),
                                                 a.keys[sa[p]]) -
                                b.keys.begin())
                  : b.size();
    }
    std::vector<Run> pieces(parts);
    std::vector<std::thread> pool;
    for (unsigned p = 0; p < parts; ++p)
      pool.emplace_back([&, p] {
        Run &o = pieces[p];
        std::size_t i = sa[p], j = sb[p];
        const std::size_t ie = sa[p + 1], je = sb[p + 1];
        o.keys.reserve(ie - i + je - j);
        o.vals.reserve(ie - i + je - j);
        o.tomb.reserve(ie - i + je - j);
        auto emit = [&](const Run &r, std::size_t x) {
          if (drop && r.tomb[x]) return;
          o.keys.push_back(r.keys[x]);
          o.vals.push_back(r.vals[x]);
          o.tomb.push_back(r.tomb[x]);
        };
        while (i < ie || j < je) {
          if (j >= je || (i < ie && a.keys[i] <= b.keys[j])) {
            if (j < je && a.keys[i] == b.keys[j]) ++j;  // shadowed
            emit(a, i++);
          } else {
            emit(b, j++);
          }
        }
      });
    for (auto &t : pool) t.join();
    Run out;
    out.ub = a.ub + b.ub;
    std::size_t total = 0;
    for (auto &p : pieces) total += p.size();
    out.keys.resize(total);
    out.vals.resize(total);
    out.tomb.resize(total);
    std::size_t at = 0;
    for (auto &p : pieces) {
      std::copy(p.keys.begin(), p.keys.end(), out.keys.begin() + at);
      std::copy(p.vals.begin(), p.vals.end(), out.vals.begin() + at);
      std::copy(p.tomb.begin(), p.tomb.end(), out.tomb.begin() + at);
      at += p.size();
    }
    return out;
  }

  void push(Run r) {
    runs.insert(runs.begin(), std::move(r));
    if (runs.size() == 1) prefix_valid = false;
    while (runs.size() >= 2 && runs[1].ub <= runs[0].ub) {
      if (runs.size() == 2) prefix_valid = false;  // the oldest is rewritten
      const bool oldest = runs.size() == 2;
      Run m = merge(runs[0], runs[1], oldest);
      runs.erase(runs.begin(), runs.begin() + 2);
      runs.insert(runs.begin(), std::move(m));
    }
  }

  // newest version of key: returns -1 absent, 0 tombstone, 1 live
  int find(u32 key, u32 &value) const {
    for (const Run &r : runs) {
      auto it = std::lower_bound(r.keys.begin(), r.keys.end(), key);
      if (it != r.keys.end() && *it == key) {
        const std::size_t x = std::size_t(it - r.keys.begin());
        if (r.tomb[x]) return 0;
        value = r.vals[x];
        return 1;
      }
    }
    return -1;
  }

  // Called before a set of range queries (single-threaded).
  void prepare_ranges() {
    if (prefix_valid) return;
    prefix.clear();
    oldest_plain = !runs.empty() && !any_tombstone(runs.back());
    if (oldest_plain) {
      const Run &o = runs.back();
      prefix.resize(o.size() + 1);
      prefix[0] = 0;
      for (std::size_t i = 0; i < o.size(); ++i)
        prefix[i + 1] = prefix[i] + o.vals[i];
    }
    prefix_valid = true;
  }

  // Live keys (count) or the sum of their values in [lo, hi].  With a
  // tombstone-free oldest run the oldest run's part is a rank difference
  // (count) or a prefix-sum difference (sum), corrected by the newest
  // versions that newer runs hold of its keys; only the newer runs' rows in
  // the range are visited.
  u64 range(u32 lo, u32 hi, bool count_only) const {
    if (lo > hi) return 0;
    if (runs.empty()) return 0;
    if (!oldest_plain || !prefix_valid) return range_visit(lo, hi, count_only);
    const std::size_t R = runs.size();
    std::vector<std::pair<std::size_t, std::size_t>> seg(R);
    for (std::size_t r = 0; r < R; ++r) {
      const auto &k = runs[r].keys;
      seg[r].first = std::size_t(std::lower_bound(k.begin(), k.end(), lo) -
                                 k.begin());
      seg[r].second = std::size_t(std::upper_bound(k.begin(), k.end(), hi) -
                                  k.begin());
    }
    const Run &old = runs[R - 1];
    u64 acc = count_only ? u64(seg[R - 1].second - seg[R - 1].first)
                         : prefix[seg[R - 1].second] - prefix[seg[R - 1].first];
    for (std::size_t r = 0; r + 1 < R; ++r) {
      const Run &run = runs[r];
      for (std::size_t x = seg[r].first; x < seg[r].second; ++x) {
        const u32 key = run.keys[x];
        bool shadowed = false;
        for (std::size_t s = 0; s < r && !shadowed; ++s) {
          const auto &k = runs[s].keys;
          shadowed = std::binary_search(k.begin() + seg[s].first,
                                        k.begin() + seg[s].second, key);
        }
        if (shadowed) continue;  // a newer run holds the newest version
        auto it = std::lower_bound(old.keys.begin() + seg[R - 1].first,
                                   old.keys.begin() + seg[R - 1].second, key);
        if (it != old.keys.begin() + seg[R - 1].second && *it == key)
          acc -= count_only ? 1u
                            : old.vals[std::size_t(it - old.keys.begin())];
        if (!run.tomb[x]) acc += count_only ? 1u : run.vals[x];
      }
    }
    return acc;
  }

  // Every row in the range of every run, each checked against the newer
  // runs (any oldest run).
  u64 range_visit(u32 lo, u32 hi, bool count_only) const {
    std::vector<std::pair<std::size_t, std::size_t>> seg(runs.size());
    for (std::size_t r = 0; r < runs.size(); ++r) {
      const auto &k = runs[r].keys;
      seg[r].first = std::size_t(std::lower_bound(k.begin(), k.end(), lo) -
                                 k.begin());
      seg[r].second = std::size_t(std::upper_bound(k.begin(), k.end(), hi) -
                                  k.begin());
    }
    u64 acc = 0;
    for (std::size_t r = 0; r < runs.size(); ++r) {
      const Run &run = runs[r];
      for (std::size_t x = seg[r].first; x < seg[r].second; ++x) {
        if (run.tomb[x]) continue;
        bool shadowed = false;
        for (std::size_t s = 0; s < r && !shadowed; ++s) {
          const auto &k = runs[s].keys;
          shadowed = std::binary_search(k.begin() + seg[s].first,
                                        k.begin() + seg[s].second,
                                        run.keys[x]);
        }
        if (!shadowed) acc += count_only ? 1u : run.vals[x];
      }
    }
    return acc;
  }
};

}  // namespace

extern "C" {

void *cpu_lsm_create(unsigned threads) {
  auto *c = new CpuLsm();
  c->threads = threads ? threads : 1;
  return c;
}

void cpu_lsm_destroy(void *h) { delete static_cast<CpuLsm *>(h); }

void cpu_lsm_clear(void *h) {
  auto *c = static_cast<CpuLsm *>(h);
  c->runs.clear();
  c->prefix_valid = false;
}

// mode: 0 insert, 1 erase, 2 mixed (flags[i] != 0 -> erase)
void cpu_lsm_apply(void *h, const u32 *keys, const u32 *vals, const u8 *flags,
                   u64 n, int mode) {
  auto *c = static_cast<CpuLsm *>(h);
  if (!n) return;
  c->push(c->make_run(keys, vals, flags, n, mode));
}

void cpu_lsm_lookup(void *h, const u32 *q, u64 n, u32 *out, u8 *found) {
  auto *c = static_cast<CpuLsm *>(h);
  parallel_for(c->threads, n, [&](std::size_t a, std::size_t b) {
    for (std::size_t i = a; i < b; ++i) {
      u32 v = 0;
      const bool live = c->find(q[i], v) == 1;
      if (found) {
        out[i] = live ? v : 0u;
        found[i] = live;
      } else {
        out[i] = live ? v : 0xffffffffu;
      }
    }
  });
}

void cpu_lsm_range(void *h, const u32 *lo, const u32 *hi, u64 n, u32 *out,
                   int count_only) {
  auto *c = static_cast<CpuLsm *>(h);
  c->prepare_ranges();
  // dynamic chunks: range widths vary by orders of magnitude
  std::atomic<u64> next{0};
  const u64 grain = 256;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < c->threads; ++t)
    pool.emplace_back([&] {
      while (true) {
        const u64 a = next.fetch_add(grain);
        if (a >= n) break;
        const u64 b = std::min(n, a + grain);
        for (u64 i = a; i < b; ++i)
          out[i] = u32(c->range(lo[i], hi[i], count_only != 0));
      }
    });
  for (auto &t : pool) t.join();
}

u64 cpu_lsm_rows(void *h) {
  u64 t = 0;
  for (const auto &r : static_cast<CpuLsm *>(h)->runs) t += r.size();
  return t;
}

unsigned cpu_lsm_runs(void *h) {
  return unsigned(static_cast<CpuLsm *>(h)->runs.size());
}

}  // extern "C"
