"""GPU parity of the 64-bit LSM dictionary (gd_lsm64_*, uint64 keys and
values) against the oracle's 64-bit replay, bit for bit: sorted contents,
lookups with and without found flags, range sums (mod 2^64) and counts,
successors.  Covers both halves of the key space (the top key bit is
carried by the run position), edge keys, the fused small-batch admission,
the radix-sort admission, merges with tombstones, bulk builds and the
serving engine with 64-bit keys."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALL = np.uint64(0xFFFFFFFFFFFFFFFF)
EDGE = np.array([0, 1, 0x7FFFFFFFFFFFFFFF, 0x8000000000000000,
                 0xFFFFFFFFFFFFFFFE, 0xFFFFFFFFFFFFFFFF, 0xFFFFFFFF,
                 0x100000000], np.uint64)


class Pair64:
    """64-bit GPU dictionary + oracle driven in lock step."""

    def __init__(self, gd, orc, io, tuning=None, **cfg):
        self.d = gd.GPULSMOpt(tuning=tuning, key_bytes=8, **cfg)
        self.o = orc.LsmOracle(key_bytes=8, **cfg)
        self.io = io

    def up(self, a):
        return self.io.up(np.asarray(a, np.uint64), np.uint64)

    def down(self, t, n):
        return self.io.down(t, n, np.uint64)

    def insert(self, keys, vals):
        self.d.insert(self.up(keys), self.up(vals), count=len(keys))
        assert self.o.insert(keys, vals) == 0

    def erase(self, keys):
        self.d.erase(self.up(keys), count=len(keys))
        assert self.o.erase(keys) == 0

    def mixed(self, keys, vals, erase):
        self.d.apply_mixed(self.up(keys), self.up(vals),
                           self.io.up(erase, np.uint8), count=len(keys))
        assert self.o.apply_mixed(keys, vals, erase) == 0

    def check_lookup(self, q, with_found):
        io = self.io
        out = io.empty(len(q), 8)
        found = io.empty(len(q), 1) if with_found else None
        self.d.lookup(self.up(q), out, found, count=len(q))
        if with_found:
            ev, ef = self.o.lookup(q, with_found=True)
            assert np.array_equal(self.down(out, len(q)), ev)
            assert np.array_equal(io.down(found, len(q), np.uint8), ef)
        else:
            assert np.array_equal(self.down(out, len(q)), self.o.lookup(q))

    def check_ranges(self, lo, hi):
        io = self.io
        s = io.empty(len(lo), 8)
        c = io.empty(len(lo), 8)
        self.d.range(self.up(lo), self.up(hi), s, count=len(lo))
        self.d.range_count(self.up(lo), self.up(hi), c, count=len(lo))
        assert np.array_equal(self.down(s, len(lo)), self.o.range_sum(lo, hi))
        assert np.array_equal(self.down(c, len(lo)),
                              self.o.range_count(lo, hi))

    def check_successor(self, q):
        out = self.io.empty(len(q), 8)
        self.d.successor(self.up(q), out, count=len(q))
        assert np.array_equal(self.down(out, len(q)), self.o.successor(q))

    def check_contents(self):
        io = self.io
        n = self.d.dump_sorted()
        ek, ev = self.o.dump()
        assert n == len(ek)
        k = io.empty(n, 8)
        v = io.empty(n, 8)
        assert self.d.dump_sorted(k, v, n) == n
        assert np.array_equal(self.down(k, n), ek)
        assert np.array_equal(self.down(v, n), ev)


def _keys(rng, n, span):
    if span == 0:  # full 64-bit range
        return rng.integers(0, 2**64, n, dtype=np.uint64, endpoint=False)
    return rng.integers(0, span, n, dtype=np.uint64)


def _queries(rng, live, n, span):
    q = _keys(rng, n, span)
    q[:len(EDGE)] = EDGE
    if len(live):
        half = n // 2
        q[len(EDGE):half] = live[rng.integers(0, len(live), half - len(EDGE))]
    return q


def _ranges(rng, n, span, width):
    lo = _keys(rng, n, span)
    w = rng.integers(0, width, n, dtype=np.uint64)
    hi = lo + np.minimum(w, ALL - lo)  # no wrap-around
    lo[:8], hi[:8] = hi[:8].copy(), lo[:8].copy()  # some lo > hi
    lo[8], hi[8] = 0, ALL  # the whole key space
    return lo, hi


@pytest.mark.parametrize("seed,span", [(1, 1 << 12), (2, 1 << 40), (3, 0)])
def test_random_mixed_stream(gd, orc, io, seed, span):
    """Small and large batches (fused and radix-sorted admission), erases of
    live keys, mixed batches; full check every few batches."""
    rng = np.random.default_rng(seed)
    p = Pair64(gd, orc, io, max_elements=1 << 19, batch_capacity=1 << 13)
    live = np.empty(0, np.uint64)
    for b in range(30):
        n = int(rng.integers(1, 1500)) if b % 3 else int(
            rng.integers(3000, 20000))
        keys = _keys(rng, n, span)
        if b == 0:
            keys[:len(EDGE)] = EDGE[:min(n, len(EDGE))]
        vals = rng.integers(0, 2**64, n, dtype=np.uint64, endpoint=False)
        kind = b % 4
        if kind in (0, 1):
            p.insert(keys, vals)
        elif kind == 2 and len(live):
            p.erase(live[rng.integers(0, len(live), n)])
        else:
            erase = (rng.random(n) < 0.3).astype(np.uint8)
            p.mixed(keys, vals, erase)
        live = p.o.dump()[0]
        if b % 5 == 4 or b == 29:
            p.check_contents()
            q = _queries(rng, live, 4000, span)
            p.check_lookup(q, with_found=bool(b % 2))
            lo, hi = _ranges(rng, 600, span,
                             max((span or 2**64) // 64, 2) if span else 2**58)
            p.check_ranges(lo, hi)
            p.check_successor(q)


def test_edges_and_empty(gd, orc, io):
    p = Pair64(gd, orc, io, max_elements=1 << 12, batch_capacity=64)
    p.check_lookup(EDGE, True)
    p.check_lookup(EDGE, False)
    p.check_successor(EDGE)
    p.check_ranges(np.array([0, 5], np.uint64), np.array([ALL, 4], np.uint64))
    p.check_contents()
    vals = np.array([1, ALL, 3, 0, 5, 6, 7, 8], np.uint64)
    p.insert(EDGE, vals)
    p.check_contents()
    p.check_lookup(EDGE, True)
    p.check_lookup(EDGE, False)
    p.check_successor(np.concatenate([EDGE, EDGE + np.uint64(1)]))
    p.check_ranges(np.array([0, 0x7FFFFFFFFFFFFFFF, ALL, 2], np.uint64),
                   np.array([ALL, 0x8000000000000000, ALL, 2], np.uint64))
    p.erase(np.array([0x8000000000000000, 0, ALL], np.uint64))
    p.check_contents()
    p.check_lookup(EDGE, True)
    p.check_successor(EDGE)


def test_bulk_build_and_merges(gd, orc, io):
    """bulk_build (last occurrence wins), then equal batches that cascade
    through the binary-counter merges, with tombstones dropped at the
    oldest run."""
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 1 << 20, 100000, dtype=np.uint64) << np.uint64(40)
    vals = np.arange(100000, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    p = Pair64(gd, orc, io, max_elements=1 << 20, batch_capacity=1 << 14)
    p.d.bulk_build(p.up(keys), p.up(vals), count=len(keys))
    assert p.o.bulk_build(keys, vals) == 0
    p.check_contents()
    for b in range(16):
        ins = rng.integers(0, 1 << 20, 1 << 14, dtype=np.uint64) << \
            np.uint64(40)
        p.insert(ins, ins ^ np.uint64(b))
        p.erase(p.o.dump()[0][rng.integers(0, p.o.size(), 2000)])
    p.check_contents()
    live = p.o.dump()[0]
    q = _queries(rng, live, 1 << 15, 0)
    p.check_lookup(q, True)
    p.check_successor(q[:8192])
    lo, hi = _ranges(rng, 2000, 0, 2**50)
    p.check_ranges(lo, hi)


def test_grouped_lookups(gd, orc, io):
    """Lookups above the grouping threshold (partition by the top key byte,
    grouped search, gather back)."""
    rng = np.random.default_rng(4)
    p = Pair64(gd, orc, io, tuning={"grouped_lookup_min_log": 10},
               max_elements=1 << 18, batch_capacity=1 << 15)
    for _ in range(4):
        k = _keys(rng, 1 << 15, 0)
        p.insert(k, ~k)
    q = _queries(rng, p.o.dump()[0], 1 << 16, 0)
    p.check_lookup(q, True)
    p.check_lookup(q, False)


def test_server_64bit_matches_replay(gd, orc, io):
    """The serving engine over 64-bit keys: every lookup and range result
    equals the oracle replayed in step order (updates of a step first)."""
    rng = np.random.default_rng(17)
    cfg = dict(max_elements=1 << 20, batch_capacity=1 << 14)
    srv = gd.Server(**cfg, key_bytes=8, max_inflight_steps=3)
    reqs = []
    inserted = [np.zeros(1, np.uint64)]
    for i in range(150):
        size = int(np.exp(rng.uniform(0, np.log(8000)))) or 1
        r = rng.random()
        if r < 0.4:
            k = _keys(rng, size, 1 << 48)
            reqs.append(("insert", k, k * np.uint64(3)))
            inserted.append(k)
        elif r < 0.5:
            pool = np.concatenate(inserted)
            reqs.append(("erase", pool[rng.integers(0, len(pool), size)],
                         None))
        elif r < 0.8:
            pool = np.concatenate(inserted)
            k = _keys(rng, size, 1 << 48)
            k[: size // 2] = pool[rng.integers(0, len(pool), size // 2)]
            reqs.append(("lookup", k, None))
        else:
            lo, hi = _ranges(rng, max(size, 10), 1 << 48, 1 << 40)
            reqs.append(("rcount" if r < 0.9 else "rsum", lo, hi))
    ops = {"insert": 0, "erase": 1, "lookup": 2, "rcount": 3, "rsum": 4}
    staged = []
    for kind, a, b in reqs:
        da = io.up(a, np.uint64)
        db = io.up(b, np.uint64) if b is not None else None
        out = io.empty(len(a), 8) if kind not in ("insert", "erase") else None
        found = io.empty(len(a), 1) if kind == "lookup" else None
        staged.append((kind, da, db, out, found, len(a)))
    io.torch.cuda.synchronize()
    tickets = [srv.submit(ops[k], da, db, out, found, count=n)
               for k, da, db, out, found, n in staged]
    steps = [srv.wait(t) for t in tickets]
    o = orc.LsmOracle(key_bytes=8, **cfg)
    by_step = {}
    for i, s in enumerate(steps):
        by_step.setdefault(s, []).append(i)
    for s in sorted(by_step):
        for i in by_step[s]:
            kind, a, b = reqs[i]
            if kind == "insert":
                assert o.insert(a, b) == 0
            elif kind == "erase":
                assert o.erase(a) == 0
        for i in by_step[s]:
            kind, a, b = reqs[i]
            out, found = staged[i][3], staged[i][4]
            if kind == "lookup":
                ev, ef = o.lookup(a, with_found=True)
                assert np.array_equal(io.down(out, len(a), np.uint64), ev)
                assert np.array_equal(io.down(found, len(a), np.uint8), ef)
            elif kind == "rcount":
                assert np.array_equal(io.down(out, len(a), np.uint64),
                                      o.range_count(a, b))
            elif kind == "rsum":
                assert np.array_equal(io.down(out, len(a), np.uint64),
                                      o.range_sum(a, b))
    srv.close()


def test_snapshot_round_trip_64(gd, orc, io, tmp_path):
    rng = np.random.default_rng(8)
    p = Pair64(gd, orc, io, max_elements=1 << 16, batch_capacity=1 << 12)
    for _ in range(6):
        k = _keys(rng, 3000, 0)
        p.mixed(k, ~k, (rng.random(3000) < 0.2).astype(np.uint8))
    path = str(tmp_path / "lsm64.snap")
    assert p.d.save(path) == p.o.size()
    fresh = Pair64(gd, orc, io, max_elements=1 << 16, batch_capacity=1 << 12)
    fresh.d.restore(path)
    k, v = p.o.dump()
    assert fresh.o.bulk_build(k, v) == 0
    fresh.check_contents()
    fresh.check_lookup(_queries(rng, k, 4000, 0), True)


def test_compact_64(gd, orc, io):
    """compact() for 64-bit keys: one run of exactly the live keys, same
    answers before and after."""
    rng = np.random.default_rng(77)
    p = Pair64(gd, orc, io, max_elements=1 << 19, batch_capacity=1 << 13)
    for b in range(19):
        n = int(rng.integers(100, 1 << 13))
        keys = _keys(rng, n, 1 << 40)
        vals = rng.integers(0, 2**64, n, dtype=np.uint64, endpoint=False)
        erase = (rng.random(n) < 0.3).astype(np.uint8)
        p.mixed(keys, vals, erase)
    runs, rows = p.d.shape()
    live = p.o.dump()[0]
    assert rows > len(live)  # tombstones / shadowed versions held
    p.d.compact()
    assert p.d.shape() == (1, len(live))
    p.check_contents()
    q = _queries(rng, live, 6000, 1 << 40)
    p.check_lookup(q, with_found=True)
    lo, hi = _ranges(rng, 800, 1 << 40, 1 << 34)
    p.check_ranges(lo, hi)
    p.check_successor(q)


def test_pool_cache_bounded_and_trim(gd, dev):
    """The device pools keep at most max(4 GiB, live / 2) cached per pool
    while a 2^28-key (4 GiB of rows) dictionary is built, merged with new
    batches and compacted; held = live + cached at every point; trim()
    returns the cache, and the answers do not change."""
    import torch
    n = 1 << 28
    mul = -7046029254386353131  # 0x9E3779B97F4A7C15 as int64 (odd)
    keys = torch.arange(n, dtype=torch.int64, device=dev) * mul
    vals = keys ^ 0x5DEECE66D
    d = gd.GPULSMOpt(key_bytes=8, max_elements=n + (1 << 22),
                     batch_capacity=1 << 20)
    floor = 4 << 30

    def check(m):
        assert m["held_bytes"] == m["live_bytes"] + m["cached_bytes"]
        assert m["peak_held_bytes"] >= m["held_bytes"]
        assert m["cached_bytes"] <= 2 * floor + m["live_bytes"] // 2, m

    d.bulk_build(keys, vals)
    check(d.memory())
    for b in range(4):
        k = keys[b << 20:(b + 1) << 20] + 1  # odd offsets: new keys
        d.insert(k, k ^ 0x5DEECE66D)
        check(d.memory())
    d.compact()
    torch.cuda.synchronize()
    m = d.memory()
    check(m)
    assert m["row_bytes_ub"] == (n + (4 << 20)) * 16
    assert m["live_bytes"] >= m["row_bytes_ub"]
    h = 1 << 16
    # present (bulk), present (inserted), present (bulk), absent
    q = torch.cat([keys[:h], keys[:h] + 1, keys[-h:], keys[:h] + 2])
    want = q ^ 0x5DEECE66D
    found_want = torch.ones(q.numel(), dtype=torch.uint8, device=dev)
    found_want[3 * h:] = 0
    out = torch.empty_like(q)
    found = torch.empty(q.numel(), dtype=torch.uint8, device=dev)
    d.lookup(q, out, found)
    torch.cuda.synchronize()
    assert torch.equal(found, found_want)
    assert torch.equal(out[found.bool()], want[found.bool()])
    d.trim()
    m = d.memory()
    assert m["cached_bytes"] == 0 and m["held_bytes"] == m["live_bytes"]
    out2 = torch.empty_like(q)
    d.lookup(q, out2, found)
    torch.cuda.synchronize()
    assert torch.equal(out2[found.bool()], want[found.bool()])
    d.close()
