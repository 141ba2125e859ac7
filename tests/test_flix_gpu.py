"""GPU parity of the FliX updatable index (libgdict.so) against the oracle,
plus the known answer of the reference's interface test (FliX/main.cu:98-156)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KT = {4: np.uint32, 8: np.uint64}


class FlixPair:
    def __init__(self, gd, orc, io, kb, vb, **cfg):
        self.f = gd.FlixIndex(key_bytes=kb, value_bytes=vb, **cfg)
        self.o = orc.FlixOracle(kb, vb)
        self.io = io
        self.kb, self.vb = kb, vb
        self.kt, self.vt = KT[kb], KT[vb]

    def up_k(self, a):
        return self.io.up(np.asarray(a, self.kt), self.kt)

    def up_v(self, a):
        return self.io.up(np.asarray(a, self.vt), self.vt)

    def build(self, keys, max_size=0):
        self.f.build_bucket_layer_only(self.up_k(keys), max_size=max_size,
                                       count=len(keys))
        self.o.build(keys)

    def insert(self, keys, vals):
        self.f.insert(self.up_k(keys), self.up_v(vals), count=len(keys))
        self.o.insert(keys, vals)

    def remove(self, keys, assume_sorted=False):
        self.f.remove(self.up_k(keys), count=len(keys),
                      assume_sorted=assume_sorted)
        self.o.remove(keys)

    def lookup(self, q, sorted_api=False):
        out = self.io.empty(len(q), self.vb)
        if sorted_api:
            self.f.lookups_ordered(self.up_k(q), out, count=len(q))
        else:
            self.f.lookup(self.up_k(q), out, count=len(q))
        return self.io.down(out, len(q), self.vt)

    def successor(self, q):
        out = self.io.empty(len(q), self.kb)
        self.f.lookups_successor(self.up_k(q), out, count=len(q))
        return self.io.down(out, len(q), self.kt)

    def range(self, lo, hi):
        n = len(lo)
        import torch
        offs = torch.zeros(n + 1, dtype=torch.int64, device=self.io.device)
        dlo, dhi = self.up_k(lo), self.up_k(hi)
        self.f.range(dlo, dhi, offs, count=n)
        o = self.io.down(offs, n + 1, np.uint64)
        total = int(o[-1])
        ok = self.io.empty(total, self.kb)
        ov = self.io.empty(total, self.vb)
        self.f.range(dlo, dhi, offs, ok, ov, count=n)
        return (o, self.io.down(ok, total, self.kt),
                self.io.down(ov, total, self.vt))

    def check_all(self, rng, span_lo, span_hi, nq=4000):
        ek, ev = self.o.dump()
        # full enumeration through the range API = sorted contents
        o, k, v = self.range(np.array([0], self.kt),
                             np.array([np.iinfo(self.kt).max], self.kt))
        assert np.array_equal(k, ek)
        assert np.array_equal(v, ev)
        assert self.f.compute_total_size() == len(ek)
        q = rng.integers(span_lo, span_hi, nq, dtype=np.uint64).astype(self.kt)
        if len(ek):
            q[: nq // 2] = ek[rng.integers(0, len(ek), nq // 2)]
        assert np.array_equal(self.lookup(q), self.o.lookup(q))
        qs = np.sort(q)
        assert np.array_equal(self.lookup(qs, True), self.o.lookup(qs))
        assert np.array_equal(self.successor(qs), self.o.successor(qs))
        lo = np.sort(rng.integers(span_lo, span_hi, 300, dtype=np.uint64)
                     .astype(self.kt))
        w = rng.integers(0, max((span_hi - span_lo) // 200, 2), 300,
                         dtype=np.uint64)
        hi = np.minimum(lo.astype(np.uint64) + w,
                        np.uint64(np.iinfo(self.kt).max)).astype(self.kt)
        lo[:10], hi[:10] = hi[:10], lo[:10]
        go, gk, gv = self.range(lo, hi)
        eo, ekk, evv = self.o.range(lo, hi)
        assert np.array_equal(go, eo)
        assert np.array_equal(gk, ekk)
        assert np.array_equal(gv, evv)


def test_interface_known_answer(gd, io):
    """FliX/main.cu:98-156, expectations derived by hand: build values are
    input positions, inserts skip nothing new, deletes remove six keys."""
    f = gd.FlixIndex(key_bytes=4, value_bytes=4)
    build = np.array([100, 90, 80, 70, 60, 50, 40, 30, 20, 10], np.uint32)
    ins = np.arange(5, 106, 10, dtype=np.uint32)
    offs = np.arange(10, 21, dtype=np.uint32)
    dels = np.array([30, 35, 40, 45, 75, 80], np.uint32)
    probes = np.arange(5, 106, 5, dtype=np.uint32)
    NF = 0xFFFFFFFF

    def show():
        out = io.empty(len(probes))
        f.lookups_ordered(io.up(probes), out, count=len(probes))
        return io.down(out, len(probes))

    f.build_bucket_layer_only(io.up(build), max_size=21, count=10)
    exp = np.array([NF if p % 10 else (100 - p) // 10 for p in probes],
                   np.uint32)
    assert np.array_equal(show(), exp)
    f.insert(io.up(ins), io.up(offs), count=len(ins))
    exp = np.array([(100 - p) // 10 if p % 10 == 0 else 10 + (p - 5) // 10
                    for p in probes], np.uint32)
    assert np.array_equal(show(), exp)
    f.remove(io.up(dels), count=len(dels))
    exp[np.isin(probes, dels)] = NF
    assert np.array_equal(show(), exp)
    assert f.compute_total_size() == 21 - 6


@pytest.mark.parametrize("kb,vb", [(4, 4), (8, 8), (4, 8), (8, 4)])
def test_random_updates_parity(gd, orc, io, kb, vb):
    rng = np.random.default_rng(kb * 10 + vb)
    kt, vt = KT[kb], KT[vb]
    hi_key = (1 << 32) - 4 if kb == 4 else (1 << 62)
    p = FlixPair(gd, orc, io, kb, vb)
    keys = rng.integers(2, hi_key, 20000, dtype=np.uint64).astype(kt)
    keys[:500] = keys[500:1000]  # duplicates in the build input
    p.build(keys, max_size=200000)
    p.check_all(rng, 2, hi_key)
    for step in range(8):
        n = int(rng.integers(1, 6000))
        ins = rng.integers(2, hi_key, n, dtype=np.uint64).astype(kt)
        live = p.o.dump()[0]
        if len(live):
            ins[: n // 4] = live[rng.integers(0, len(live), n // 4)]  # present
        ins[n // 4: n // 4 + n // 8] = ins[:n // 8]  # in-batch duplicates
        vals = rng.integers(0, 1 << 31, n, dtype=np.uint64).astype(vt)
        p.insert(ins, vals)
        live = p.o.dump()[0]
        m = int(rng.integers(1, 5000))
        dels = rng.integers(2, hi_key, m, dtype=np.uint64).astype(kt)
        dels[: m * 3 // 4] = live[rng.integers(0, len(live), m * 3 // 4)]
        p.remove(dels)
        p.check_all(rng, 2, hi_key)
    p.f.rebuild_bucket_layer_only()
    p.check_all(rng, 2, hi_key)
    p.insert(rng.integers(2, hi_key, 3000, dtype=np.uint64).astype(kt),
             np.arange(3000, dtype=vt))
    p.check_all(rng, 2, hi_key)


def test_clustered_runs_and_skewed_inserts(gd, orc, io):
    """Config 3 shape, scaled: clustered runs of consecutive uint64 keys at
    random offsets; inserts concentrated into one bucket (node splits over
    many pieces); deletes that empty whole buckets."""
    rng = np.random.default_rng(3)
    p = FlixPair(gd, orc, io, 8, 8, node_slots=32, fill_percent=50)
    starts = rng.integers(2, 1 << 62, 64, dtype=np.uint64)
    keys = (starts[:, None] + np.arange(256, dtype=np.uint64)).ravel()
    p.build(keys, max_size=1 << 16)
    p.check_all(rng, 2, 1 << 62)
    # 10000 new keys inside one run's gap: all land in a single bucket
    base = int(starts[0]) + 300
    dense = (np.uint64(base) + np.arange(0, 20000, 2, dtype=np.uint64))
    p.insert(dense, dense * np.uint64(3))
    p.check_all(rng, 2, 1 << 62)
    # delete two whole runs and half of the dense block
    p.remove(np.concatenate([keys[:512], dense[::2]]))
    p.check_all(rng, 2, 1 << 62)
    p.f.rebuild_bucket_layer_only()
    p.check_all(rng, 2, 1 << 62)


def test_allocation_region_growth(gd, orc, io):
    """A batch needing more nodes than a too-small max_size left in the
    allocation region is not refused and not applied in part: the index
    rebuilds itself with a larger region first, then takes the whole
    batch (the oracle's contents after it)."""
    rng = np.random.default_rng(9)
    p = FlixPair(gd, orc, io, 4, 4, node_slots=16)
    keys = np.arange(2, 4002, dtype=np.uint32) * 16
    p.build(keys, max_size=4000)
    before = p.f.gpu_resident_bytes()
    ins = rng.integers(2, 70000, 60000, dtype=np.uint64).astype(np.uint32)
    p.insert(ins, ins)
    assert p.f.gpu_resident_bytes() > before
    p.check_all(rng, 2, 70000)
    p.insert(rng.integers(2, 1 << 20, 200000, dtype=np.uint64)
             .astype(np.uint32), np.arange(200000, dtype=np.uint32))
    p.check_all(rng, 2, 1 << 20)


@pytest.mark.parametrize("kb", [4, 8])
def test_sorted_remove(gd, orc, io, kb):
    """remove(assume_sorted=True), the reference's sorted-keys contract:
    ascending batches with duplicates and absent keys, slices small enough
    for the wave's LDS copy and a slice of 3000 keys inside one bucket
    (searched in global memory), all against the oracle."""
    rng = np.random.default_rng(31 + kb)
    kt = KT[kb]
    p = FlixPair(gd, orc, io, kb, kb, node_slots=32)
    keys = (np.arange(1, 30001, dtype=np.uint64) * 100003).astype(kt)
    p.build(keys, max_size=60000)
    ins = rng.integers(2, 30001 * 100003, 20000, dtype=np.uint64).astype(kt)
    p.insert(ins, ins)
    live = p.o.dump()[0]
    for r in range(3):
        d = np.sort(np.concatenate([
            live[rng.integers(0, len(live), 5000)],  # with duplicates
            rng.integers(2, 30001 * 100003, 2000, dtype=np.uint64)
            .astype(kt)]))
        p.remove(d, assume_sorted=True)
        live = p.o.dump()[0]
        p.check_all(rng, 2, 30001 * 100003)
    # 3000 absent keys between two stored keys, plus the keys around them
    base = int(keys[777])
    d = np.concatenate([[base], np.arange(base + 1, base + 3001),
                        [int(keys[778])]]).astype(kt)
    p.remove(d, assume_sorted=True)
    p.check_all(rng, 2, 30001 * 100003)


@pytest.mark.parametrize("kb,vb", [(4, 4), (8, 8)])
def test_grouped_any_order_lookups(gd, orc, io, kb, vb):
    """Any-order lookups of 2^18 queries or more are grouped by their top
    varying key byte, looked up and gathered back: every value (hits,
    misses, duplicate queries) equals the oracle's, and equals the sorted
    API on the sorted queries."""
    rng = np.random.default_rng(41 + kb)
    kt = KT[kb]
    p = FlixPair(gd, orc, io, kb, vb, node_slots=32)
    top = (1 << 31) if kb == 4 else (1 << 60)
    keys = np.unique(rng.integers(2, top, 1 << 20, dtype=np.uint64)).astype(kt)
    p.build(keys)
    p.insert(rng.integers(2, top, 1 << 16, dtype=np.uint64).astype(kt),
             rng.integers(0, 1 << 31, 1 << 16, dtype=np.uint64).astype(KT[vb]))
    p.remove(keys[rng.integers(0, len(keys), 1 << 15)])
    nq = (1 << 18) + 12345
    q = rng.integers(2, top, nq, dtype=np.uint64).astype(kt)
    live = p.o.dump()[0]
    q[: nq // 2] = live[rng.integers(0, len(live), nq // 2)]
    q[nq // 2: nq // 2 + 1000] = q[:1000]  # repeated queries
    got = p.lookup(q)
    assert np.array_equal(got, p.o.lookup(q))
    qs = np.sort(q)
    assert np.array_equal(p.lookup(qs, True), p.o.lookup(qs))


def test_memory_stats_and_trim(gd, orc, io):
    """gd_flix_memory: held = live + cached, the peak bounds the held bytes,
    the node layer's bytes grow with a growth rebuild; trim returns the
    cache and changes no answer."""
    rng = np.random.default_rng(21)
    p = FlixPair(gd, orc, io, 8, 8, node_slots=32)
    keys = np.unique(rng.integers(2, 1 << 40, 50000, dtype=np.uint64))
    p.build(keys, max_size=60000)

    def check(m):
        assert m["held_bytes"] == m["live_bytes"] + m["cached_bytes"]
        assert m["peak_held_bytes"] >= m["held_bytes"]
        assert m["held_bytes"] == p.f.gpu_resident_bytes()
        assert m["live_bytes"] >= m["row_bytes_ub"] > 0

    m0 = p.f.memory()
    check(m0)
    ins = rng.integers(2, 1 << 40, 400000, dtype=np.uint64)
    p.insert(ins, ins ^ np.uint64(0x5DEECE66D))  # grows the region
    m1 = p.f.memory()
    check(m1)
    assert m1["row_bytes_ub"] > m0["row_bytes_ub"]
    p.remove(ins[:100000])
    check(p.f.memory())
    p.f.trim()
    m2 = p.f.memory()
    assert m2["cached_bytes"] == 0 and m2["held_bytes"] == m2["live_bytes"]
    check(m2)
    p.check_all(rng, 2, 1 << 40)


@pytest.mark.parametrize("slots", [8, 32])
def test_long_chains(gd, orc, io, slots):
    """Small batches aimed at a few buckets grow their chains past what one
    wave rewrites in registers (16 nodes), so later batches take the
    workgroup path with small slices, and large batches take it too; the
    contents and answers stay the oracle's after every batch."""
    rng = np.random.default_rng(slots)
    p = FlixPair(gd, orc, io, 4, 4, node_slots=slots, fill_percent=50)
    keys = np.arange(1, 2001, dtype=np.uint32) * 1000
    p.build(keys, max_size=1 << 12)
    for t in range(24):
        # keys inside the gaps of three buckets' ranges
        lo = np.array([5000, 900000, 1500000], np.uint64)
        k = (lo[rng.integers(0, 3, 40 + 30 * t)] +
             rng.integers(1, 20000, 40 + 30 * t, dtype=np.uint64))
        k = k.astype(np.uint32)
        p.insert(k, k ^ np.uint32(0x5A5A5A5A))
        if t % 6 == 5:
            p.check_all(rng, 2, 2100000)
    p.remove(p.o.dump()[0][::3])
    p.check_all(rng, 2, 2100000)


def test_rebuild_policy(gd, orc, io):
    """rebuild_chain_nodes: inserts that grow a chain past the threshold
    rebuild the bucket layer; the answers stay the oracle's."""
    rng = np.random.default_rng(41)
    p = FlixPair(gd, orc, io, 8, 8, node_slots=16, rebuild_chain_nodes=3)
    keys = np.unique(rng.integers(2, 1 << 40, 20000, dtype=np.uint64))
    p.build(keys, max_size=len(keys) + 100000)
    for t in range(6):
        ins = rng.integers(2, 1 << 38, 15000, dtype=np.uint64)
        p.insert(ins, ins + np.uint64(7))
        p.remove(p.o.dump()[0][t::9])
        p.check_all(rng, 2, 1 << 40)


def _check_queries(p, q, lo, hi):
    """Lookups (any-order and sorted API), successors of the sorted
    queries, and range enumeration, all against the oracle."""
    assert np.array_equal(p.lookup(q), p.o.lookup(q))
    qs = np.sort(q)
    assert np.array_equal(p.lookup(qs, True), p.o.lookup(qs))
    assert np.array_equal(p.successor(qs), p.o.successor(qs))
    go, gk, gv = p.range(lo, hi)
    eo, ek, ev = p.o.range(lo, hi)
    assert np.array_equal(go, eo)
    assert np.array_equal(gk, ek)
    assert np.array_equal(gv, ev)
    return int(eo[-1])


@pytest.mark.parametrize("scale_log,batches", [(6, 8)])
def test_config3_scaled(gd, orc, io, scale_log, batches):
    """Config 3 (FliX, uint64 keys and values) at 1/64 of its size: a build
    of 2^21 keys in clustered runs of 2^12 (131,072 buckets: the
    multi-block work list), then batches of 2^16 mixed ops (75% inserts of
    new keys, 25% deletes of live keys), each followed by 2^18 lookups,
    successors and 2^12 range enumerations of 16..1024 pairs, all against
    the oracle; full contents and total size every other batch."""
    import importlib
    wl = importlib.import_module("gdict_amd.workload")
    w = wl.config3(scale_log)
    built = w.build_keys()
    p = FlixPair(gd, orc, io, 8, 8)
    p.build(built, max_size=len(built) + batches * w.batch)
    assert p.f.compute_total_size() == len(built)
    pairs = 0
    for t in range(batches):
        ins, vals, dels = w.batch_ops(t, built)
        p.insert(ins, vals)
        p.remove(dels)
        q, lo, hi = w.queries(t, built)
        pairs += _check_queries(p, q, lo, hi)
        if t % 2 == 1 or t == batches - 1:
            ek, ev = p.o.dump()
            assert p.f.compute_total_size() == len(ek)
            o, k, v = p.range(np.array([0], np.uint64),
                              np.array([2**64 - 1], np.uint64))
            assert np.array_equal(k, ek)
            assert np.array_equal(v, ev)
    assert pairs > 16 * w.ranges * batches // 2  # k-pair ranges were served


def test_reference_shape_scaled(gd, orc, io):
    """The shape of FliX's recorded run (FliX/build/sample_output: 32-bit
    keys, 32-slot nodes at 50%, inserts with 90% of the keys in the lowest
    quarter of the key range, then deletes of live keys) at 1/32 of its
    size: many buckets split in every batch, so many waves take nodes of
    the allocation region at once; contents and answers after every step."""
    rng = np.random.default_rng(25)
    top = (1 << 32) - 4
    p = FlixPair(gd, orc, io, 4, 4, node_slots=32, fill_percent=50)
    build = np.unique(rng.integers(2, top, 1 << 20, dtype=np.uint64)
                      .astype(np.uint32))
    n_step = 419430  # 13,421,772 / 32
    p.build(build, max_size=len(build) + 5 * n_step)

    def check():
        ek, ev = p.o.dump()
        o, k, v = p.range(np.array([0], np.uint32),
                          np.array([0xFFFFFFFF], np.uint32))
        assert np.array_equal(k, ek) and np.array_equal(v, ev)
        assert p.f.compute_total_size() == len(ek)
        q = ek[rng.integers(0, len(ek), 1 << 16)]
        q[::3] = rng.integers(2, top, len(q[::3]), dtype=np.uint64)
        lo = np.sort(rng.integers(2, top, 2000, dtype=np.uint64)
                     .astype(np.uint32))
        hi = np.minimum(lo.astype(np.uint64) + rng.integers(
            0, 1 << 16, 2000, dtype=np.uint64), top).astype(np.uint32)
        _check_queries(p, q, lo, hi)

    for step in range(5):
        dense = rng.random(n_step) < 0.9
        keys = np.where(dense,
                        rng.integers(2, top // 4, n_step, dtype=np.uint64),
                        rng.integers(top // 4, top, n_step, dtype=np.uint64)
                        ).astype(np.uint32)
        vals = rng.integers(0, 1 << 32, n_step, dtype=np.uint64).astype(
            np.uint32)
        p.insert(keys, vals)
        check()
    for step in range(5):
        live = p.o.dump()[0]
        p.remove(np.sort(live[rng.choice(len(live), n_step, replace=False)]))
        check()


def test_operations_on_alternating_streams(gd, orc, io):
    """Inserts and removes of one FliX index issued on alternating streams,
    two calls per stream, with no host wait between them: each call sees
    the state every earlier call left, and the contents and lookups match
    the oracle."""
    import torch
    rng = np.random.default_rng(71)
    p = FlixPair(gd, orc, io, 8, 8, node_slots=32)
    keys = np.unique(rng.integers(2, 1 << 40, 1 << 16, dtype=np.uint64))
    p.build(keys, max_size=1 << 19)
    streams = [torch.cuda.Stream(), torch.cuda.Stream(), None]
    keep = []  # device inputs stay alive until the streams finish
    for b in range(16):
        s = streams[(b // 2) % 3]
        if b % 4 == 3:
            live = p.o.dump()[0]
            k = live[rng.integers(0, len(live), 4000)]
            dk = p.up_k(k)
            torch.cuda.synchronize()
            keep.append(dk)
            p.f.remove(dk, count=len(k), stream=s)
            p.o.remove(k)
        else:
            k = rng.integers(2, 1 << 40, 6000, dtype=np.uint64)
            v = rng.integers(0, 1 << 62, 6000, dtype=np.uint64)
            dk, dv = p.up_k(k), p.up_v(v)
            torch.cuda.synchronize()
            keep += [dk, dv]
            p.f.insert(dk, dv, count=len(k), stream=s)
            p.o.insert(k, v)
    torch.cuda.synchronize()
    p.check_all(rng, 2, 1 << 40)


@pytest.mark.parametrize("kb", [4, 8])
def test_range_edges(gd, orc, io, kb):
    """Range enumeration where its walk changes course: bounds on, just
    below and just above stored keys, single-key and empty ranges, ranges
    that start or end inside emptied buckets or span several of them,
    ranges that start in the middle of a bucket chain (a node after the
    head), below the smallest and above the largest key, and lo > hi;
    offsets, keys and values equal the oracle's."""
    rng = np.random.default_rng(51 + kb)
    kt = KT[kb]
    p = FlixPair(gd, orc, io, kb, kb, node_slots=16, fill_percent=50)
    keys = (np.arange(1, 4001, dtype=np.uint64) * 1000).astype(kt)
    p.build(keys, max_size=1 << 16)
    # 600 keys between two stored keys: one bucket's chain of many nodes
    chain = (np.uint64(20001) + np.arange(600, dtype=np.uint64)).astype(kt)
    p.insert(chain, chain)
    # whole buckets emptied (8 keys per head at 50% of 16 slots)
    gone = keys[((keys >= 1000000) & (keys <= 1200000)) |
                ((keys >= 1500000) & (keys <= 1520000))]
    p.remove(gone)
    live = p.o.dump()[0].astype(np.uint64)
    top = int(np.iinfo(kt).max)
    lo, hi = [], []
    for k in live[rng.integers(0, len(live), 60)].tolist():
        for a, z in ((k, k), (k - 1, k), (k, k + 1), (k + 1, k + 2500),
                     (k - 5, k + 5), (k, k + 20000)):
            lo.append(a)
            hi.append(z)
    edges = [(0, top), (0, 5), (999990, 1000010), (1000005, 1199990),
             (1199995, 1200500), (1400000, 1600000), (20000, 21000),
             (20500, 20700), (20999, 21001), (20300, 1000000),
             (3999000, top), (4000001, top), (5, 9), (40010, 39990)]
    lo += [a for a, _ in edges]
    hi += [z for _, z in edges]
    lo = np.array(lo, np.uint64).astype(kt)
    hi = np.minimum(np.array(hi, np.uint64), np.uint64(top)).astype(kt)
    go, gk, gv = p.range(lo, hi)
    eo, ek, ev = p.o.range(lo, hi)
    assert np.array_equal(go, eo)
    assert np.array_equal(gk, ek)
    assert np.array_equal(gv, ev)
    # the mid-chain ranges hold what they should
    assert int(eo[-7] - eo[-8]) == 101  # 20500..20600 of the chain


@pytest.mark.parametrize("kb,vb", [(4, 8), (8, 4)])
def test_rebuild_after_emptying(gd, orc, io, kb, vb):
    """A rebuild writes each chain's rows straight into the new layer: an
    index emptied by deletes rebuilds into one empty bucket, answers every
    query as empty, and takes inserts again; a half-emptied one rebuilds
    into fewer buckets with the same contents."""
    rng = np.random.default_rng(61 + kb)
    kt, vt = KT[kb], KT[vb]
    p = FlixPair(gd, orc, io, kb, vb, node_slots=16, fill_percent=50)
    keys = np.unique(rng.integers(2, 1 << 30, 30000, dtype=np.uint64)
                     ).astype(kt)
    p.build(keys, max_size=1 << 16)
    p.insert(rng.integers(2, 1 << 30, 5000, dtype=np.uint64).astype(kt),
             rng.integers(0, 1 << 31, 5000, dtype=np.uint64).astype(vt))
    live = p.o.dump()[0]
    p.remove(live[::2])
    p.f.rebuild_bucket_layer_only()
    p.check_all(rng, 2, 1 << 30)
    p.remove(p.o.dump()[0])
    assert p.f.compute_total_size() == 0
    p.f.rebuild_bucket_layer_only()
    assert p.f.compute_total_size() == 0
    p.check_all(rng, 2, 1 << 30)
    p.insert(rng.integers(2, 1 << 30, 3000, dtype=np.uint64).astype(kt),
             np.arange(3000, dtype=vt))
    p.check_all(rng, 2, 1 << 30)


@pytest.mark.parametrize("kb", [4, 8])
def test_insert_of_stored_keys_only(gd, orc, io, kb):
    """A batch whose keys are all stored already (with duplicates inside
    it) inserts nothing and changes no value; a batch mixing such keys with
    new ones takes only the new ones."""
    rng = np.random.default_rng(81 + kb)
    kt = KT[kb]
    p = FlixPair(gd, orc, io, kb, kb, node_slots=32)
    keys = np.unique(rng.integers(2, 1 << 30, 20000, dtype=np.uint64)
                     ).astype(kt)
    p.build(keys, max_size=1 << 16)
    again = keys[rng.integers(0, len(keys), 5000)]
    p.insert(again, np.full(len(again), 7, kt))
    assert p.f.compute_total_size() == len(keys)
    p.check_all(rng, 2, 1 << 30)
    mixed = np.concatenate([again[:1000], rng.integers(
        2, 1 << 30, 1000, dtype=np.uint64).astype(kt)])
    p.insert(mixed, np.arange(len(mixed), dtype=kt))
    p.check_all(rng, 2, 1 << 30)
