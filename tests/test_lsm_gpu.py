"""GPU parity of the LSM dictionary (libgdict.so through its C ABI) against
the oracle, at sizes the oracle finishes in seconds, plus closed-form pins
taken from the reference's own validation (FliX/paper_lsmu_sweep.cu)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NOT_FOUND = 0xFFFFFFFF


class Pair:
    """GPU dictionary + oracle driven in lock step."""

    def __init__(self, gd, orc, io, tuning=None, **cfg):
        self.d = gd.GPULSMOpt(tuning=tuning, **cfg)
        self.o = orc.LsmOracle(**cfg)
        self.io = io
        self.gd = gd

    def insert(self, keys, vals):
        self.d.insert(self.io.up(keys), self.io.up(vals), count=len(keys))
        assert self.o.insert(keys, vals) == 0

    def erase(self, keys):
        self.d.erase(self.io.up(keys), count=len(keys))
        assert self.o.erase(keys) == 0

    def mixed(self, keys, vals, erase):
        self.d.apply_mixed(self.io.up(keys), self.io.up(vals),
                           self.io.up(erase, np.uint8), count=len(keys))
        assert self.o.apply_mixed(keys, vals, erase) == 0

    def check_lookup(self, q, with_found):
        io = self.io
        out = io.empty(len(q))
        found = io.empty(len(q), 1) if with_found else None
        self.d.lookup(io.up(q), out, found, count=len(q))
        if with_found:
            ev, ef = self.o.lookup(q, with_found=True)
            assert np.array_equal(io.down(out, len(q)), ev)
            assert np.array_equal(io.down(found, len(q), np.uint8), ef)
        else:
            assert np.array_equal(io.down(out, len(q)), self.o.lookup(q))

    def check_ranges(self, lo, hi):
        io = self.io
        s = io.empty(len(lo))
        c = io.empty(len(lo))
        self.d.range(io.up(lo), io.up(hi), s, count=len(lo))
        self.d.range_count(io.up(lo), io.up(hi), c, count=len(lo))
        assert np.array_equal(io.down(s, len(lo)), self.o.range_sum(lo, hi))
        assert np.array_equal(io.down(c, len(lo)), self.o.range_count(lo, hi))

    def check_successor(self, q):
        io = self.io
        out = io.empty(len(q))
        self.d.successor(io.up(q), out, count=len(q))
        assert np.array_equal(io.down(out, len(q)), self.o.successor(q))

    def check_contents(self):
        io = self.io
        n = self.d.dump_sorted()
        ek, ev = self.o.dump()
        assert n == len(ek)
        k = io.empty(n)
        v = io.empty(n)
        assert self.d.dump_sorted(k, v, n) == n
        assert np.array_equal(io.down(k, n), ek)
        assert np.array_equal(io.down(v, n), ev)


def _queries(rng, keys, n, span):
    q = rng.integers(0, span, n, dtype=np.uint64).astype(np.uint32)
    if len(keys):
        take = rng.integers(0, len(keys), n // 2)
        q[: n // 2] = keys[take]
    return q


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_mixed_stream_every_batch(gd, orc, io, seed):
    """Inserts, erases and mixed batches with duplicates; full check of the
    sorted contents, lookups (with and without found), ranges and successors
    after every batch."""
    rng = np.random.default_rng(seed)
    span = [1 << 12, 1 << 20, 1 << 32][seed - 1]
    p = Pair(gd, orc, io, max_elements=1 << 18, batch_capacity=1 << 11)
    live = np.empty(0, np.uint32)
    for b in range(40):
        n = int(rng.integers(1, 5000))
        keys = rng.integers(0, span, n, dtype=np.uint64).astype(np.uint32)
        vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        kind = b % 5
        if kind in (0, 1, 2):
            p.insert(keys, vals)
        elif kind == 3 and len(live):
            p.erase(live[rng.integers(0, len(live), n)])
        else:
            erase = (rng.random(n) < 0.3).astype(np.uint8)
            p.mixed(keys, vals, erase)
        live = p.o.dump()[0]
        if b % 4 == 3 or b == 39:
            p.check_contents()
            q = _queries(rng, live, 3000, span)
            p.check_lookup(q, with_found=bool(b % 8 == 3))
            lo = rng.integers(0, span, 500, dtype=np.uint64)
            w = rng.integers(0, max(span // 64, 2), 500, dtype=np.uint64)
            hi = np.minimum(lo + w, 0xFFFFFFFF)
            lo[:20], hi[:20] = hi[:20], lo[:20]  # some lo > hi
            p.check_ranges(lo.astype(np.uint32), hi.astype(np.uint32))
            p.check_successor(q)


def test_edge_keys_and_empty(gd, orc, io):
    p = Pair(gd, orc, io, max_elements=1 << 12, batch_capacity=64)
    q = np.array([0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFE, 0xFFFFFFFF],
                 np.uint32)
    # empty dictionary
    p.check_lookup(q, True)
    p.check_lookup(q, False)
    p.check_successor(q)
    p.check_ranges(np.array([0, 5], np.uint32),
                   np.array([0xFFFFFFFF, 4], np.uint32))
    p.check_contents()
    # zero-count calls are no-ops
    p.d.insert(io.up(q), io.up(q), count=0)
    p.d.erase(io.up(q), count=0)
    # extreme keys, both halves of the key space, and a value of 0xFFFFFFFF
    keys = np.array([0, 0xFFFFFFFF, 0x80000000, 0x7FFFFFFF, 5, 5],
                    np.uint32)
    vals = np.array([10, 0xFFFFFFFF, 12, 13, 14, 15], np.uint32)
    p.insert(keys, vals)
    p.check_contents()
    p.check_lookup(q, True)
    p.check_successor(np.array([0, 6, 0x7FFFFFFF, 0x80000001, 0xFFFFFFFF],
                               np.uint32))
    p.check_ranges(np.array([0, 0x7FFFFFFF, 0, 0xFFFFFFFF], np.uint32),
                   np.array([0xFFFFFFFF, 0x80000000, 0, 0xFFFFFFFF],
                            np.uint32))
    p.erase(np.array([0x80000000, 0], np.uint32))
    p.check_contents()
    p.check_lookup(q, True)
    p.check_successor(q)


def test_invalid_arguments(gd, io):
    d = gd.GPULSMOpt(max_elements=1024, batch_capacity=16)
    with pytest.raises(gd.InvalidArgument):
        d.insert(None, None, count=4)
    with pytest.raises(gd.InvalidArgument):
        d.lookup(None, None, count=4)
    with pytest.raises(ValueError):
        gd.GPULSMOpt(max_elements=1 << 33)


def test_bulk_build_last_occurrence_wins(gd, orc, io):
    rng = np.random.default_rng(11)
    keys = rng.integers(0, 1 << 16, 200000, dtype=np.uint32)
    vals = np.arange(200000, dtype=np.uint32)
    p = Pair(gd, orc, io, max_elements=1 << 20, batch_capacity=1 << 14)
    p.d.bulk_build(io.up(keys), io.up(vals), count=len(keys))
    assert p.o.bulk_build(keys, vals) == 0
    p.check_contents()
    # closed form: the value of key k is the last index i with keys[i] == k
    uniq, first_rev = np.unique(keys[::-1], return_index=True)
    expect = (len(keys) - 1 - first_rev).astype(np.uint32)
    p.check_lookup(uniq, False)
    assert np.array_equal(p.o.lookup(uniq), expect)
    # updates after a bulk build
    p.insert(keys[:5000], vals[:5000] + 7)
    p.erase(keys[5000:9000])
    p.check_contents()
    p.check_successor(_queries(rng, keys, 4000, 1 << 17))


def _sparse_validation_key(quotient):
    # FliX/paper_lsmu_sweep.cu:434-439
    q = np.asarray(quotient, dtype=np.uint32)
    with np.errstate(over="ignore"):
        suffix = (q * np.uint32(40503) + np.uint32(17)) & np.uint32(0xFFFF)
    return (q << np.uint32(16)) | suffix


def test_forced_unified_scenarios(gd, orc, io):
    """FliX/paper_lsmu_sweep.cu:441-589: the scenario data and expectations
    are restated from the reference driver (closed forms, no oracle)."""
    d = gd.GPULSMOpt(max_elements=1 << 20, batch_capacity=1 << 16,
                     level_zero_capacity=(1 << 16) * 16)
    expected = {}
    removed = []

    def validate():
        qk = np.array(list(expected.keys()) + removed, np.uint32)
        qv = np.array(list(expected.values()) + [NOT_FOUND] * len(removed),
                      np.uint32)
        extra = []
        s = 0
        while len(extra) < 4096:
            k = (25000 << 16) | (s & 0xFFFF)
            if k not in expected:
                extra.append(k)
            s += 1
        qk = np.concatenate([qk, np.array(extra, np.uint32)])
        qv = np.concatenate([qv, np.full(len(extra), NOT_FOUND, np.uint32)])
        out = io.empty(len(qk))
        d.lookup(io.up(qk), out, count=len(qk))
        got = io.down(out, len(qk))
        assert int((got != qv).sum()) == 0

    for b in range(16):  # sparse_unique
        q = np.arange(b * 2048, (b + 1) * 2048, dtype=np.uint32)
        k = _sparse_validation_key(q)
        v = (0x01000000 + q).astype(np.uint32)
        expected.update(zip(k.tolist(), v.tolist()))
        d.insert(io.up(k), io.up(v), count=len(k))
    validate()
    for b in range(16):  # dense_unique
        sfx = np.arange(b * 4096, (b + 1) * 4096, dtype=np.uint32)
        k = (np.uint32(30000) << np.uint32(16)) | sfx
        v = (0x02000000 + sfx).astype(np.uint32)
        expected.update(zip(k.tolist(), v.tolist()))
        d.insert(io.up(k), io.up(v), count=len(k))
    validate()
    for b in range(16):  # duplicate_newest_wins
        i = np.arange(4096, dtype=np.uint32)
        k = _sparse_validation_key(i % 1024)
        v = (0x10000000 + b * 4096 + i).astype(np.uint32)
        for kk, vv in zip(k.tolist(), v.tolist()):
            expected[kk] = vv
        d.insert(io.up(k), io.up(v), count=len(k))
    validate()
    k = _sparse_validation_key(np.arange(1024, dtype=np.uint32))
    for kk in k.tolist():  # tombstone_visibility
        expected.pop(kk, None)
        removed.append(kk)
    kd = io.up(k)
    for _ in range(16):
        d.erase(kd, count=len(k))
    validate()


def test_multilevel_rollover(gd, io):
    """FliX/paper_lsmu_sweep.cu:591-641: max=512, batch=4, L0=64, 512
    batches of keys ordinal % 256.  Expected value of key k is
    0x28000000 + (last ordinal with ordinal % 256 == k) = 0x28000000+1792+k."""
    d = gd.GPULSMOpt(max_elements=512, batch_capacity=4,
                     level_zero_capacity=64)
    ordinals = np.arange(512 * 4, dtype=np.uint32)
    keys = (ordinals % 256).astype(np.uint32)
    vals = (0x28000000 + ordinals).astype(np.uint32)
    kd, vd = io.up(keys), io.up(vals)
    for b in range(512):
        d.insert(kd[4 * b:4 * b + 4], vd[4 * b:4 * b + 4], count=4)
    q = np.arange(256, dtype=np.uint32)
    out = io.empty(256)
    d.lookup(io.up(q), out, count=256)
    expect = (0x28000000 + 1792 + q).astype(np.uint32)
    assert np.array_equal(io.down(out, 256), expect)
    assert d.canonical_carry_status() == 0


def test_capacity_exhaustion(gd, orc, io):
    """FliX/paper_lsmu_sweep.cu:643-695: 17 epochs of 16 batches of 4 unique
    keys into a 1024-key dictionary fail with kPublicationOutputOverflow,
    and the next update throws an error that mentions the capacity."""
    d = gd.GPULSMOpt(max_elements=1024, batch_capacity=4,
                     level_zero_capacity=1024)
    o = orc.LsmOracle(max_elements=1024, batch_capacity=4,
                      level_zero_capacity=1024)
    keys = np.arange(17 * 16 * 4, dtype=np.uint32)
    vals = (0x30000000 + keys).astype(np.uint32)
    kd, vd = io.up(keys), io.up(vals)
    for b in range(17 * 16):
        d.insert(kd[4 * b:4 * b + 4], vd[4 * b:4 * b + 4], count=4)
        assert o.insert(keys[4 * b:4 * b + 4], vals[4 * b:4 * b + 4]) == 0
    status = d.canonical_carry_status()
    assert status & gd.PUBLICATION_OUTPUT_OVERFLOW
    assert status == o.status()
    with pytest.raises(gd.PublicationFailed) as err:
        d.insert(kd[:4], vd[:4], count=4)
    assert "capacity" in str(err.value)
    assert o.insert(keys[:4], vals[:4]) != 0
    # accepted updates stay readable
    out = io.empty(len(keys))
    d.lookup(kd, out, count=len(keys))
    assert np.array_equal(io.down(out, len(keys)), vals)
    # clear() recovers
    d.clear()
    d.insert(kd[:4], vd[:4], count=4)


def test_tiled_update_reports_failure(gd, orc, io):
    """A multi-tile update that fails mid-way throws after keeping the
    accepted tiles (GPULSMOpt.cuh:6913-6922)."""
    cfg = dict(max_elements=64, batch_capacity=4)
    d = gd.GPULSMOpt(**cfg)
    o = orc.LsmOracle(**cfg)
    keys = np.arange(200, dtype=np.uint32)
    vals = keys + 1
    with pytest.raises(gd.PublicationFailed):
        d.insert(io.up(keys), io.up(vals), count=200)
    assert o.insert(keys, vals) != 0
    q = np.arange(256, dtype=np.uint32)
    out = io.empty(256)
    d.lookup(io.up(q), out, count=256)
    assert np.array_equal(io.down(out, 256), o.lookup(q))


def test_paper_sweep_closed_form(gd, io):
    """FliX/paper_lsmu_sweep.cu:64-121,1137-1153: keys key_for_index(i) are
    unique (odd multiplier, bijection on 2^30); values are 1, so hits return
    1, misses (indices >= 2^27) return not_found, and range sums equal the
    number of keys in range."""
    import importlib
    wl = importlib.import_module("gdict_amd.workload")
    n = 1 << 18
    keys = wl.paper_key_for_index(np.arange(n, dtype=np.uint32))
    ones = np.ones(n, np.uint32)
    d = gd.GPULSMOpt(max_elements=n, batch_capacity=1 << 14)
    kd, vd = io.up(keys), io.up(ones)
    for b in range(n // (1 << 14)):
        s = slice(b << 14, (b + 1) << 14)
        d.insert(kd[s], vd[s], count=1 << 14)
    rng = np.random.default_rng(5)
    hits = keys[rng.integers(0, n, 1 << 16)]
    misses = wl.paper_key_for_index(
        rng.integers(1 << 27, 1 << 30, 1 << 16, dtype=np.uint64)
        .astype(np.uint32))
    out = io.empty(1 << 16)
    d.lookup(io.up(hits), out, count=1 << 16)
    assert np.all(io.down(out, 1 << 16) == 1)
    d.lookup(io.up(misses), out, count=1 << 16)
    assert np.all(io.down(out, 1 << 16) == NOT_FOUND)
    lo = rng.integers(0, 1 << 30, 4096, dtype=np.uint64).astype(np.uint32)
    hi = np.minimum(lo.astype(np.uint64) + (1 << 14), (1 << 30) - 1) \
        .astype(np.uint32)
    s = io.empty(4096)
    d.range(io.up(lo), io.up(hi), s, count=4096)
    sk = np.sort(keys)
    expect = (np.searchsorted(sk, hi, "right") -
              np.searchsorted(sk, lo, "left")).astype(np.uint32)
    assert np.array_equal(io.down(s, 4096), expect)


def test_config1_plain_lsm(gd, orc, io):
    """Config 1 at full size: 2^20 keys in 16 batches of 2^16, 2^20 lookups
    (50% present), 2^14 range counts with widths in [2^10, 2^16]."""
    import importlib
    wl = importlib.import_module("gdict_amd.workload")
    w = wl.config1()
    p = Pair(gd, orc, io, max_elements=w.n_keys, batch_capacity=w.batch)
    batches = list(w.batches())
    for keys, vals, _ in batches:
        p.insert(keys, vals)
    p.check_contents()
    q = w.lookups(w.live_keys(batches))
    p.check_lookup(q, False)
    p.check_lookup(q[:1 << 16], True)
    lo, hi = w.ranges()
    p.check_ranges(lo, hi)
    p.check_successor(q[:1 << 16])


def test_config2_scaled(gd, orc, io):
    """Config 2 shape (mixed batches with 10% tombstones of earlier keys,
    70% hit lookups, widths in [2^4, 2^20]) at 1/64 of its size."""
    import importlib
    wl = importlib.import_module("gdict_amd.workload")
    w = wl.config2(scale_log=6)
    p = Pair(gd, orc, io, max_elements=w.n_keys, batch_capacity=w.batch)
    batches = list(w.batches())
    for t, (keys, vals, erase) in enumerate(batches):
        p.mixed(keys, vals, erase)
        if t % 16 == 15:
            p.check_contents()
    q = w.lookups(w.live_keys(batches))
    p.check_lookup(q, True)
    lo, hi = w.ranges()
    p.check_ranges(lo, hi)
    p.check_successor(q[:1 << 14])


def test_config2_quarter(gd, orc, io):
    """Config 2 at 1/4 of its size: 2^24 keys in 64 mixed batches of 2^18
    (10% tombstones of earlier keys), full sorted contents after every 16
    batches, then all 2^24 lookups (70% live hits, found flags), all 2^18
    range counts and sums, and 2^20 successors against the oracle."""
    import importlib
    wl = importlib.import_module("gdict_amd.workload")
    w = wl.config2(scale_log=2)
    p = Pair(gd, orc, io, max_elements=w.n_keys, batch_capacity=w.batch)
    batches = list(w.batches())
    for t, (keys, vals, erase) in enumerate(batches):
        p.mixed(keys, vals, erase)
        if t % 16 == 15:
            p.check_contents()
    live = w.live_keys(batches)
    assert p.o.size() == len(live)
    q = w.lookups(live)
    p.check_lookup(q, True)
    lo, hi = w.ranges()
    p.check_ranges(lo, hi)
    p.check_successor(q[:1 << 20])


@pytest.mark.parametrize("knobs", [
    {"dir_rows_log": 1, "merge_tile_log": 11, "sort_block": 256},
    {"dir_rows_log": 5, "grouped_lookup_min_log": 8, "small_batch_max": 1},
    {"max_runs": 2, "small_batch_max": 1024, "sort_block": 512},
    {"max_runs": 40, "grouped_lookup_min_log": 63, "merge_tile_log": 12},
    {"segmented_admission": 1, "dir_rows_log": 2, "small_batch_max": 1024},
    {"segmented_admission": 2, "dir_rows_log": 4, "small_batch_max": 1024},
    {"merge_block": 512, "merge_tile_log": 12, "max_runs": 3},
    {"merge_block": 512, "merge_tile_log": 11, "dir_rows_log": 2},
    {"merge_ways": 2, "dir_rows_log": 3},
    {"merge_ways": 3, "dir_rows_log": 1, "max_runs": 4},
    {"merge_ways": 8, "dir_rows_log": 6, "small_batch_max": 1},
    {"sort_block": 1024, "segmented_admission": 2, "max_runs": 5},
    {"merge_fanout": 4, "merge_ways": 8},
    {"merge_fanout": 3, "merge_ways": 3, "max_runs": 6},
])
def test_tuning_knobs_keep_results(gd, orc, io, knobs):
    """Launch-shape knobs change speed only: a mixed stream with small and
    large batches gives the oracle's contents and answers under each."""
    rng = np.random.default_rng(77)
    p = Pair(gd, orc, io, tuning=knobs, max_elements=1 << 19,
             batch_capacity=1 << 13)
    got = p.d.get_tuning()
    for k, v in knobs.items():
        assert got[k] == v
    for b in range(24):
        n = int(rng.integers(1, 1 << 15)) if b % 3 else int(
            rng.integers(1, 3000))
        keys = rng.integers(0, 1 << 22, n, dtype=np.uint64).astype(np.uint32)
        vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        erase = (rng.random(n) < 0.2).astype(np.uint8)
        p.mixed(keys, vals, erase)
    p.check_contents()
    live = p.o.dump()[0]
    q = _queries(rng, live, 1 << 14, 1 << 22)
    p.check_lookup(q, True)
    lo = rng.integers(0, 1 << 22, 2000, dtype=np.uint64)
    hi = np.minimum(lo + rng.integers(0, 1 << 16, 2000, dtype=np.uint64),
                    0xFFFFFFFF)
    p.check_ranges(lo.astype(np.uint32), hi.astype(np.uint32))
    p.check_successor(q[:4096])


def test_invalid_tuning_rejected(gd):
    for bad in ({"sort_block": 384}, {"merge_tile_log": 13},
                {"segmented_admission": 3}, {"merge_block": 384},
                {"merge_ways": 1}, {"merge_ways": 9}, {"merge_fanout": 1},
                {"merge_fanout": 9}, {"no_such_knob": 1}):
        with pytest.raises(gd.InvalidArgument):
            gd.GPULSMOpt(max_elements=1 << 16, batch_capacity=1 << 10,
                         tuning=bad)


@pytest.mark.parametrize("ways", [2, 8])
def test_kernel_profile_counts(gd, orc, io, ways):
    """The opt-in kernel profile: merge rows in/out agree with the run
    sizes it merged and lookups count every query; reset clears it."""
    p = Pair(gd, orc, io, max_elements=1 << 18, batch_capacity=1 << 12,
             tuning={"merge_ways": ways})
    p.d.set_profiling(True)
    rng = np.random.default_rng(5)
    # distinct keys (odd multiplier: a bijection mod 2^32), so merges keep
    # every row
    distinct = ((np.arange(8 * 4096, dtype=np.uint64) * 2654435761) %
                (1 << 32)).astype(np.uint32)
    rng.shuffle(distinct)
    for b in range(8):
        keys = distinct[b * 4096:(b + 1) * 4096]
        p.insert(keys, keys)
    q = _queries(rng, p.o.dump()[0], 5000, 1 << 32)
    p.check_lookup(q, False)
    p.check_lookup(q[:3000], True)
    lo = np.sort(rng.integers(0, 1 << 32, 700, dtype=np.uint64)).astype(
        np.uint32)
    hi = (lo.astype(np.uint64) + (1 << 24)).clip(max=(1 << 32) - 1).astype(
        np.uint32)
    p.check_ranges(lo, hi)  # a sum launch and a count launch
    runs, _ = p.d.shape()
    prof = p.d.read_profile(reset=True)
    assert prof["range_launches"] == 2
    assert prof["range_queries"] == 1400
    assert prof["range_runs"] == 2 * runs
    assert prof["range_ms"] > 0
    if ways == 2:
        # eight equal batches under the binary counter: 4+2+1 merges
        assert prof["merge_launches"] == 7
        assert prof["merge_rows_in"] == 4096 * (2 * 4 + 4 * 2 + 8 * 1)
    else:
        # each cascade in one pass: batches 2 and 6 merge 2 runs, batch 4
        # merges 3 (1+1+2 batches), batch 8 merges 4 (1+1+2+4)
        assert prof["merge_launches"] == 4
        assert prof["merge_rows_in"] == 4096 * (2 + 4 + 2 + 8)
    assert prof["merge_rows_out"] == prof["merge_rows_in"]  # distinct keys
    assert prof["merge_ms"] > 0
    assert prof["admission_batches"] == 8
    assert prof["admission_keys"] == 8 * 4096
    assert prof["admission_ms"] > 0
    assert prof["lookup_launches"] == 2
    assert prof["lookup_queries"] == 8000
    assert prof["lookup_ms"] > 0
    again = p.d.read_profile(reset=False)
    assert again["merge_launches"] == 0 and again["lookup_launches"] == 0
    assert again["range_launches"] == 0
    p.d.set_profiling(False)


def test_snapshot_save_restore(gd, orc, io, tmp_path):
    """A snapshot holds the live contents (tombstones resolved); restoring
    it into a fresh dictionary gives the same contents and answers.
    Foreign, truncated, corrupted or other-width files are rejected."""
    rng = np.random.default_rng(31)
    p = Pair(gd, orc, io, max_elements=1 << 18, batch_capacity=1 << 12)
    for b in range(12):
        keys = rng.integers(0, 1 << 32, 6000, dtype=np.uint64).astype(
            np.uint32)
        erase = (rng.random(6000) < 0.25).astype(np.uint8)
        p.mixed(keys, keys ^ np.uint32(0xA5A5A5A5), erase)
    path = str(tmp_path / "lsm.snap")
    n = p.d.save(path)
    assert n == p.o.size()
    fresh = Pair(gd, orc, io, max_elements=1 << 18, batch_capacity=1 << 12)
    assert fresh.d.restore(path) == n
    k, v = p.o.dump()
    assert fresh.o.bulk_build(k, v) == 0
    fresh.check_contents()
    q = _queries(rng, k, 5000, 1 << 32)
    fresh.check_lookup(q, True)
    fresh.check_successor(q[:2000])
    # updates after a restore
    fresh.insert(q[:3000], q[:3000])
    fresh.check_contents()
    # empty dictionary round trip
    empty = gd.GPULSMOpt(max_elements=1 << 12, batch_capacity=64)
    epath = str(tmp_path / "empty.snap")
    assert empty.save(epath) == 0
    assert empty.restore(epath) == 0
    assert empty.dump_sorted() == 0
    # rejections
    raw = open(path, "rb").read()
    bad = tmp_path / "bad.snap"
    for blob in (b"not a snapshot at all", raw[: len(raw) - 5],
                 raw[:100] + bytes([raw[100] ^ 1]) + raw[101:]):
        bad.write_bytes(blob)
        with pytest.raises(gd.InvalidArgument):
            fresh.d.restore(str(bad))
    wide = gd.GPULSMOpt(max_elements=1 << 18, batch_capacity=1 << 12,
                        key_bytes=8)
    with pytest.raises(gd.InvalidArgument):
        wide.restore(path)
    small = gd.GPULSMOpt(max_elements=1 << 10, batch_capacity=16)
    with pytest.raises(gd.OutOfMemory):
        small.restore(path)


def test_snapshot_write_failures(gd, orc, io, tmp_path):
    """A snapshot that cannot be written (no such directory, or a device
    that is always full, where the final flush fails) raises
    InvalidArgument; the merged copy it made goes back to the pool, so
    repeated failures do not grow the resident bytes."""
    rng = np.random.default_rng(41)
    p = Pair(gd, orc, io, max_elements=1 << 16, batch_capacity=1 << 12)
    for _ in range(5):
        keys = rng.integers(0, 1 << 20, 3000, dtype=np.uint64).astype(
            np.uint32)
        p.insert(keys, keys)
    bad = str(tmp_path / "missing" / "x.snap")
    with pytest.raises(gd.InvalidArgument):
        p.d.save(bad)
    held = p.d.gpu_resident_bytes()
    for _ in range(3):
        with pytest.raises(gd.InvalidArgument):
            p.d.save(bad)
    assert p.d.gpu_resident_bytes() == held
    import os
    if os.path.exists("/dev/full"):
        with pytest.raises(gd.InvalidArgument):
            p.d.save("/dev/full")
        assert p.d.gpu_resident_bytes() == held
    p.check_contents()  # the dictionary is untouched


@pytest.mark.parametrize("tombstone_rate", [0.0, 0.3, 1.0])
def test_compact(gd, orc, io, tombstone_rate):
    """compact() (LSMu's cleanup, FliX/impl_lsm_tree.cuh:1255-1262) leaves
    one run holding exactly the live keys (tombstones and shadowed versions
    gone), with the same answers before and after, and updates after it
    behave as before."""
    rng = np.random.default_rng(int(tombstone_rate * 10) + 50)
    p = Pair(gd, orc, io, max_elements=1 << 20, batch_capacity=1 << 14)
    for b in range(23):
        n = int(rng.integers(100, 1 << 14))
        keys = rng.integers(0, 1 << 18, n, dtype=np.uint64).astype(np.uint32)
        erase = (rng.random(n) < tombstone_rate).astype(np.uint8)
        p.mixed(keys, keys * np.uint32(7), erase)
    runs_before, rows_before = p.d.shape()
    assert runs_before >= 1
    live = p.o.size()
    if tombstone_rate > 0:
        assert rows_before > live  # tombstones / shadowed versions held
    q = _queries(rng, p.o.dump()[0], 20000, 1 << 19)
    lo = rng.integers(0, 1 << 18, 3000, dtype=np.uint64).astype(np.uint32)
    hi = (lo + rng.integers(0, 1 << 12, 3000, dtype=np.uint64).astype(
        np.uint32))
    p.check_contents()
    p.d.compact()
    assert p.d.shape() == (1, live)
    p.check_contents()
    p.check_lookup(q, True)
    p.check_lookup(q, False)
    p.check_ranges(lo, hi)
    p.check_successor(q[:5000])
    p.d.compact()  # idempotent
    assert p.d.shape() == (1, live)
    for b in range(5):
        keys = rng.integers(0, 1 << 18, 5000, dtype=np.uint64).astype(
            np.uint32)
        erase = (rng.random(5000) < 0.2).astype(np.uint8)
        p.mixed(keys, keys + np.uint32(b), erase)
    p.check_contents()
    p.check_ranges(lo, hi)


def test_compact_empty(gd, io):
    d = gd.GPULSMOpt(max_elements=1 << 12, batch_capacity=64)
    d.compact()
    assert d.dump_sorted() == 0
    k = io.up(np.arange(10, dtype=np.uint32))
    d.insert(k, k)
    d.erase(k)
    d.compact()
    assert d.dump_sorted() == 0
    assert d.shape() == (1, 0)


@pytest.mark.parametrize("seed", [5, 6])
def test_ranges_over_many_runs(gd, orc, io, seed):
    """Range counts and sums with up to 12 runs holding shadowed versions
    and tombstones: narrow ranges (the round's newer rows are staged in
    LDS), ranges over the whole key space (staged rows beyond the LDS
    budget: one query at a time), empty and inverted ranges."""
    rng = np.random.default_rng(seed)
    p = Pair(gd, orc, io, tuning={"max_runs": 12, "small_batch_max": 1024},
             max_elements=1 << 20, batch_capacity=1 << 15)
    span = 1 << 24
    for b in range(11):
        # each batch smaller than the run before it: no merges
        n = 1 << (15 - b)
        keys = rng.integers(0, span, n, dtype=np.uint64).astype(np.uint32)
        vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        if b % 3 == 2:
            p.mixed(keys, vals, (rng.random(n) < 0.4).astype(np.uint8))
        else:
            p.insert(keys, vals)
    assert p.d.shape()[0] >= 8
    lo = rng.integers(0, span, 6000, dtype=np.uint64)
    width = np.where(rng.random(6000) < 0.5,
                     rng.integers(0, 1 << 12, 6000, dtype=np.uint64),
                     rng.integers(0, span, 6000, dtype=np.uint64))
    hi = np.minimum(lo + width, 0xFFFFFFFF)
    lo[:50], hi[:50] = 0, 0xFFFFFFFF      # whole key space
    lo[50:60], hi[50:60] = 1000, 10       # inverted: empty
    p.check_ranges(lo.astype(np.uint32), hi.astype(np.uint32))


def test_large_range_batch_compacts_first(gd, orc, io):
    """2^15 ranges over a stack whose newer runs hold more than 1/8 of the
    oldest run's rows: the range call merges the stack into one run first
    (tombstones drop); sums, counts, lookups and the contents stay the
    oracle's, and the stack is one run afterwards."""
    rng = np.random.default_rng(808)
    p = Pair(gd, orc, io, tuning={"max_runs": 12},
             max_elements=1 << 21, batch_capacity=1 << 16)
    span = 1 << 26
    for b in range(7):
        n = 1 << (16 - b)
        keys = rng.integers(0, span, n, dtype=np.uint64).astype(np.uint32)
        vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        if b % 2 == 1:
            p.mixed(keys, vals, (rng.random(n) < 0.3).astype(np.uint8))
        else:
            p.insert(keys, vals)
    assert p.d.shape()[0] >= 5
    n = 1 << 15
    lo = rng.integers(0, span, n, dtype=np.uint64)
    hi = np.minimum(lo + rng.integers(0, 1 << 16, n, dtype=np.uint64),
                    0xFFFFFFFF)
    lo[:8], hi[:8] = 0, 0xFFFFFFFF
    lo[8:12], hi[8:12] = 1000, 10
    p.check_ranges(lo.astype(np.uint32), hi.astype(np.uint32))
    assert p.d.shape()[0] == 1
    p.check_lookup(_queries(rng, p.o.dump()[0], 20000, span), True)
    p.check_contents()


@pytest.mark.parametrize("hot_keys", [3, 4000, 20000])
def test_segmented_admission_oversized_segments(gd, orc, io, hot_keys):
    """Segments beyond the LDS tile: 90% of each batch falls into one
    segment, drawn from `hot_keys` distinct keys (3 and 4000: deduplicated
    in LDS first; 20000: more distinct keys than the LDS table takes, so
    the segment is sorted in global memory), in insert, erase and mixed
    batches, with erase flags carried as positions or in the free top key
    bit (4000)."""
    rng = np.random.default_rng(hot_keys)
    p = Pair(gd, orc, io, tuning={"segmented_admission": 1},
             max_elements=1 << 20, batch_capacity=1 << 17)
    hot = rng.integers(0, 1 << 20, hot_keys, dtype=np.uint64) \
        .astype(np.uint32)
    for b in range(6):
        n = int(rng.integers(1 << 16, 1 << 17))
        # keys below 2^31 leave the top key bit free for the erase flag
        span = 1 << 31 if hot_keys == 4000 else 1 << 32
        keys = rng.integers(0, span, n, dtype=np.uint64).astype(np.uint32)
        sel = rng.random(n) < 0.9
        keys[sel] = hot[rng.integers(0, hot_keys, int(sel.sum()))]
        vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        kind = b % 3
        if kind == 0:
            p.insert(keys, vals)
        elif kind == 1:
            p.erase(keys[: n // 3])
        else:
            p.mixed(keys, vals, (rng.random(n) < 0.3).astype(np.uint8))
        p.check_contents()
    q = _queries(rng, p.o.dump()[0], 1 << 14, 1 << 32)
    p.check_lookup(q, True)


@pytest.mark.parametrize("span,dup,base", [
    (1 << 12, False, 0), (1 << 20, False, 0), (1 << 32, False, 0),
    (1 << 32, True, 0), (1 << 31, True, 0), (1 << 20, False, 0xFFF00000),
    (1 << 31, True, 1 << 31), (1 << 8, True, 0x9000_0000)])
def test_segmented_admission(gd, orc, io, span, dup, base):
    """The partition + per-segment LDS sort admission (32-bit batches above
    16384 keys) against the oracle: narrow key spreads (few varying bits),
    the full key space (erase flags travel as positions), keys that all
    have the top bit set (the flag rides in that bit during the sort), and
    hot keys whose segment exceeds the LDS tile (sorted in global memory by
    its workgroup)."""
    rng = np.random.default_rng(span % 1000 + dup + base % 997)
    p = Pair(gd, orc, io, tuning={"segmented_admission": 1},
             max_elements=1 << 21, batch_capacity=1 << 17)
    assert p.d.get_tuning()["segmented_admission"] == 1

    def draw(size):
        return (base + rng.integers(0, span, size, dtype=np.uint64)) \
            .astype(np.uint32)

    for b in range(10):
        n = int(rng.integers(16385, 1 << 17))
        keys = draw(n)
        if dup:  # ~40% of the batch on a handful of hot keys
            hot = draw(5)
            sel = rng.random(n) < 0.4
            keys[sel] = hot[rng.integers(0, 5, int(sel.sum()))]
        vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        kind = b % 3
        if kind == 0:
            p.insert(keys, vals)
        elif kind == 1:
            live = p.o.dump()[0]
            p.erase(np.concatenate([live[rng.integers(0, len(live), n // 2)],
                                    keys[: n - n // 2]]))
        else:
            erase = (rng.random(n) < 0.2).astype(np.uint8)
            p.mixed(keys, vals, erase)
        if b % 3 == 2 or b == 9:
            p.check_contents()
            q = _queries(rng, p.o.dump()[0], 1 << 14, span) + np.uint32(base)
            q[: 1 << 13] -= np.uint32(base)  # the live keys drawn above
            p.check_lookup(q, True)
            lo = base + rng.integers(0, span, 1000, dtype=np.uint64)
            hi = np.minimum(lo + rng.integers(0, max(span // 256, 2), 1000,
                                              dtype=np.uint64), 0xFFFFFFFF)
            p.check_ranges(lo.astype(np.uint32), hi.astype(np.uint32))
            p.check_successor(q[:4096])


def test_operations_on_alternating_streams(gd, orc, io):
    """Consecutive calls of one handle on different streams (and runs of
    calls on one stream) stay ordered: every batch is enqueued without a
    host wait, queries read the state after all earlier updates, and the
    results match the oracle."""
    import torch
    rng = np.random.default_rng(8)
    d = gd.GPULSMOpt(max_elements=1 << 22, batch_capacity=1 << 16)
    o = orc.LsmOracle(max_elements=1 << 22, batch_capacity=1 << 16)
    streams = [torch.cuda.Stream(), torch.cuda.Stream(), None]
    keep = []  # device inputs stay alive until the streams finish
    for b in range(24):
        n = int(rng.integers(1000, 1 << 16))
        k = rng.integers(0, 1 << 24, n, dtype=np.uint64).astype(np.uint32)
        v = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        e = (rng.random(n) < 0.2).astype(np.uint8)
        dk, dv, de = io.up(k), io.up(v), io.up(e, np.uint8)
        torch.cuda.synchronize()  # the uploads are done before any stream
        keep += [dk, dv, de]
        s = streams[(b // 2) % 3]  # two calls per stream, then the next
        d.apply_mixed(dk, dv, de, count=n, stream=s)
        assert o.apply_mixed(k, v, e) == 0
    q = _queries(rng, o.dump()[0], 20000, 1 << 24)
    dq = io.up(q)
    out = io.empty(len(q))
    torch.cuda.synchronize()
    d.lookup(dq, out, None, count=len(q), stream=streams[1])
    streams[1].synchronize()
    assert np.array_equal(io.down(out, len(q)), o.lookup(q))
    d.close()
