"""CPU pins of the oracle (no GPU): every oracle function is checked against
something it does not share code with -- published known-answer vectors,
closed forms from the reference's own validation scenarios, or a naive
brute-force replay in plain Python on small inputs."""

import ctypes
import importlib

import numpy as np
import pytest

NF = 0xFFFFFFFF


# ------------------------------------------------------------- generators
def test_splitmix64_published_sequence(orc):
    # SplitMix64 seeded with 0 (Vigna's reference generator): first outputs
    lib = orc.lib()
    assert lib.orc_splitmix64(0) == 0xE220A8397B1DCDAF
    assert lib.orc_splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    assert lib.orc_splitmix64((2 * 0x9E3779B97F4A7C15) % 2**64) == \
        0x06C45D188009454F


@pytest.mark.parametrize("data,seed,expect", [
    (b"", 0, 0), (b"", 1, 0x514E28B7), (b"", 0xFFFFFFFF, 0x81F16F39),
    (b"\0\0\0\0", 0, 0x2362F9DE), (b"aaaa", 0x9747B28C, 0x5A97808A),
    (b"abc", 0, 0xB3DD93FA), (b"Hello, world!", 0x9747B28C, 0x24884CBA),
    (b"The quick brown fox jumps over the lazy dog", 0x9747B28C,
     0x2FA826CD)])
def test_murmur3_published_vectors(orc, data, seed, expect):
    buf = ctypes.create_string_buffer(data, max(len(data), 1))
    assert orc.lib().orc_murmur3_bytes(ctypes.addressof(buf), len(data),
                                       seed) == expect


def test_murmur3_u32_is_the_4_byte_case(orc):
    rng = np.random.default_rng(1)
    for k in rng.integers(0, 2**32, 200, dtype=np.uint64):
        k = int(k)
        buf = ctypes.create_string_buffer(k.to_bytes(4, "little"), 4)
        assert orc.lib().orc_murmur3_u32(k, 0) == \
            orc.lib().orc_murmur3_bytes(ctypes.addressof(buf), 4, 0)


def test_product_generators_match_the_oracle(orc, gd):
    wl = importlib.import_module("gdict_amd.workload")
    seed = 0xC0F10002
    assert np.array_equal(wl.stream(seed, 5000),
                          orc.splitmix64_stream(seed, 5000))
    keys = (wl.stream(seed, 5000) >> np.uint64(33)).astype(np.uint32)
    assert np.array_equal(wl.murmur3_u32(keys), orc.murmur3(keys))
    a = wl.log_uniform(wl.stream(7, 20000), 16, 1 << 20)
    b = orc.log_uniform_stream(7, 20000, 16, 1 << 20)
    assert a.min() >= 16 and a.max() <= 1 << 20
    assert (a != b).mean() < 1e-3  # libm rounding may differ by one ulp


def test_log_uniform_bounds_and_median(orc):
    x = orc.log_uniform_stream(3, 100000, 1 << 4, 1 << 20)
    assert x.min() >= 16 and x.max() <= 1 << 20
    # median of a log-uniform on [2^4, 2^20] is 2^12
    assert abs(np.log2(np.median(x)) - 12) < 0.12  # 4.5 standard errors
    lib = orc.lib()
    assert lib.orc_log_uniform(0, 16, 1 << 20) == 16


def test_zipf_frequencies(orc):
    s, n = 1.1, 20
    r = orc.zipf_ranks(s, n, 11, 200000)
    assert r.min() >= 1 and r.max() <= n
    pmf = 1.0 / np.arange(1, n + 1) ** s
    pmf /= pmf.sum()
    freq = np.bincount(r.astype(np.int64), minlength=n + 1)[1:] / len(r)
    assert np.max(np.abs(freq - pmf)) < 0.01


def test_product_zipf_matches_oracle(orc, gd):
    """The product's numpy Zipf sampler (config 4/5 keys) draws the same
    ranks as the oracle's rejection-inversion, sample for sample."""
    import importlib
    wl = importlib.import_module("gdict_amd.workload")
    for s, n, seed in ((1.1, 20, 11), (0.99, 1 << 30, 5), (0.8, 1 << 31, 9)):
        assert np.array_equal(wl.Zipf(s, n).ranks(seed, 50000),
                              orc.zipf_ranks(s, n, seed, 50000))
    # the scrambling is a bijection of the 30-bit key space
    r = np.arange(1 << 16, dtype=np.uint64) * np.uint64(16411)
    k = wl.scrambled_keys(r, 30)
    assert len(np.unique(k)) == len(r) and k.max() < (1 << 30)


def test_paper_key_bijection(orc):
    # FliX/paper_lsmu_sweep.cu:64-67: odd multiplier -> bijection on 2^30
    wl = importlib.import_module("gdict_amd.workload")
    idx = np.arange(1 << 20, dtype=np.uint32)
    k = wl.paper_key_for_index(idx)
    assert len(np.unique(k)) == len(k)
    assert int(k.max()) < 1 << 30
    lib = orc.lib()
    for i in (0, 1, 12345, (1 << 30) - 1):
        assert lib.orc_paper_key_for_index(i) == int(k[i]) if i < len(k) \
            else lib.orc_paper_key_for_index(i) < 1 << 30


def test_range_digest_composes(orc):
    rng = np.random.default_rng(2)
    r = rng.integers(0, 2**32, 1000, dtype=np.uint64).astype(np.uint32)
    s_all, x_all = orc.range_digest(r, 0)
    s1, x1 = orc.range_digest(r[:400], 0)
    s2, x2 = orc.range_digest(r[400:], 400)
    assert s_all == s1 + s2 == int(r.astype(np.uint64).sum())
    assert x_all == x1 ^ x2


# ------------------------------------------------------------- LSM oracle
class DictReplay:
    """Naive replay: a dict, newest write wins, erase deletes."""

    def __init__(self):
        self.d = {}

    def apply(self, keys, vals, erase):
        for k, v, e in zip(keys.tolist(), vals.tolist(), erase.tolist()):
            if e:
                self.d.pop(k, None)
            else:
                self.d[k] = v

    def lookup(self, q):
        return np.array([self.d.get(k, NF) for k in q.tolist()], np.uint32)

    def range_sum(self, lo, hi, count=False):
        out = []
        for a, z in zip(lo.tolist(), hi.tolist()):
            s = sum((1 if count else v) for k, v in self.d.items()
                    if a <= k <= z)
            out.append(s & 0xFFFFFFFF)
        return np.array(out, np.uint32)

    def successor(self, q):
        ks = sorted(self.d)
        out = []
        for x in q.tolist():
            c = [k for k in ks if k >= x]
            out.append(c[0] if c else NF)
        return np.array(out, np.uint32)


@pytest.mark.parametrize("seed", range(4))
def test_lsm_oracle_against_dict_replay(orc, seed):
    rng = np.random.default_rng(100 + seed)
    span = [64, 1000, 1 << 20, 1 << 32][seed]
    o = orc.LsmOracle(max_elements=1 << 16, batch_capacity=7)
    ref = DictReplay()
    for b in range(30):
        n = int(rng.integers(1, 40))
        k = rng.integers(0, span, n, dtype=np.uint64).astype(np.uint32)
        v = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        e = (rng.random(n) < [0.0, 1.0, 0.3][b % 3]).astype(np.uint8)
        if b % 3 == 0:
            assert o.insert(k, v) == 0
        elif b % 3 == 1:
            assert o.erase(k) == 0
        else:
            assert o.apply_mixed(k, v, e) == 0
        ref.apply(k, np.where(e > 0, 0, v) if b % 3 else v,
                  np.ones(n, np.uint8) if b % 3 == 1 else
                  (e if b % 3 == 2 else np.zeros(n, np.uint8)))
    q = rng.integers(0, span, 300, dtype=np.uint64).astype(np.uint32)
    if ref.d:
        q[:100] = rng.choice(np.array(list(ref.d), np.uint32), 100)
    assert np.array_equal(o.lookup(q), ref.lookup(q))
    vals, found = o.lookup(q, with_found=True)
    assert np.array_equal(found, (ref.lookup(q) != NF) | np.isin(
        q, [k for k, v in ref.d.items() if v == NF]))
    lo = rng.integers(0, span, 60, dtype=np.uint64).astype(np.uint32)
    hi = (lo.astype(np.uint64) + rng.integers(0, span, 60, dtype=np.uint64)) \
        .clip(0, NF).astype(np.uint32)
    lo[:5], hi[:5] = hi[:5], lo[:5]
    assert np.array_equal(o.range_sum(lo, hi), ref.range_sum(lo, hi))
    assert np.array_equal(o.range_count(lo, hi),
                          ref.range_sum(lo, hi, count=True))
    assert np.array_equal(o.successor(q), ref.successor(q))
    k, v = o.dump()
    assert k.tolist() == sorted(ref.d)
    assert v.tolist() == [ref.d[x] for x in sorted(ref.d)]


ALL64 = (1 << 64) - 1


@pytest.mark.parametrize("seed", range(4))
def test_lsm64_oracle_against_dict_replay(orc, seed):
    """The 64-bit extension: dict replay with sums modulo 2^64, all-ones
    misses and successor sentinel, keys in both halves of the key space."""
    rng = np.random.default_rng(200 + seed)
    span = [64, 1 << 33, 1 << 63, 0][seed]
    o = orc.LsmOracle(max_elements=1 << 16, batch_capacity=5, key_bytes=8)
    d = {}

    def draw(n):
        if span:
            return rng.integers(0, span, n, dtype=np.uint64) | (
                np.uint64(1 << 63) if seed == 2 else np.uint64(0))
        return rng.integers(0, 2**64, n, dtype=np.uint64, endpoint=False)

    for b in range(30):
        n = int(rng.integers(1, 30))
        k = draw(n)
        if b == 0:
            k[:2] = [0, ALL64][:n] if n >= 2 else [ALL64]
        v = rng.integers(0, 2**64, n, dtype=np.uint64, endpoint=False)
        e = (rng.random(n) < 0.3).astype(np.uint8)
        if b % 3 == 0:
            assert o.insert(k, v) == 0
            for kk, vv in zip(k.tolist(), v.tolist()):
                d[kk] = vv
        elif b % 3 == 1:
            assert o.erase(k) == 0
            for kk in k.tolist():
                d.pop(kk, None)
        else:
            assert o.apply_mixed(k, v, e) == 0
            for kk, vv, ee in zip(k.tolist(), v.tolist(), e.tolist()):
                if ee:
                    d.pop(kk, None)
                else:
                    d[kk] = vv
    q = draw(200)
    if d:
        q[:80] = rng.choice(np.array(list(d), np.uint64), 80)
    q[80:82] = [0, ALL64]
    expect = np.array([d.get(x, ALL64) for x in q.tolist()], np.uint64)
    assert np.array_equal(o.lookup(q), expect)
    vals, found = o.lookup(q, with_found=True)
    assert found.tolist() == [int(x in d) for x in q.tolist()]
    assert vals.tolist() == [d.get(x, 0) for x in q.tolist()]
    lo = draw(50)
    w = rng.integers(0, 2**62, 50, dtype=np.uint64)
    hi = lo + np.minimum(w, np.uint64(ALL64) - lo)
    lo[:4], hi[:4] = hi[:4].copy(), lo[:4].copy()
    lo[4], hi[4] = 0, ALL64
    sums = [sum(v for k, v in d.items() if a <= k <= z) % (1 << 64)
            for a, z in zip(lo.tolist(), hi.tolist())]
    counts = [sum(1 for k in d if a <= k <= z)
              for a, z in zip(lo.tolist(), hi.tolist())]
    assert o.range_sum(lo, hi).tolist() == sums
    assert o.range_count(lo, hi).tolist() == counts
    ks = sorted(d)
    succ = [next((k for k in ks if k >= x), ALL64) for x in q.tolist()]
    assert o.successor(q).tolist() == succ
    k, v = o.dump()
    assert k.tolist() == ks
    assert v.tolist() == [d[x] for x in ks]


def test_lsm_oracle_bulk_build_last_wins(orc):
    keys = np.array([5, 3, 5, 9, 3, 5], np.uint32)
    vals = np.arange(6, dtype=np.uint32)
    o = orc.LsmOracle()
    assert o.bulk_build(keys, vals) == 0
    k, v = o.dump()
    assert k.tolist() == [3, 5, 9] and v.tolist() == [4, 5, 3]


def test_lsm_oracle_rollover_closed_form(orc):
    # FliX/paper_lsmu_sweep.cu:591-641
    o = orc.LsmOracle(max_elements=512, batch_capacity=4,
                      level_zero_capacity=64)
    for b in range(512):
        ordn = np.arange(4 * b, 4 * b + 4, dtype=np.uint32)
        assert o.insert(ordn % 256, 0x28000000 + ordn) == 0
    q = np.arange(256, dtype=np.uint32)
    assert np.array_equal(o.lookup(q), 0x28000000 + 1792 + q)
    assert o.status() == 0


def test_lsm_oracle_capacity_closed_form(orc):
    # FliX/paper_lsmu_sweep.cu:643-695: 1088 unique keys > 1024 capacity
    o = orc.LsmOracle(max_elements=1024, batch_capacity=4,
                      level_zero_capacity=1024)
    for b in range(17 * 16 - 1):
        k = np.arange(4 * b, 4 * b + 4, dtype=np.uint32)
        assert o.insert(k, k) == 0
        assert o.status() == 0  # 16 epochs = 1024 keys fit
    k = np.arange(4 * 271, 4 * 272, dtype=np.uint32)
    assert o.insert(k, k) == 0  # the failing publication is reported later
    assert o.status() == 8
    assert o.insert(k, k) != 0
    assert "capacity" in o.message()
    assert o.size() == 1088  # accepted updates are preserved


def test_planted_oracle_mistake_is_caught(orc):
    """A deliberately wrong replay (oldest write wins) must disagree with the
    oracle on the duplicate scenario, so the parity check has teeth."""
    o = orc.LsmOracle(max_elements=1 << 12, batch_capacity=16)
    keys = np.array([1, 2, 1, 1], np.uint32)
    vals = np.array([10, 20, 11, 12], np.uint32)
    o.insert(keys, vals)
    wrong = {}
    for k, v in zip(keys.tolist(), vals.tolist()):
        wrong.setdefault(k, v)
    q = np.array([1], np.uint32)
    assert o.lookup(q)[0] == 12 and wrong[1] == 10


# ------------------------------------------------------------ FliX oracle
def test_flix_oracle_interface_known_answer(orc):
    # FliX/main.cu:98-156
    o = orc.FlixOracle(4, 4)
    o.build(np.array([100, 90, 80, 70, 60, 50, 40, 30, 20, 10], np.uint32))
    probes = np.arange(5, 106, 5, dtype=np.uint32)
    exp = np.array([NF if p % 10 else (100 - p) // 10 for p in probes],
                   np.uint32)
    assert np.array_equal(o.lookup(probes), exp)
    o.insert(np.arange(5, 106, 10, dtype=np.uint32),
             np.arange(10, 21, dtype=np.uint32))
    exp = np.array([(100 - p) // 10 if p % 10 == 0 else 10 + (p - 5) // 10
                    for p in probes], np.uint32)
    assert np.array_equal(o.lookup(probes), exp)
    o.remove(np.array([30, 35, 40, 45, 75, 80], np.uint32))
    exp[np.isin(probes, [30, 35, 40, 45, 75, 80])] = NF
    assert np.array_equal(o.lookup(probes), exp)
    assert o.size() == 15


@pytest.mark.parametrize("kb,vb", [(4, 4), (8, 8)])
def test_flix_oracle_against_dict(orc, kb, vb):
    rng = np.random.default_rng(kb + vb)
    kt = np.uint32 if kb == 4 else np.uint64
    vt = np.uint32 if vb == 4 else np.uint64
    o = orc.FlixOracle(kb, vb)
    build = rng.integers(2, 500, 200).astype(kt)
    o.build(build)
    d = {}
    for i, k in enumerate(build.tolist()):
        d.setdefault(k, i)  # first occurrence wins
    for step in range(10):
        ins = rng.integers(2, 600, 50).astype(kt)
        vals = rng.integers(0, 1000, 50).astype(vt)
        o.insert(ins, vals)
        for k, v in zip(ins.tolist(), vals.tolist()):
            d.setdefault(k, v)  # insert-if-absent
        dels = rng.integers(2, 600, 30).astype(kt)
        o.remove(dels)
        for k in dels.tolist():
            d.pop(k, None)
    q = np.arange(0, 700, dtype=kt)
    nf = np.iinfo(vt).max
    assert np.array_equal(o.lookup(q), np.array([d.get(k, nf) for k in
                                                 q.tolist()], vt))
    ks = sorted(d)
    succ = [next((k for k in ks if k >= x), 0xFFFFFFFF) for x in q.tolist()]
    assert np.array_equal(o.successor(q), np.array(succ, kt))
    lo = np.array([0, 10, 300, 650, 50], kt)
    hi = np.array([700, 20, 310, 700, 40], kt)
    offs, rk, rv = o.range(lo, hi)
    for i in range(len(lo)):
        sel = [k for k in ks if lo[i] <= k <= hi[i]]
        assert rk[offs[i]:offs[i + 1]].tolist() == sel
        assert rv[offs[i]:offs[i + 1]].tolist() == [d[k] for k in sel]


# ------------------------------------------------------- CPU LSM baseline
def test_cpu_baseline_matches_oracle(orc):
    """The CPU baseline against the oracle; its ranges are checked after
    several batches, so both range paths run: the rank/prefix difference
    over a tombstone-free oldest run, and the row visit when the oldest run
    still holds tombstones (a single run made from a batch with erases)."""
    rng = np.random.default_rng(77)
    c = orc.CpuLsm(4)
    o = orc.LsmOracle(max_elements=1 << 22, batch_capacity=1 << 14)

    def check_ranges():
        lo = rng.integers(0, 1 << 18, 1500, dtype=np.uint64).astype(np.uint32)
        hi = lo + rng.integers(0, 9000, 1500, dtype=np.uint64).astype(
            np.uint32)
        lo[:4] = [0, 5, 0, 1 << 17]
        hi[:4] = [0xFFFFFFFF, 4, 0, 1 << 17]
        assert np.array_equal(c.range(lo, hi), o.range_sum(lo, hi))
        assert np.array_equal(c.range(lo, hi, True), o.range_count(lo, hi))

    for b in range(24):
        n = int(rng.integers(1, 20000))
        k = rng.integers(0, 1 << 18, n, dtype=np.uint64).astype(np.uint32)
        v = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        e = (rng.random(n) < 0.2).astype(np.uint8)
        c.apply(k, v, e)
        o.apply_mixed(k, v, e)
        if b in (0, 1, 2, 5, 11, 17):
            check_ranges()
    q = rng.integers(0, 1 << 18, 20000, dtype=np.uint64).astype(np.uint32)
    a, fa = c.lookup(q, True)
    b, fb = o.lookup(q, True)
    assert np.array_equal(a, b) and np.array_equal(fa, fb)
    lo = rng.integers(0, 1 << 18, 2000, dtype=np.uint64).astype(np.uint32)
    hi = lo + rng.integers(0, 5000, 2000, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(c.range(lo, hi), o.range_sum(lo, hi))
    assert np.array_equal(c.range(lo, hi, True), o.range_count(lo, hi))
