"""GPU parity of the one-pass merges of run cascades (up to eight runs per
merge) against the oracle, bit for bit, for 32- and 64-bit keys.  Equal
batches drive the binary counter through cascades of 2..7 runs; the key
patterns cover uniform keys over both halves of the key space, dense keys
(directory buckets far above a merge tile, which the partition refines by
key), disjoint sequential runs, hot keys updated in every run, tombstones
that drop at the oldest run and erase-everything streams that leave empty
runs.  Each case is also checked with merge_ways = 2 (pairwise merges) and
with merge_fanout = 4 (four newest runs of one size merge) for the same
answers."""

import numpy as np
import pytest

from test_lsm_gpu import Pair, _queries
from test_lsm64_gpu import Pair64

pytestmark = pytest.mark.gpu

PATTERNS = ["uniform", "dense", "sequential", "hot", "erase_all"]


def _batch(rng, pattern, b, n, bits):
    top = 1 << bits
    erase = np.zeros(n, np.uint8)
    if pattern == "uniform":
        keys = rng.integers(0, top, n, dtype=np.uint64)
        erase = (rng.random(n) < 0.15).astype(np.uint8)
    elif pattern == "dense":
        # every batch draws from one window of 3n keys around the top bit
        base = (top >> 1) - n
        keys = base + rng.integers(0, 3 * n, n, dtype=np.uint64)
        erase = (rng.random(n) < 0.1).astype(np.uint8)
    elif pattern == "sequential":
        keys = np.arange(b * n, (b + 1) * n, dtype=np.uint64) + (top >> 2)
    elif pattern == "hot":
        hot = np.array([0, 5, top >> 1, (top >> 1) + 7, top - 1] +
                       list(range(1000, 1100)), np.uint64)
        keys = np.where(rng.random(n) < 0.6, hot[rng.integers(0, len(hot), n)],
                        rng.integers(0, top, n, dtype=np.uint64))
        erase = (rng.random(n) < 0.2).astype(np.uint8)
    else:  # erase_all: the second half erases the keys of the first
        half = b // 2 if b % 2 == 0 else (b - 1) // 2
        r = np.random.default_rng(1000 + half)
        keys = r.integers(0, top, n, dtype=np.uint64)
        erase = np.full(n, b % 2, np.uint8)
    vals = rng.integers(0, top, n, dtype=np.uint64)
    return keys, vals, erase


def _run(p, rng, pattern, batches, n, bits, dt):
    for b in range(batches):
        keys, vals, erase = _batch(rng, pattern, b, n, bits)
        p.mixed(keys.astype(dt), vals.astype(dt), erase)
        if b in (2, 6, 13, batches - 1):
            p.check_contents()
    live = p.o.dump()[0]
    return live


@pytest.mark.parametrize("ways,fanout", [(8, 2), (2, 2), (8, 4)])
@pytest.mark.parametrize("pattern", PATTERNS)
def test_cascades_32(gd, orc, io, pattern, ways, fanout):
    rng = np.random.default_rng(PATTERNS.index(pattern) + 11)
    n = 6000
    p = Pair(gd, orc, io, tuning={"merge_ways": ways, "merge_fanout": fanout},
             max_elements=1 << 20, batch_capacity=n)
    live = _run(p, rng, pattern, 32, n, 32, np.uint32)
    q = _queries(rng, live, 1 << 13, 1 << 32)
    p.check_lookup(q, True)
    lo = rng.integers(0, 1 << 32, 3000, dtype=np.uint64)
    hi = np.minimum(lo + rng.integers(0, 1 << 24, 3000, dtype=np.uint64),
                    0xFFFFFFFF)
    p.check_ranges(lo.astype(np.uint32), hi.astype(np.uint32))
    p.check_successor(q[:4096])


@pytest.mark.parametrize("pattern", ["uniform", "dense", "hot", "erase_all"])
def test_cascades_64(gd, orc, io, pattern):
    rng = np.random.default_rng(PATTERNS.index(pattern) + 21)
    n = 3000
    p = Pair64(gd, orc, io, max_elements=1 << 19, batch_capacity=n)
    live = _run(p, rng, pattern, 24, n, 64, np.uint64)
    q = np.concatenate([live[rng.integers(0, max(len(live), 1), 4000)]
                        if len(live) else np.zeros(0, np.uint64),
                        rng.integers(0, 1 << 64, 2000, dtype=np.uint64)])
    p.check_lookup(q, True)
    lo = rng.integers(0, 1 << 64, 2000, dtype=np.uint64)
    hi = lo + rng.integers(0, 1 << 56, 2000, dtype=np.uint64)
    hi = np.where(hi < lo, np.uint64(0xFFFFFFFFFFFFFFFF), hi)
    p.check_ranges(lo, hi)
    p.check_successor(q[:2048])


def test_large_cascade_tiles(gd, orc, io):
    """Batches of 2^16 keys: the 8-batch cascade spans hundreds of merge
    tiles, with dense and uniform keys mixed in one stream."""
    rng = np.random.default_rng(99)
    n = 1 << 16
    p = Pair(gd, orc, io, max_elements=1 << 21, batch_capacity=n)
    for b in range(16):
        pattern = "dense" if b % 3 == 0 else "uniform"
        keys, vals, erase = _batch(rng, pattern, b, n, 32)
        p.mixed(keys.astype(np.uint32), vals.astype(np.uint32), erase)
    p.check_contents()
    live = p.o.dump()[0]
    q = _queries(rng, live, 1 << 15, 1 << 32)
    p.check_lookup(q, False)
