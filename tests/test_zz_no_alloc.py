"""gd_lsm_lookup_noalloc (the allocation-free lookup) against gd_lsm_lookup
and the oracle, bit for bit, through the C ABI: two calls in a row on the
library's shared device scratch, the second smaller than the first (the
scratch keeps the first call's contents), and the handle's device pool
does not grow across them."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_QUERIES = 1 << 20


def _dictionary(gd, orc, io, seed, tuning=None):
    """Four 2^18-key batches (a mixed one with tombstones of earlier keys)
    on the GPU and in the oracle."""
    rng = np.random.default_rng(seed)
    cfg = dict(max_elements=1 << 22, batch_capacity=1 << 18)
    d = gd.GPULSMOpt(tuning=tuning, **cfg)
    o = orc.LsmOracle(**cfg)
    inserted = []
    for b in range(4):
        keys = rng.integers(0, 1 << 31, 1 << 18, dtype=np.uint64).astype(
            np.uint32)
        vals = rng.integers(0, 1 << 32, 1 << 18, dtype=np.uint64).astype(
            np.uint32)
        erase = np.zeros(len(keys), np.uint8)
        if inserted:
            prev = np.concatenate(inserted)
            pick = rng.integers(0, len(prev), len(keys) // 10)
            keys[: len(pick)] = prev[pick]
            erase[: len(pick)] = 1
        d.apply_mixed(io.up(keys), io.up(vals), io.up(erase, np.uint8),
                      count=len(keys))
        assert o.apply_mixed(keys, vals, erase) == 0
        inserted.append(keys[erase == 0])
    return d, o, np.concatenate(inserted), rng


def _queries(rng, keys, n):
    q = rng.integers(0, 1 << 31, n, dtype=np.uint64).astype(np.uint32)
    take = rng.integers(0, len(keys), n // 2)
    q[: n // 2] = keys[take]
    return q


def _noalloc(gd, d, q_dev, out, found, n):
    """Calls the entry point through the C ABI directly."""
    import torch
    lib = gd.lib()
    st = lib.gd_lsm_lookup_noalloc(
        d._h, q_dev.data_ptr(), n, out.data_ptr(),
        found.data_ptr() if found is not None else None,
        torch.cuda.current_stream().cuda_stream)
    return st


@pytest.mark.parametrize("tuning", [None, {"grouped_lookup_min_log": 10}])
def test_noalloc_lookup_matches_lookup(gd, orc, io, tuning):
    d, o, keys, rng = _dictionary(gd, orc, io, 11, tuning)
    io.torch.cuda.synchronize()
    # a large call, then a smaller one (both on the grouped path by
    # default: 2^18 queries and up), then one on the direct path
    sizes = [MAX_QUERIES, 300_001, 4099]
    batches = [_queries(rng, keys, n) for n in sizes]
    devs = [io.up(q) for q in batches]
    outs = [(io.empty(n), io.empty(n, 1)) for n in sizes]
    resident_before = d.gpu_resident_bytes()
    for i, n in enumerate(sizes):
        found = outs[i][1] if i != 1 else None  # the second without found
        assert _noalloc(gd, d, devs[i], outs[i][0], found, n) == gd.GD_OK
    io.torch.cuda.synchronize()
    assert d.gpu_resident_bytes() == resident_before  # nothing was allocated
    for i, n in enumerate(sizes):
        q = batches[i]
        ref = io.empty(n)
        ref_found = io.empty(n, 1) if i != 1 else None
        d.lookup(devs[i], ref, ref_found, count=n)
        got = io.down(outs[i][0], n)
        assert np.array_equal(got, io.down(ref, n)), ("values", n)
        if i != 1:
            ev, ef = o.lookup(q, with_found=True)
            assert np.array_equal(io.down(outs[i][1], n, np.uint8), ef)
            assert np.array_equal(
                io.down(outs[i][1], n, np.uint8),
                io.down(ref_found, n, np.uint8))
            assert np.array_equal(got, ev)
        else:
            assert np.array_equal(got, o.lookup(q))
    d.close()


def test_noalloc_lookup_rejects_oversized_batches(gd, orc, io):
    d = gd.GPULSMOpt(max_elements=1 << 16, batch_capacity=1 << 10)
    n = MAX_QUERIES + 1
    q = io.empty(n)
    out = io.empty(n)
    assert _noalloc(gd, d, q, out, None, n) == gd.GD_INVALID_ARGUMENT
    msg = gd.lib().gd_last_error_message(d._h).decode()
    assert "at most" in msg
    with pytest.raises(gd.InvalidArgument):
        d.lookup_noalloc(q, out)
    # an empty dictionary answers all misses without scratch
    k = io.up(np.arange(5000, dtype=np.uint32))
    out = io.empty(5000)
    d.lookup_noalloc(k, out)
    assert np.all(io.down(out, 5000) == 0xFFFFFFFF)
    d.close()


def test_noalloc_lookup_is_declared(gd):
    assert ctypes.CDLL(gd.LIB_PATH).gd_lsm_lookup_noalloc is not None
