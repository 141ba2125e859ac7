"""The serving engine on the request streams of configs 4 and 5 (reduced
sizes), each result checked twice: against the oracle replayed in step
order (updates of a step before its queries, GPULSMOpt.cuh:6852-6858
ordering), and against the same query run alone on a plain dictionary that
received the same updates in the same order (the result is the same alone
and in the mix)."""

import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPS = {"insert": 0, "erase": 1, "lookup": 2, "lookup_nf": 2, "rcount": 3,
       "rsum": 4}


def _serve_and_check(gd, orc, io, reqs, kb, cfg, srv_kw):
    """reqs: (kind, a, b) with kind in OPS.  Submits every request (inputs
    written first), waits for all, then replays in step order."""
    kt = np.uint32 if kb == 4 else np.uint64
    srv = gd.Server(**cfg, key_bytes=kb, **srv_kw)
    staged = []
    for kind, a, b in reqs:
        da = io.up(a, kt)
        db = io.up(b, kt) if b is not None else None
        out = io.empty(len(a), kb) if kind not in ("insert", "erase") else None
        found = io.empty(len(a), 1) if kind == "lookup" else None
        staged.append((kind, da, db, out, found, len(a)))
    io.torch.cuda.synchronize()
    tickets = [srv.submit(OPS[k], da, db, out, found, count=n)
               for k, da, db, out, found, n in staged]
    steps = [srv.wait(t) for t in tickets]
    stats = srv.stats()
    assert stats["requests"] == len(reqs)
    assert stats["steps"] < len(reqs)  # requests were coalesced
    srv.close()

    o = orc.LsmOracle(key_bytes=kb, **cfg)
    replica = gd.GPULSMOpt(key_bytes=kb, **cfg)
    by_step = {}
    for i, s in enumerate(steps):
        by_step.setdefault(s, []).append(i)  # ticket (= arrival) order
    checked = 0
    for s in sorted(by_step):
        for i in by_step[s]:
            kind, a, b = reqs[i]
            da = staged[i][1]
            if kind == "insert":
                assert o.insert(a, b) == 0
                replica.insert(da, staged[i][2], count=len(a))
            elif kind == "erase":
                assert o.erase(a) == 0
                replica.erase(da, count=len(a))
        for i in by_step[s]:
            kind, a, b = reqs[i]
            if kind in ("insert", "erase"):
                continue
            _, da, db, out, found, n = staged[i]
            got = io.down(out, n, kt)
            alone = io.empty(n, kb)
            if kind == "lookup":
                ev, ef = o.lookup(a, with_found=True)
                assert np.array_equal(got, ev)
                assert np.array_equal(io.down(found, n, np.uint8), ef)
                af = io.empty(n, 1)
                replica.lookup(da, alone, af, count=n)
                assert np.array_equal(io.down(af, n, np.uint8), ef)
            elif kind == "lookup_nf":
                assert np.array_equal(got, o.lookup(a))
                replica.lookup(da, alone, count=n)
            elif kind == "rcount":
                assert np.array_equal(got, o.range_count(a, b))
                replica.range_count(da, db, alone, count=n)
            else:
                assert np.array_equal(got, o.range_sum(a, b))
                replica.range(da, db, alone, count=n)
            assert np.array_equal(io.down(alone, n, kt), got)
            checked += 1
    return checked, stats


@pytest.mark.parametrize("n_req,scale_log,inflight", [(400, 4, 2),
                                                      (150, 1, 3)])
def test_config4_stream(gd, orc, io, n_req, scale_log, inflight):
    """Config 4's stream: sizes log-uniform in [32, 2^18 >> scale_log],
    40% insert / 10% erase / 40% lookup / 10% range count, uint32 keys
    Zipf(0.99) over 2^30 scrambled bijectively (hot keys repeat across and
    inside requests)."""
    wl = importlib.import_module("gdict_amd.workload")
    w = wl.config4(n_requests=n_req, scale_log=scale_log)
    reqs = []
    for i, (kind, a, b) in enumerate(w.requests()):
        if kind == "lookup":
            kind = "lookup" if i % 2 else "lookup_nf"
        elif kind == "range":
            kind = "rcount"
        reqs.append((kind, a, b))
    cfg = dict(max_elements=1 << 24, batch_capacity=1 << 16)
    checked, _ = _serve_and_check(gd, orc, io, reqs, 4, cfg,
                                  dict(max_inflight_steps=inflight))
    assert checked > n_req // 3


def test_config5_stream(gd, orc, io):
    """Config 5's stream on a uint64 index grown to 2^18 keys through the
    engine (the full config grows it to 2^31): sizes log-uniform in
    [1, 2^16 >> 4], 50% lookup / 30% insert / 10% erase / 10% range sum,
    keys Zipf(0.8) over the grown keys, under an HBM budget that splits
    steps."""
    wl = importlib.import_module("gdict_amd.workload")
    scale = 13
    universe = 1 << (31 - scale)
    grow = wl.splitmix64(np.arange(universe, dtype=np.uint64))
    reqs = []
    for part in np.array_split(grow, 8):
        reqs.append(("insert", part, wl.splitmix64(part)))
    w = wl.ServingStream64(wl.SEEDS[5], 300, universe, smax=1 << 12)
    for i, (kind, a, b) in enumerate(w.requests()):
        if kind == "lookup":
            kind = "lookup" if i % 2 else "lookup_nf"
        elif kind == "range":
            kind = "rsum"
        reqs.append((kind, a, b))
    cfg = dict(max_elements=1 << 20, batch_capacity=1 << 15)
    checked, stats = _serve_and_check(
        gd, orc, io, reqs, 8, cfg,
        dict(max_inflight_steps=3, hbm_budget_bytes=48 << 20))
    assert checked > 100
    assert stats["peak_staged_bytes"] <= 48 << 20
