"""GPU tests of the serving engine (gd_serve_*): every request's result must
equal a sequential replay in step order (updates before queries inside a
step), both on the oracle and on a plain-API replica that runs each request
alone."""

import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NF = 0xFFFFFFFF


def _make_requests(rng, n_req, span, max_size):
    reqs = []
    inserted = []
    for i in range(n_req):
        size = int(np.exp(rng.uniform(0, np.log(max_size)))) or 1
        r = rng.random()
        if r < 0.4:
            k = rng.integers(0, span, size, dtype=np.uint64).astype(np.uint32)
            v = rng.integers(0, 2**32, size, dtype=np.uint64).astype(np.uint32)
            reqs.append(("insert", k, v))
            inserted.append(k)
        elif r < 0.5:
            pool = np.concatenate(inserted) if inserted else \
                np.zeros(1, np.uint32)
            k = pool[rng.integers(0, len(pool), size)]
            reqs.append(("erase", k, None))
        elif r < 0.85:
            k = rng.integers(0, span, size, dtype=np.uint64).astype(np.uint32)
            if inserted:
                pool = np.concatenate(inserted)
                k[: size // 2] = pool[rng.integers(0, len(pool), size // 2)]
            reqs.append(("lookup" if r < 0.7 else "lookup_nf", k, None))
        else:
            lo = rng.integers(0, span, size, dtype=np.uint64)
            hi = np.minimum(lo + rng.integers(0, span // 16, size,
                                              dtype=np.uint64), NF)
            reqs.append(("rcount" if r < 0.93 else "rsum",
                         lo.astype(np.uint32), hi.astype(np.uint32)))
    return reqs


OPS = {"insert": 0, "erase": 1, "lookup": 2, "lookup_nf": 2, "rcount": 3,
       "rsum": 4}


def _submit_all(gd, io, srv, reqs, order=None):
    handles = []
    staged = []
    for idx in (order if order is not None else range(len(reqs))):
        kind, a, b = reqs[idx]
        da = io.up(a)
        db = io.up(b) if b is not None else None
        out = found = None
        if kind in ("lookup", "lookup_nf", "rcount", "rsum"):
            out = io.empty(len(a))
        if kind == "lookup":
            found = io.empty(len(a), 1)
        staged.append((idx, kind, da, db, out, found, len(a)))
    # request buffers must be fully written before they are submitted: the
    # engine's streams do not wait for the caller's stream
    io.torch.cuda.synchronize()
    for idx, kind, da, db, out, found, n in staged:
        t = srv.submit(OPS[kind], da, db, out, found, count=n)
        handles.append((idx, t, out, found))
    return handles


def _check_against_replay(gd, orc, io, reqs, served, cfg):
    """served: list of (idx, ticket order, step, out, found)."""
    o = orc.LsmOracle(**cfg)
    replica = gd.GPULSMOpt(**cfg)
    by_step = {}
    for rec in served:
        by_step.setdefault(rec[2], []).append(rec)
    for step in sorted(by_step):
        recs = sorted(by_step[step], key=lambda r: r[1])  # arrival order
        for idx, _, _, _, _ in recs:
            kind, a, b = reqs[idx]
            if kind == "insert":
                assert o.insert(a, b) == 0
                replica.insert(io.up(a), io.up(b), count=len(a))
            elif kind == "erase":
                assert o.erase(a) == 0
                replica.erase(io.up(a), count=len(a))
        for idx, _, _, out, found in recs:
            kind, a, b = reqs[idx]
            if kind in ("insert", "erase"):
                continue
            got = io.down(out, len(a))
            alone = io.empty(len(a))
            if kind == "lookup":
                ev, ef = o.lookup(a, with_found=True)
                assert np.array_equal(got, ev)
                assert np.array_equal(io.down(found, len(a), np.uint8), ef)
                af = io.empty(len(a), 1)
                replica.lookup(io.up(a), alone, af, count=len(a))
                assert np.array_equal(io.down(alone, len(a)), got)
            elif kind == "lookup_nf":
                assert np.array_equal(got, o.lookup(a))
                replica.lookup(io.up(a), alone, count=len(a))
                assert np.array_equal(io.down(alone, len(a)), got)
            elif kind == "rcount":
                assert np.array_equal(got, o.range_count(a, b))
                replica.range_count(io.up(a), io.up(b), alone, count=len(a))
                assert np.array_equal(io.down(alone, len(a)), got)
            else:
                assert np.array_equal(got, o.range_sum(a, b))
                replica.range(io.up(a), io.up(b), alone, count=len(a))
                assert np.array_equal(io.down(alone, len(a)), got)


@pytest.mark.parametrize("seed,span,max_size,inflight,n_req,launch", [
    (1, 1 << 12, 300, 2, 300, "direct"), (2, 1 << 20, 5000, 3, 300, "direct"),
    (3, 1 << 32, 20000, 1, 300, "direct"),
    # large inserts: merges of several 2^16-row runs are still running on
    # the merge stream while later steps query their inputs
    (4, 1 << 32, 1 << 16, 3, 120, "direct"),
    # small steps as captured graphs, and every step kernel by kernel
    (5, 1 << 20, 5000, 3, 300, "graph"), (6, 1 << 20, 5000, 3, 300,
                                          "kernels")])
def test_served_results_equal_step_replay(gd, orc, io, seed, span, max_size,
                                          inflight, n_req, launch):
    rng = np.random.default_rng(seed)
    cfg = dict(max_elements=1 << 22, batch_capacity=1 << 16)
    srv = gd.Server(**cfg, max_inflight_steps=inflight, step_launch=launch)
    reqs = _make_requests(rng, n_req, span, max_size)
    handles = _submit_all(gd, io, srv, reqs)
    served = []
    for order, (idx, t, out, found) in enumerate(handles):
        served.append((idx, order, srv.wait(t), out, found))
    st = srv.stats()
    assert st["requests"] == len(reqs)
    assert st["steps"] < len(reqs)  # requests were coalesced
    # device memory as of the last completed step
    assert 0 < st["live_bytes"] <= st["resident_bytes"]
    assert st["peak_resident_bytes"] >= st["resident_bytes"]
    _check_against_replay(gd, orc, io, reqs, served, cfg)
    srv.close()


@pytest.mark.parametrize("span,inflight,launch", [
    (1 << 12, 3, "graph"), (1 << 32, 2, "graph"), (1 << 32, 3, "direct")])
def test_small_steps_run_fused(gd, orc, io, span, inflight, launch):
    """Small requests submitted as they come: most steps are fused steps
    (descriptor copy, admission reading the requests in place, persistent
    query kernel), in graph mode one captured graph replayed step after
    step with the engine's own launches and waits in between; every result
    equals the step-order replay and the same request run alone."""
    rng = np.random.default_rng(span % 1000 + inflight)
    cfg = dict(max_elements=1 << 20, batch_capacity=1 << 14)
    srv = gd.Server(**cfg, max_inflight_steps=inflight, step_launch=launch)
    reqs = _make_requests(rng, 600, span, 200)
    served, pending = [], []
    for idx in range(len(reqs)):
        # submit in small groups, waiting in between, so steps stay small
        (i2, t, out, found), = _submit_all(gd, io, srv, reqs, order=[idx])
        pending.append((i2, idx, t, out, found))
        if len(pending) == 7 or idx == len(reqs) - 1:
            for i3, order, t3, out3, found3 in pending:
                served.append((i3, order, srv.wait(t3), out3, found3))
            pending = []
    st = srv.stats()
    assert st["graph_steps"] > st["steps"] // 2
    _check_against_replay(gd, orc, io, reqs, served, cfg)
    srv.close()


def test_large_range_steps_compact(gd, orc, io):
    """Steps with at least 2^14 ranges over a stack whose newer runs hold
    more than 1/8 of the oldest run's rows merge the stack on the merge
    stream first and the step's queries wait for it; every result still
    equals the step-order replay and the same request run alone."""
    rng = np.random.default_rng(77)
    span = 1 << 26
    cfg = dict(max_elements=1 << 22, batch_capacity=1 << 16)
    srv = gd.Server(**cfg, max_inflight_steps=2, max_step_keys=1 << 17)

    def keys(n):
        return rng.integers(0, span, n, dtype=np.uint64).astype(np.uint32)

    def ranges(n, kind):
        lo = rng.integers(0, span, n, dtype=np.uint64)
        hi = np.minimum(lo + rng.integers(0, 1 << 16, n, dtype=np.uint64), NF)
        return (kind, lo.astype(np.uint32), hi.astype(np.uint32))

    reqs = [("insert", keys(1 << 18), keys(1 << 18))]
    for t in range(6):
        reqs.append(("insert", keys(1 << (15 - t)), keys(1 << (15 - t))))
        reqs.append(("erase", reqs[0][1][rng.integers(0, 1 << 18, 3000)],
                     None))
        reqs.append(ranges(1 << 15, "rcount" if t % 2 else "rsum"))
        reqs.append(("lookup", keys(5000), None))
    handles = _submit_all(gd, io, srv, reqs)
    served = []
    for order, (idx, t, out, found) in enumerate(handles):
        served.append((idx, order, srv.wait(t), out, found))
    _check_against_replay(gd, orc, io, reqs, served, cfg)
    srv.close()


def test_concurrent_submitters(gd, orc, io):
    rng = np.random.default_rng(21)
    cfg = dict(max_elements=1 << 20, batch_capacity=1 << 14)
    srv = gd.Server(**cfg)
    reqs = _make_requests(rng, 240, 1 << 16, 2000)
    # device copies are made up front so both threads only submit
    staged = []
    for kind, a, b in reqs:
        da = io.up(a)
        db = io.up(b) if b is not None else None
        out = io.empty(len(a)) if kind not in ("insert", "erase") else None
        found = io.empty(len(a), 1) if kind == "lookup" else None
        staged.append((kind, da, db, out, found, len(a)))
    io.torch.cuda.synchronize()
    tickets = [None] * len(reqs)
    lock = threading.Lock()
    order = []

    def worker(ids):
        for i in ids:
            kind, da, db, out, found, n = staged[i]
            with lock:  # arrival order = ticket order
                t = srv.submit(OPS[kind], da, db, out, found, count=n)
                order.append(i)
            tickets[i] = t

    th = [threading.Thread(target=worker, args=(range(k, len(reqs), 2),))
          for k in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    served = []
    for pos, i in enumerate(order):
        step = srv.wait(tickets[i])
        served.append((i, pos, step, staged[i][3], staged[i][4]))
    _check_against_replay(gd, orc, io, reqs, served, cfg)
    srv.close()


def test_budget_admission(gd, io):
    srv = gd.Server(max_elements=1 << 16, batch_capacity=1 << 12,
                    hbm_budget_bytes=1 << 20)
    big = io.up(np.arange(1 << 20, dtype=np.uint32))
    with pytest.raises(gd.CapacityExhausted):
        srv.submit(gd.Server.INSERT, big, big, count=1 << 20)
    small = io.up(np.arange(1000, dtype=np.uint32))
    out = io.empty(1000)
    io.torch.cuda.synchronize()
    t1 = srv.submit(gd.Server.INSERT, small, small, count=1000)
    t2 = srv.submit(gd.Server.LOOKUP, small, None, out, None, count=1000)
    srv.wait(t1)
    srv.wait(t2)
    assert np.array_equal(io.down(out, 1000), np.arange(1000))
    srv.close()
