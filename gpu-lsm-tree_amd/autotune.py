"""Autotuner for the launch-shape knobs (run on the target GPU).

It measures, on the GPU at hand, the knobs that change speed and never a
result:

* LSM (gd_lsm_tuning): directory density, merge tile, onesweep block shape,
  the batch size from which lookups are grouped, the fused small-batch
  limit, the run cap, the admission path (segmented or radix) and the most
  runs one merge pass takes;
* serving: keys coalesced per step and steps in flight;
* FliX: slots per node, initial fill and the chain length that triggers a
  rebuild.

It writes gpu-lsm-tree_amd/tuned.json, which GPULSMOpt(tuning="auto"),
Server(tuning="auto") and FlixIndex(tuning="auto") load.  The LSM and
serving knobs are searched by coordinate descent from the defaults over the
bulk LSM step (2^26 uint32 keys in 2^20-key mixed batches, lookups, range
counts), the FliX knobs by a full grid over the FliX update workload
(clustered uint64 build, mixed batches, lookups, range enumerations), both
scaled by 2^-scale_log (a few minutes).

    python gpu-lsm-tree_amd/autotune.py [--scale-log 2] [--out PATH]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(_HERE, "tuned.json")

LSM_SPACE = {
    "dir_rows_log": [2, 3, 4],
    "merge_tile_log": [11, 12],
    "sort_block": [256, 512, 1024],
    "grouped_lookup_min_log": [16, 18, 20, 63],
    "small_batch_max": [1024, 4096],
    "max_runs": [6, 10, 16],
    "segmented_admission": [1, 2],
    "merge_block": [256, 512],
    "merge_ways": [2, 4, 8],
}
LSM_DEFAULT = {"dir_rows_log": 3, "merge_tile_log": 12, "sort_block": 1024,
               "grouped_lookup_min_log": 18, "small_batch_max": 4096,
               "max_runs": 10, "segmented_admission": 1, "merge_block": 512,
               "merge_ways": 8}
SERVE_SPACE = {"max_step_keys": [1 << 20, 1 << 22, 1 << 24],
               "max_inflight_steps": [1, 2, 3]}
FLIX_SPACE = {"node_slots": [16, 32, 64], "fill_percent": [50, 75],
              "rebuild_chain_nodes": [0, 16, 32, 64, 128]}


def _pkg():
    if "gdict_amd" in sys.modules:
        return sys.modules["gdict_amd"]
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "gdict_amd", os.path.join(_HERE, "__init__.py"),
        submodule_search_locations=[_HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gdict_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


class LsmBench:
    """The bulk LSM step (2^26 uint32 keys in 2^20-key mixed batches with
    10% erases, 2^26 lookups, 2^20 range counts), scaled: mixed batches,
    grouped lookups, ranges."""

    def __init__(self, pkg, torch, dev, scale_log):
        import gdict_amd.workload as wl
        self.pkg, self.torch, self.dev = pkg, torch, dev
        w = wl.config2(scale_log)
        b = list(w.batches())
        keys = np.concatenate([x[0] for x in b])
        vals = np.concatenate([x[1] for x in b])
        erase = np.concatenate([x[2] for x in b])
        q = w.lookups(w.live_keys(b))
        lo, hi = w.ranges()
        up = lambda a, dt=np.int32: torch.from_numpy(  # noqa: E731
            np.ascontiguousarray(a).view(dt)).to(dev)
        self.w = w
        self.k, self.v, self.e = up(keys), up(vals), up(erase, np.uint8)
        self.q, self.lo, self.hi = up(q), up(lo), up(hi)
        self.out = torch.empty(len(q), dtype=torch.int32, device=dev)
        self.found = torch.empty(len(q), dtype=torch.uint8, device=dev)
        self.cnt = torch.empty(len(lo), dtype=torch.int32, device=dev)
        torch.cuda.synchronize()

    def step(self, d):
        w = self.w
        d.clear()
        for t in range(w.n_batches):
            a = t * w.batch
            n = min(w.batch, w.n_keys - a)
            d.apply_mixed(self.k[a:a + n], self.v[a:a + n],
                          self.e[a:a + n], count=n)
        d.lookup(self.q, self.out, self.found)
        d.range_count(self.lo, self.hi, self.cnt)

    def time_ms(self, knobs, reps=3):
        d = self.pkg.GPULSMOpt(max_elements=self.w.n_keys,
                               batch_capacity=self.w.batch, tuning=knobs)
        self.step(d)
        self.torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            self.step(d)
        self.torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        d.close()
        return ms


def coordinate_descent(space, start, measure, log):
    best = dict(start)
    measure(best)  # warm-up: the first run pays one-time costs
    best_ms = measure(best)
    log.append({"knobs": dict(best), "ms": best_ms})
    for knob, values in space.items():
        for v in values:
            if v == best[knob]:
                continue
            trial = dict(best)
            trial[knob] = v
            ms = measure(trial)
            log.append({"knobs": dict(trial), "ms": ms})
            if ms < best_ms * 0.99:  # ignore noise-level differences
                best, best_ms = trial, ms
    return best, best_ms


def tune_serve(pkg, torch, dev, lsm_knobs, scale_log, log):
    import gdict_amd.workload as wl
    n_base = 1 << (24 - scale_log)
    base = wl.uniform31(0xBA5E, n_base)
    reqs = wl.RequestStream(wl.SEEDS[4], 400, 32, 1 << (18 - scale_log)) \
        .requests(base)
    up = lambda a: torch.from_numpy(  # noqa: E731
        np.ascontiguousarray(a).view(np.int32)).to(dev)
    staged = []
    for kind, a, b in reqs:
        out = torch.empty(len(a), dtype=torch.int32, device=dev) \
            if kind in ("lookup", "range") else None
        staged.append((kind, up(a), up(b) if b is not None else None, out,
                       len(a)))
    bk, bv = up(base), up(wl.murmur3_u32(base))
    torch.cuda.synchronize()
    S = pkg.Server
    opmap = {"insert": S.INSERT, "erase": S.ERASE, "lookup": S.LOOKUP,
             "range": S.RANGE_COUNT}

    def measure(knobs):
        srv = S(max_elements=4 * n_base, batch_capacity=1 << 20,
                tuning=lsm_knobs, **knobs)
        srv.wait(srv.submit(S.INSERT, bk, bv))
        srv.drain()
        t0 = time.perf_counter()
        for kind, a, b, out, n in staged:
            srv.submit(opmap[kind], a, b, out, None, count=n)
        srv.drain()
        ms = (time.perf_counter() - t0) * 1e3
        srv.close()
        return ms

    return coordinate_descent(SERVE_SPACE, {"max_step_keys": 1 << 22,
                                            "max_inflight_steps": 2},
                              measure, log)


def tune_flix(pkg, torch, dev, scale_log, log, batches=8):
    """Every combination of the FliX knobs (a grid: a fuller node wants
    another rebuild threshold, which a descent along one knob at a time
    from a single start misses) on the FliX update workload's timed work
    (2^27 keys, 2^22-op batches, 2^24 lookups and 2^18 range enumerations
    per batch) scaled by 2^-scale_log: a build of clustered uint64 keys
    (untimed), then
    `batches` batches of mixed updates (75% inserts of uniform keys, most
    of which land in the gaps between the clustered runs, so those chains
    grow from batch to batch; 25% deletes), each followed by its lookups
    and range enumerations (count pass, then fill pass).  Best of two runs
    per combination."""
    import itertools
    import gdict_amd.workload as wl
    w = wl.config3(scale_log)
    built = w.build_keys()
    up = lambda a: torch.from_numpy(  # noqa: E731
        np.ascontiguousarray(a).view(np.int64)).to(dev)
    dk = up(built)
    work = []
    for t in range(batches):
        ins, vals, dels = w.batch_ops(t, built)
        q, lo, hi = w.queries(t, built)
        work.append(tuple(up(a) for a in (ins, vals, dels, q, lo, hi)))
    out = torch.empty(w.lookups, dtype=torch.int64, device=dev)
    offs = torch.empty(w.ranges + 1, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()

    def measure(knobs):
        f = pkg.FlixIndex(key_bytes=8, value_bytes=8, **knobs)
        f.build_bucket_layer_only(dk, max_size=len(built) + 2 * batches *
                                  w.batch)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for di, dv, dd, dq, dlo, dhi in work:
            f.insert(di, dv)
            f.remove(dd)
            f.lookup(dq, out[:len(dq)])
            f.range(dlo, dhi, offs)
            total = max(int(offs[-1].item()), 1)
            ok = torch.empty(total, dtype=torch.int64, device=dev)
            ov = torch.empty(total, dtype=torch.int64, device=dev)
            f.range(dlo, dhi, offs, ok, ov)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        f.close()
        return ms

    best, best_ms = None, None
    names = list(FLIX_SPACE)
    for combo in itertools.product(*(FLIX_SPACE[k] for k in names)):
        knobs = dict(zip(names, combo))
        ms = min(measure(knobs), measure(knobs))
        log.append({"knobs": knobs, "ms": ms})
        if best_ms is None or ms < best_ms:
            best, best_ms = knobs, ms
    return best, best_ms


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale-log", type=int, default=2)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args(argv)
    import torch
    pkg = _pkg()
    dev = torch.device("cuda:0")
    log = {"lsm": [], "serve": [], "flix": []}
    bench = LsmBench(pkg, torch, dev, args.scale_log)
    lsm, lsm_ms = coordinate_descent(LSM_SPACE, LSM_DEFAULT, bench.time_ms,
                                     log["lsm"])
    print("lsm", lsm, round(lsm_ms, 3), flush=True)
    serve, serve_ms = tune_serve(pkg, torch, dev, lsm, args.scale_log,
                                 log["serve"])
    print("serve", serve, round(serve_ms, 3), flush=True)
    flix, flix_ms = tune_flix(pkg, torch, dev, args.scale_log, log["flix"])
    print("flix", flix, round(flix_ms, 3), flush=True)
    result = {"device": torch.cuda.get_device_name(0),
              "scale_log": args.scale_log, "lsm": lsm, "serve": serve,
              "flix": flix, "measurements_ms": log}
    with open(args.out, "w") as f:
        json.dump(result, f, indent=1)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
