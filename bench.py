"""Benchmark of the LSM dictionary on one MI355X.

Workload (one step, config 2: GPULSMOpt at scale): clear, then 2^26 uint32
keys in 64 mixed batches of 2^20 (10% of the entries of every batch after
the first are tombstones of earlier keys, values murmur3(key)), then 2^26
lookups (70% drawn from inserted keys) and 2^20 range counts with widths
log-uniform in [2^4, 2^20].  Plain bulk form: every batch and query set is
one API call in arrival order, each synchronised, so per-call latency is
completion - submission.

Prints ONE JSON line (see the driver contract): value = whole-step mixed-op
throughput in Mops/s with inputs resident in HBM; roofline of the dominant
kernel from HIP events on the launch stream; cpu_baseline = the
multithreaded CPU LSM on the host cores (the full step once, same inputs).
"""

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__ as entry  # noqa: E402

METRIC = ("mixed-op Mops/s + p99 latency, serving vs plain bulk, 2^26 keys, "
          "bit-exact")
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale-log", type=int, default=0,
                    help="shrink every size by 2^s (quick runs)")
    ap.add_argument("--cpu-threads", type=int,
                    default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def pmc_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, each divided by
    the copy-kernel calibration factor stated in the file)."""
    path = os.path.join(ROOT, "profiles", kernel + "_pmc.txt")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    vals = {}
    for key in ("FETCH_KIB_PER_DISPATCH", "WRITE_KIB_PER_DISPATCH",
                "CALIBRATION_FACTOR"):
        m = re.search(key + r"\s*[:=]\s*([0-9.eE+]+)", text)
        if not m:
            return None
        vals[key] = float(m.group(1))
    kib = vals["FETCH_KIB_PER_DISPATCH"] + vals["WRITE_KIB_PER_DISPATCH"]
    return kib * 1024.0 / vals["CALIBRATION_FACTOR"]


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: every box here has one GPU, "
              "and the dictionary is single-GPU (see the sharding note in "
              "DESIGN.md). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    pkg = entry.load_package()
    import gdict_amd.workload as wl

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()

    # ---------------------------------------------------------- workload
    t0 = time.time()
    w = wl.config2(args.scale_log)
    batches = list(w.batches())
    keys = np.concatenate([b[0] for b in batches])
    vals = np.concatenate([b[1] for b in batches])
    erase = np.concatenate([b[2] for b in batches])
    inserted = keys[erase == 0]
    queries = w.lookups(inserted)
    lo, hi = w.ranges()
    gen_s = time.time() - t0

    def up(a, dt=np.int32):
        return torch.from_numpy(np.ascontiguousarray(a).view(dt)).to(dev)

    d_keys, d_vals = up(keys), up(vals)
    d_erase = up(erase, np.uint8)
    d_q = up(queries)
    d_lo, d_hi = up(lo), up(hi)
    d_out = torch.empty(len(queries), dtype=torch.int32, device=dev)
    d_found = torch.empty(len(queries), dtype=torch.uint8, device=dev)
    d_cnt = torch.empty(len(lo), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    n_upd = len(keys)
    n_ops = n_upd + len(queries) + len(lo)
    d = pkg.GPULSMOpt(max_elements=w.n_keys, batch_capacity=w.batch)
    s_ptr = stream.cuda_stream

    def ev():
        e = torch.cuda.Event(enable_timing=True)
        e.record(stream)
        return e

    def step(record):
        lat = []
        phases = {"update": 0.0, "lookup": 0.0, "range": 0.0}
        d.clear(stream=s_ptr)
        for t in range(w.n_batches):
            a = t * w.batch
            n = min(w.batch, n_upd - a)
            h0 = time.perf_counter()
            e0 = ev()
            d.apply_mixed(d_keys[a:a + n], d_vals[a:a + n],
                          d_erase[a:a + n], count=n, stream=s_ptr)
            e1 = ev()
            e1.synchronize()
            lat.append((time.perf_counter() - h0) * 1e3)
            phases["update"] += e0.elapsed_time(e1)
        shape = d.shape() if record else None
        h0 = time.perf_counter()
        e0 = ev()
        d.lookup(d_q, d_out, d_found, stream=s_ptr)
        e1 = ev()
        e1.synchronize()
        lat.append((time.perf_counter() - h0) * 1e3)
        phases["lookup"] = e0.elapsed_time(e1)
        h0 = time.perf_counter()
        e0 = ev()
        d.range_count(d_lo, d_hi, d_cnt, stream=s_ptr)
        e1 = ev()
        e1.synchronize()
        lat.append((time.perf_counter() - h0) * 1e3)
        phases["range"] = e0.elapsed_time(e1)
        return lat, phases, shape

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    lats, ph_tot = [], {"update": 0.0, "lookup": 0.0, "range": 0.0}
    shape = None
    t_start = time.perf_counter()
    for k in range(args.steps):
        lat, ph, sh = step(k == 0)
        lats += lat
        for key in ph_tot:
            ph_tot[key] += ph[key]
        shape = sh or shape
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    # the shape() call of the first timed step synchronises once; it is
    # outside every phase event and inside the wall clock (conservative)
    ms_per_step = elapsed * 1e3 / args.steps
    value = n_ops / (ms_per_step * 1e-3) / 1e6
    lat_arr = np.array(lats)

    # ---------------------------------------------------------- roofline
    runs, rows = shape
    found = d_found.cpu().numpy()
    counts = d_cnt.cpu().numpy().view(np.uint32).astype(np.uint64)
    Q = len(queries)
    lookup_ms = ph_tot["lookup"] / args.steps
    lookup_bytes = Q * (4 + 4 + 1) + min(rows * 8, Q * runs * 64)
    R = len(lo)
    range_ms = ph_tot["range"] / args.steps
    range_bytes = R * (2 * 4 + 4) + min(rows * 8, int(counts.sum()) * 8 +
                                        R * runs * 64)
    kernels = {
        "lookup_kernel": (lookup_ms, lookup_bytes),
        "range_kernel": (range_ms, range_bytes),
    }
    dom = max(kernels, key=lambda k: kernels[k][0])
    dom_ms, dom_bytes = kernels[dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom,
                "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": pmc_traffic(dom),
                "algorithmic_bytes": int(dom_bytes),
                "avg_launch_ms": round(dom_ms, 4),
                "terms": {"Q": Q, "R_rows": int(rows), "L_runs": int(runs)}}

    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "config2: GPULSMOpt at scale, plain bulk "
                        "(synchronous calls)",
            "keys": int(n_upd), "batch": int(w.batch),
            "erase_fraction": 0.1, "lookups": int(Q),
            "lookup_hit_draw": 0.7, "range_counts": int(R),
            "range_width": [w.wmin, w.wmax], "scale_log": args.scale_log,
        },
        "latency_ms": {"p50": round(float(np.percentile(lat_arr, 50)), 4),
                       "p99": round(float(np.percentile(lat_arr, 99)), 4),
                       "calls_per_step": len(lats) // args.steps},
        "phase_ms": {k: round(v / args.steps, 4) for k, v in ph_tot.items()},
        "lookup_hit_rate": round(float(found.mean()), 4),
        "roofline": roofline,
        "workload_gen_s": round(gen_s, 2),
    }

    # ------------------------------------------------------ CPU baseline
    if not args.no_cpu:
        import oracle as orc
        cpu = orc.CpuLsm(args.cpu_threads)
        frac = 8  # lookups and ranges are timed on a 1/8 prefix
        qs, rs = Q // frac, R // frac
        c0 = time.perf_counter()
        for t in range(w.n_batches):
            a = t * w.batch
            n = min(w.batch, n_upd - a)
            cpu.apply(keys[a:a + n], vals[a:a + n], erase[a:a + n])
        c1 = time.perf_counter()
        c_out, c_found = cpu.lookup(queries[:qs], with_found=True)
        c2 = time.perf_counter()
        c_cnt = cpu.range(lo[:rs], hi[:rs], count_only=True)
        c3 = time.perf_counter()
        cpu_s = (c1 - c0) + (c2 - c1) * Q / qs + (c3 - c2) * R / rs
        g_out = d_out.cpu().numpy().view(np.uint32)
        exact = (np.array_equal(c_out, g_out[:qs]) and
                 np.array_equal(c_found, found[:qs]) and
                 np.array_equal(c_cnt, counts[:rs].astype(np.uint32)))
        result["cpu_baseline"] = {
            "value": round(n_ops / cpu_s / 1e6, 3), "unit": "Mops/s",
            "cores": args.cpu_threads, "kind": "port",
            "sample": "same inputs: all %d updates timed in full, the first "
                      "%d of %d lookups and %d of %d range counts timed and "
                      "scaled by their op counts; multithreaded sorted-array "
                      "LSM (oracle/cpu_lsm.cpp)" % (n_upd, qs, Q, rs, R),
            "seconds_measured": round(c3 - c0, 3),
            "seconds_full_step_est": round(cpu_s, 3)}
        result["bit_exact_vs_cpu"] = bool(exact)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
