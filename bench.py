#!/usr/bin/env python3
"""GTEPS per label-propagation iteration on one MI355X.

Workload: RMAT scale 22, edge factor 16 (a=.57, b=.19, c=.19, seed 22),
symmetrised, deduplicated, self-loops dropped, ids permuted; int32 CSR with
m = 128,302,318 directed edges.  One step = one full LP run of 10
synchronous iterations from labels[i] = i (the reference's timed main loop,
lfht.cuh:87-106), inputs resident in HBM.  Prints one JSON line.

value      = m * iterations / wall time of the timed steps, in GTEPS
roofline   = the dominant kernel's algorithmic bytes / its average launch
             time (HIP events on the launch stream), against 8 TB/s
cpu_baseline = the CPU oracle (oracle/lp_oracle.c, OpenMP) on the same
             graph and iterations; its final labels are also the parity check
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "GTEPS per LP iteration + % HBM peak; learned vs plain dispatch, bit-exact labels"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_variant(ctx, off, nbr, steps, warmup, niter, **opts):
    from galp_amd import _abi
    with _abi.Graph(ctx, off, nbr, profile=1, **opts) as g:
        for _ in range(warmup):
            g.run(niter, want_labels=False)
        g.profile_reset()
        iters = 0
        dev_ms = 0.0
        t0 = time.perf_counter()
        for _ in range(steps):
            _, res = g.run(niter, want_labels=False)
            iters += res["iterations"]
            dev_ms += res["device_ms"]
        wall = time.perf_counter() - t0
        prof = g.profile()
        labels, _ = g.run(niter)
        info = g.info()
    return {"wall_s": wall, "iters": iters, "device_ms": dev_ms, "profile": prof, "labels": labels,
            "info": info}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the parity check")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import numpy as np
    from galp_amd import _abi

    t = time.perf_counter()
    off, nbr = _abi.gen_rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=args.scale)
    n, m = len(off) - 1, len(nbr)
    log(f"bench: RMAT-{args.scale} n={n} m={m} generated in {time.perf_counter() - t:.1f}s")

    ctx = _abi.Context(0)
    dev = ctx.device_info()
    variants = {}
    for name, opts in (("hash_plain", dict(method=_abi.METHOD_HASH, dispatch=_abi.DISPATCH_PLAIN)),):
        r = time_variant(ctx, off, nbr, args.steps, args.warmup, args.iters, **opts)
        gteps = m * r["iters"] / r["wall_s"] / 1e9
        variants[name] = r
        log(f"bench: {name}: {r['iters']} iterations in {r['wall_s'] * 1e3:.2f} ms wall "
            f"({r['device_ms']:.2f} ms device) -> {gteps:.2f} GTEPS")
        for k in sorted(r["profile"], key=lambda k: -k["total_ms"]):
            if k["launches"]:
                log(f"    {k['name']:18s} {k['launches']:5d} launches {k['total_ms'] / k['launches'] * 1e3:9.1f} us avg "
                    f"{k['alg_bytes'] / (k['total_ms'] * 1e-3) / 1e9 if k['total_ms'] > 0 else 0:8.1f} GB/s alg")

    main_v = variants["hash_plain"]
    wall = main_v["wall_s"]
    iters = main_v["iters"]
    gteps = m * iters / wall / 1e9
    per_iter_s = wall / iters
    b_iter = 8.0 * m + 12.0 * n + 4.0
    prof = [k for k in main_v["profile"] if k["launches"]]
    dom = max(prof, key=lambda k: k["total_ms"])
    dom_avg_s = dom["total_ms"] / dom["launches"] * 1e-3
    dom_bytes = dom["alg_bytes"] / dom["launches"]
    achieved = dom_bytes / dom_avg_s / 1e9

    out = {
        "metric": METRIC,
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"RMAT scale {args.scale}, edge factor 16, a=0.57 b=0.19 c=0.19, seed {args.scale}, "
                               f"{args.iters} synchronous LP iterations per step, in-core",
                   "n": n, "m": m, "method": "hash (method 1)", "dispatch": "plain",
                   "iterations_per_step": iters // args.steps},
        "roofline": {"bound": "hbm", "kernel": dom["name"], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "alg_bytes_per_launch": dom_bytes,
                     "avg_launch_us": round(dom_avg_s * 1e6, 2)},
        "iteration_roofline": {"alg_bytes": b_iter, "achieved_GBps": round(b_iter / per_iter_s / 1e9, 1),
                               "frac": round(b_iter / per_iter_s / 1e9 / HBM_PEAK_GBS, 4)},
        "variants": {k: round(m * v["iters"] / v["wall_s"] / 1e9, 3) for k, v in variants.items()},
        "device": {"name": dev["name"], "arch": dev["arch"], "cus": dev["compute_units"]},
        "preprocess_ms": round(main_v["info"]["preprocess_ms"], 1),
    }

    if not args.no_cpu:
        import oracle
        cores = min(16, os.cpu_count() or 1)
        t = time.perf_counter()
        final, per, counts = oracle.lp_run(off, nbr, niter=args.iters, nthreads=cores, keep_iters=False)
        cpu_s = time.perf_counter() - t
        k = len(counts)
        out["cpu_baseline"] = {"value": round(m * k / cpu_s / 1e9, 4), "unit": "GTEPS", "cores": cores,
                               "kind": "port",
                               "sample": f"the full workload: RMAT-{args.scale}, {k} iterations, "
                                         f"{cpu_s:.2f} s wall on {cores} threads"}
        out["labels_bit_exact"] = {name: bool(np.array_equal(v["labels"], final)) for name, v in variants.items()}
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
