#!/usr/bin/env python3
"""GTEPS per label-propagation iteration on one MI355X.

Workload: RMAT scale 26, edge factor 16 (a=.57, b=.19, c=.19, seed 26),
symmetrised, deduplicated, self-loops dropped, ids permuted; int32
neighbour ids, int64 offsets, m = 2,103,849,724 directed edges -- the scale
the learned-vs-plain and %-of-HBM targets are quoted on (--scale 22 for the
RMAT-22 configuration; --graph lfr for the LFR-style graph of 8M
vertices).  One step = one full LP run of 10 synchronous iterations from
labels[i] = i (the reference's timed main loop, lfht.cuh:87-106), inputs
resident in HBM.  Prints one JSON line.

Variants, each timed over the same steps:
  hash_learned  method 1, learned dispatch: the 4-32-32-1 network is trained
                on the GPU from distinct-label counts logged during one
                untimed plain run of the same graph (training and logging
                are outside the timed region; inference is inside)
  hash_plain    method 1, degree-class dispatch
  *_tuned       the same after the autotuner picked each class's kernel
                configuration (set-up, outside the timed region)
  sort          method 0 (segmented sorts), degree-class dispatch
  sort_tuned    the same after the autotuner picked each class's sort
                configuration
  *_ooc_sync, *_ooc_async
                out-of-core (policies 1 and 2): the neighbour array stays in
                pinned host memory and streams in batches of < 2^b edges
                (--buffer-pow), without and with copy/compute overlap; not in
                the default list
  batched       512 planted-partition graphs (n up to 20000; --batch-graphs,
                --batch-nmax) in one batched run (separate JSON object, not
                the value)
Default variants: hash_learned_tuned (the value) and hash_tuned, the pair
learned_vs_plain compares.
value        = GTEPS of the first variant (hash_learned_tuned by default)
             = m * iterations / wall time of the timed
               steps (host clock around the steps, device synchronised, no
               instrumentation)
gather_ceiling = k_gather_probe on the same graph and layout: neighbour-id reads
               and label gathers only (8 B per edge), the best rate an LP kernel
               can reach on this access pattern; dominant_kernel_frac and
               iteration_frac compare the roofline kernel and the whole
               iteration (algorithmic bytes per second) with it
roofline     = the dominant kernel of the first variant: its algorithmic bytes per
               launch / its average launch time, from HIP events on the launch
               stream over a second, instrumented pass of the same steps,
               against 8 TB/s; traffic = the HBM bytes per launch of the same
               kernel, workload and variant measured with rocprofv3 counters
               and recorded in profiles/hbm_traffic.json together with a hash
               of the kernel sources they were measured on: null (and
               traffic_stale true) when no entry matches or the kernels have
               changed since
cpu_baseline = the std::map-structured CPU baseline (oracle/lp_baseline.cpp,
               label_propagator.h:52-66 made synchronous, OpenMP
               schedule(dynamic,1)) on a bounded sample of the workload: one
               iteration from the labels of the run's 5th iteration over every
               k-th vertex (k sized for ~10 s), on all OpenMP threads (value)
               and on one; its labels are checked against the oracle's.  The
               full run of the dense CPU oracle (oracle/lp_oracle.c) gives
               the final labels every variant must equal (labels_bit_exact)
"""
import argparse
import glob
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")
METRIC = "GTEPS per LP iteration + % HBM peak; learned vs plain dispatch, bit-exact labels"
MLP_HIDDEN = 32
CAND_THRESH = 14.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_source_hash():
    """sha256 (first 16 hex digits) of the device sources, galp_amd/csrc/*.hip
    and *.hpp in name order: the key that ties a traffic measurement to the
    kernels it was taken on."""
    h = hashlib.sha256()
    for path in sorted(glob.glob(os.path.join(ROOT, "galp_amd", "csrc", "*.hip")) +
                       glob.glob(os.path.join(ROOT, "galp_amd", "csrc", "*.hpp"))):
        h.update(os.path.basename(path).encode())
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(workload, variant, kernel, source_hash=None, alg_bytes=None):
    """HBM bytes per launch of `kernel` (FETCH_SIZE and WRITE_SIZE counters,
    each scaled by what it reads for a copy of known size), measured for this
    workload and variant on kernel sources whose hash is source_hash (the
    current sources when None).  alg_bytes: this run's algorithmic bytes per
    launch; the measured bytes are then rescaled by it (bytes per algorithmic
    byte of the measured launches), because a run's mix of launches -- how
    many iterations the learned dispatcher routes to the kernel, and with
    how much work -- can differ from the counter run's.  Returns (bytes or
    None, stale): stale is True when an entry exists but was measured on
    other kernel sources, and the figure is then withheld."""
    try:
        with open(TRAFFIC_FILE) as f:
            entries = json.load(f).get("entries", [])
    except (OSError, ValueError):
        return None, False
    want = source_hash or kernel_source_hash()
    stale = False
    for e in entries:
        if e.get("workload") == workload and e.get("variant") == variant and e.get("kernel") == kernel:
            if e.get("source_hash") == want:
                b = float(e["bytes_per_launch"])
                if alg_bytes:
                    b *= float(alg_bytes) / float(e["algorithmic_bytes_per_launch"])
                return b, False
            stale = True
    return None, stale


def train_network(ctx, off, nbr, niter):
    """Logs (features, distinct-count) samples during one plain run and
    trains the dispatcher's network on them, all on the GPU."""
    import numpy as np
    from galp_amd import _abi
    t = time.perf_counter()
    with _abi.Graph(ctx, off, nbr) as g:
        big = int((np.diff(off) > 64).sum())
        g.log_enable((niter + 1) * max(big, 1))
        g.run(niter, want_labels=False)
        x, y = g.log_read()
    # the device appends samples in no fixed order: a canonical order first,
    # so that the seeded subsample and training give the same network every run
    order = np.lexsort((y, x[:, 3], x[:, 2], x[:, 1], x[:, 0]))
    x, y = x[order], y[order]
    if len(x) > 3_000_000:   # a seeded subsample keeps training time flat on large graphs
        idx = np.random.default_rng(0).choice(len(x), 3_000_000, replace=False)
        x, y = x[idx], y[idx]
    p0 = _abi.mlp_init(MLP_HIDDEN, 1)
    p, loss = _abi.mlp_train(ctx, MLP_HIDDEN, p0, x, y, epochs=8, batch=512, lr=3e-3, seed=2)
    log(f"bench: network trained on {len(x)} samples, loss {loss:.4f}, {time.perf_counter() - t:.1f}s")
    return p, len(x), loss


def make_batch(count, nmax, seed=4096):
    """Planted-partition graphs of uneven size as one disjoint-union CSR
    (n ~ U[1000, nmax], community size ~ U[32, 1024], intra degree
    ~ U[8, 48], inter degree ~ intra / 4; graph k from the seed stream
    (seed, k)), generated by the library's host threads."""
    from galp_amd import _abi
    return _abi.gen_batch_planted(count, n_range=(1000, nmax), comm_range=(32, 1024), deg_range=(8, 48),
                                  inter_frac=0.25, seed=seed)


def time_batched(ctx, count, nmax, check, steps, niter):
    """One batched run of `count` graphs (one launch per degree class, the
    10-iteration loop captured in a hipGraph) against running graphs one by
    one: `check` graphs spread over the batch are run alone, their labels and
    iteration counts compared with the batched run's, and their one-by-one
    rate measured."""
    import numpy as np
    from galp_amd import _abi
    t = time.perf_counter()
    off, nbr, vbase, spec = make_batch(count, nmax)
    m = len(nbr)
    log(f"bench: batched: {count} graphs, n={len(off) - 1} m={m} generated in {time.perf_counter() - t:.1f}s")
    mk = off[vbase[1:]] - off[vbase[:-1]]   # edges of each graph
    with _abi.Graph(ctx, off, nbr, reorder=0, graph_capture=1) as g:
        g.set_batch(vbase)
        t = time.perf_counter()
        tuning = g.autotune(reps=3).tolist()   # set-up, outside the timed region
        streams = g.tuned_streams
        log(f"bench: batched: autotuned {tuning}, {streams} stream(s), in {time.perf_counter() - t:.1f}s")
        g.run(niter, want_labels=False)
        edge_iters = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            _, res = g.run(niter, want_labels=False)
            edge_iters += int((mk * g.batch_iterations()).sum())
        wall = time.perf_counter() - t0
        labels, _ = g.run(niter)
        per_graph_iters = g.batch_iterations()
    sample = sorted(set(np.linspace(0, count - 1, min(check, count)).astype(int).tolist()))
    edges_iters, t1, ok = 0, 0.0, True
    for k in sample:
        a, b = int(vbase[k]), int(vbase[k + 1])
        o = off[a:b + 1] - off[a]
        nb = nbr[off[a]:off[b]] - np.int32(a)
        with _abi.Graph(ctx, o, nb, reorder=0) as gk:
            gk.run(niter, want_labels=False)
            t0 = time.perf_counter()
            lab, res = gk.run(niter)
            t1 += time.perf_counter() - t0
            edges_iters += len(nb) * res["iterations"]
            ok = ok and np.array_equal(lab, labels[a:b] - a) and res["iterations"] == per_graph_iters[k]
    return {"graphs": count, "n": int(len(off) - 1), "m": int(m),
            "gteps": round(edge_iters / wall / 1e9, 3),
            "mean_iterations_per_graph": round(float(per_graph_iters.mean()), 2),
            "one_by_one_gteps": round(edges_iters / t1 / 1e9, 3),
            "ms_per_run": round(wall / steps * 1e3, 3),
            "graphs_checked_one_by_one": len(sample),
            "tuning": tuning,
            "streams": streams,
            "labels_equal_one_by_one": bool(ok),
            "workload": f"{count} planted-partition graphs, n~U[1000,{nmax}], community~U[32,1024], "
                        f"intra degree~U[8,48], inter degree~intra/4, graph k from seed stream (4096, k); hash "
                        f"method, plain dispatch, autotuned class kernels, hipGraph-captured {niter}-iteration "
                        f"loop, per-graph stop rule, stopped graphs dropped from the work lists"}


def parse_mask(s):
    """The learned-iteration mask from its printed form: character i is
    iteration i + 1 ('1': routed by the network); the last one holds for
    every later iteration."""
    mask = 0
    for i, c in enumerate(s):
        if c == "1":
            mask |= 1 << i
    if s and s[-1] == "1":
        mask |= ((1 << 64) - 1) ^ ((1 << len(s)) - 1)
    return mask


def time_variant(ctx, off, nbr, steps, warmup, niter, mlp=None, tune=False, mask=None, **opts):
    """Times `steps` runs without instrumentation (the value), then the same
    runs again with a HIP event pair around every kernel launch (the
    per-kernel roofline; events cost a few percent of wall time).  tune:
    autotune the class kernels first (set-up, outside the timed region).
    mask: the learned-iteration mask to use instead of the autotuner's (the
    counter runs, where profiling skews the autotuner's timings, pass the
    one a plain run chose so that both launch the same kernels)."""
    from galp_amd import _abi
    out = {}
    choice = None
    for profile in (0, 1):
        with _abi.Graph(ctx, off, nbr, profile=profile, **opts) as g:
            if mlp is not None:
                g.set_mlp(MLP_HIDDEN, mlp, CAND_THRESH)
            if tune:
                if choice is None:
                    choice = g.autotune(reps=3)
                    out["tuning"] = choice.tolist()
                    out["tuned_streams"] = g.tuned_streams
                    if mlp is not None:   # the judge's per-iteration choice of router
                        out["learned_iterations"] = g.learned_iterations
                        if mask is not None:
                            out["learned_iterations"] = mask
                            g.set_learned_iterations(mask)
                else:
                    g.set_tuning(choice)
                    if mlp is not None:
                        g.set_learned_iterations(out["learned_iterations"])
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
            if profile:
                out["profile"] = g.profile()
                out["profiled_wall_s"] = wall
            else:
                out.update({"wall_s": wall, "iters": iters, "device_ms": dev_ms})
                out["labels"], _ = g.run(niter)
                out["info"] = g.info()
    return out


def cpu_legs(off, nbr, gname, niter, variants):
    """The dense oracle's full run (final labels: labels_bit_exact of every
    variant) and the std::map-structured CPU baseline on a bounded sample."""
    import numpy as np
    import oracle
    n, m = len(off) - 1, len(nbr)
    t = time.perf_counter()
    prev = np.arange(n, dtype=np.int32)
    mid = None
    k = 0
    for it in range(niter):
        nxt, c = oracle.lp_iterate(off, nbr, prev, nlabels=n)
        prev = nxt
        k += 1
        if it + 1 == min(5, niter):
            mid = prev.copy()
        if c <= n * 1e-5:
            break
    log(f"bench: CPU oracle: {k} iterations in {time.perf_counter() - t:.1f} s")
    res = {"labels_bit_exact": {name: bool(np.array_equal(v["labels"], prev)) for name, v in variants.items()}}
    if mid is None:
        mid = prev
    deg = np.diff(off)
    nthreads = oracle.baseline_threads()

    def timed(target_s, threads):
        # every k-th vertex: first a small sample for the rate, then one sized
        # for target_s seconds
        for target in (1e6, None):
            if target is None:
                rate = edges / max(secs, 1e-6)
                target = rate * target_s
            stride = max(1, int(round(m / max(target, 1.0))))
            vl = np.arange(0, n, stride, dtype=np.int64)
            edges = int(deg[vl].sum())
            t0 = time.perf_counter()
            got, _ = oracle.map_iterate(off, nbr, mid, vlist=vl, nthreads=threads)
            secs = time.perf_counter() - t0
        return edges, secs, stride, got, vl

    e_all, s_all, st_all, got, vl = timed(10.0, nthreads)
    e_one, s_one, st_one, got1, vl1 = timed(10.0, 1)
    # parity of the baseline itself on its sample: the oracle's next labels
    want, _ = oracle.lp_iterate(off, nbr, mid, nlabels=n)
    ok = bool(np.array_equal(got[vl], want[vl]) and np.array_equal(got1[vl1], want[vl1]))
    it_mid = min(5, niter) + 1
    res["cpu_baseline"] = {
        "value": round(e_all / s_all / 1e9, 5), "unit": "GTEPS", "cores": nthreads, "kind": "port",
        "sample": f"{gname}, iteration {it_mid} (from the labels after iteration {it_mid - 1}) over every "
                  f"{st_all}-th vertex ({e_all} edges, {s_all:.2f} s) on {nthreads} OpenMP threads, "
                  f"std::map histogram per vertex, schedule(dynamic,1)",
        "single_thread": {"value": round(e_one / s_one / 1e9, 5), "unit": "GTEPS", "cores": 1,
                          "sample": f"every {st_one}-th vertex ({e_one} edges, {s_one:.2f} s)"},
        "labels_equal_oracle": ok,
        "cpu_model": oracle.cpu_model(), "nproc": os.cpu_count(),
        "threads_available": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
    }
    log(f"bench: CPU baseline: {res['cpu_baseline']}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--graph", choices=("rmat", "lfr"), default="rmat",
                    help="rmat: RMAT --scale; lfr: LFR-style graph of --lfr-n vertices (gamma 2.5, mean degree 20, "
                         "max degree 2000, community sizes 50..50000 with exponent 1.5, mu 0.3, seed 8)")
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--lfr-n", type=int, default=8_000_000)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--variants", default="hash_learned_tuned,hash_tuned",
                    help="comma-separated; the first is the value.  Others: hash_learned, hash_plain, sort, "
                         "sort_tuned, *_graph (captured loop), *_ooc_sync/async")
    ap.add_argument("--buffer-pow", type=int, default=24,
                    help="out-of-core variants: batches of fewer than 2^b edges (-b)")
    ap.add_argument("--batch-graphs", type=int, default=512,
                    help="graphs of the batched measurement (0: skip it)")
    ap.add_argument("--batch-nmax", type=int, default=20000,
                    help="largest graph of the batch (200000 with --batch-graphs 4096: the full batched config)")
    ap.add_argument("--batch-check", type=int, default=64,
                    help="graphs of the batch also run one by one (labels compared, one-by-one rate)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the parity check")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched measurement (its GTEPS is then the value)")
    ap.add_argument("--save-mlp", help="write the trained dispatcher network (float32 .npy)")
    ap.add_argument("--load-mlp", help="use this network instead of training one (float32 .npy)")
    ap.add_argument("--learned-mask", metavar="BITS",
                    help="learned-iteration mask as printed in learned_iterations (e.g. 1100111111), used instead of "
                         "the autotuner's judge (counter runs)")
    ap.add_argument("--calibrate-copy", type=int, metavar="MIB", default=0,
                    help="only copy MIB MiB device-to-device (a reference kernel for the HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import numpy as np
    from galp_amd import _abi

    if args.calibrate_copy:
        import torch
        a = torch.ones(args.calibrate_copy * (1 << 18), dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": 4 * a.numel(), "launches": 3}), flush=True)
        return

    if args.batch_only:
        ctx = _abi.Context(0)
        bt = time_batched(ctx, args.batch_graphs, args.batch_nmax, args.batch_check, args.steps, args.iters)
        print(json.dumps({"metric": METRIC, "value": bt["gteps"], "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
                          "warmup": 1, "ms_per_step": bt["ms_per_run"], "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "int32", "data": "synthetic",
                          "config": {"workload": bt["workload"], "n": bt["n"], "m": bt["m"]}, "batched": bt}),
              flush=True)
        ctx.close()
        return

    t = time.perf_counter()
    if args.graph == "rmat":
        off, nbr = _abi.gen_rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=args.scale)
        gname = f"RMAT-{args.scale}"
        workload = f"RMAT scale {args.scale}, edge factor 16, a=0.57 b=0.19 c=0.19, seed {args.scale}"
    else:
        off, nbr = _abi.gen_lfr(args.lfr_n, 2.5, 20.0, 2000, 1.5, 50, 50000, 0.3, seed=8)
        gname = f"LFR-{args.lfr_n}"
        workload = (f"LFR-style graph, n={args.lfr_n}, degree exponent 2.5, mean degree 20, max degree 2000, "
                    f"community sizes 50..50000 (exponent 1.5), mu=0.3, seed 8")
    n, m = len(off) - 1, len(nbr)
    log(f"bench: {gname} n={n} m={m} generated in {time.perf_counter() - t:.1f}s")

    ctx = _abi.Context(0)
    dev = ctx.device_info()
    names = [v for v in args.variants.split(",") if v]
    mlp = None
    train_info = None
    if any(nm.startswith("hash_learned") for nm in names):
        if args.load_mlp:
            mlp = np.load(args.load_mlp).astype(np.float32)
            train_info = {"loaded": os.path.basename(args.load_mlp), "hidden": MLP_HIDDEN,
                          "cand_thresh": CAND_THRESH}
        else:
            mlp, nsamples, loss = train_network(ctx, off, nbr, args.iters)
            train_info = {"samples": nsamples, "loss": round(loss, 5), "hidden": MLP_HIDDEN,
                          "cand_thresh": CAND_THRESH}
        if args.save_mlp:
            np.save(args.save_mlp, mlp)
    table = {"hash_learned": dict(method=_abi.METHOD_HASH, dispatch=_abi.DISPATCH_LEARNED),
             "hash_plain": dict(method=_abi.METHOD_HASH, dispatch=_abi.DISPATCH_PLAIN),
             "hash_tuned": dict(method=_abi.METHOD_HASH, dispatch=_abi.DISPATCH_PLAIN, tune=True),
             "hash_learned_tuned_graph": dict(method=_abi.METHOD_HASH, dispatch=_abi.DISPATCH_LEARNED, tune=True,
                                              graph_capture=1),
             "hash_tuned_graph": dict(method=_abi.METHOD_HASH, dispatch=_abi.DISPATCH_PLAIN, tune=True,
                                      graph_capture=1),
             "hash_learned_tuned": dict(method=_abi.METHOD_HASH, dispatch=_abi.DISPATCH_LEARNED, tune=True),
             "sort": dict(method=_abi.METHOD_SORT, dispatch=_abi.DISPATCH_PLAIN),
             "sort_tuned": dict(method=_abi.METHOD_SORT, dispatch=_abi.DISPATCH_PLAIN, tune=True),
             "hash_ooc_sync": dict(method=_abi.METHOD_HASH, policy=_abi.POLICY_OOC_SYNC, buffer_pow=args.buffer_pow),
             "hash_ooc_async": dict(method=_abi.METHOD_HASH, policy=_abi.POLICY_OOC_ASYNC, buffer_pow=args.buffer_pow),
             "sort_ooc_async": dict(method=_abi.METHOD_SORT, policy=_abi.POLICY_OOC_ASYNC, buffer_pow=args.buffer_pow),
             "hash_learned_ooc_sync": dict(method=_abi.METHOD_HASH, dispatch=_abi.DISPATCH_LEARNED,
                                           policy=_abi.POLICY_OOC_SYNC, buffer_pow=args.buffer_pow),
             "hash_learned_ooc_async": dict(method=_abi.METHOD_HASH, dispatch=_abi.DISPATCH_LEARNED,
                                            policy=_abi.POLICY_OOC_ASYNC, buffer_pow=args.buffer_pow)}
    variants = {}
    for name in names:
        r = time_variant(ctx, off, nbr, args.steps, args.warmup, args.iters,
                         mlp=mlp if name.startswith("hash_learned") else None,
                         mask=parse_mask(args.learned_mask) if args.learned_mask else None, **table[name])
        gteps = m * r["iters"] / r["wall_s"] / 1e9
        variants[name] = r
        log(f"bench: {name}: {r['iters']} iterations in {r['wall_s'] * 1e3:.2f} ms wall "
            f"({r['device_ms']:.2f} ms device) -> {gteps:.2f} GTEPS")
        for k in sorted(r["profile"], key=lambda k: -k["total_ms"]):
            if k["launches"]:
                log(f"    {k['name']:18s} {k['launches']:5d} launches {k['total_ms'] / k['launches'] * 1e3:9.1f} us avg "
                    f"{k['alg_bytes'] / (k['total_ms'] * 1e-3) / 1e9 if k['total_ms'] > 0 else 0:8.1f} GB/s alg")

    # the gather ceiling: one pass of neighbour-id reads and label gathers
    # only (k_gather_probe), on the layout every in-core variant uses
    probe_ms = None
    if not any("_ooc_" in nm for nm in names):
        with _abi.Graph(ctx, off, nbr) as g:
            g.init()
            probe_ms, _ = g.gather_probe(reps=5)
        log(f"bench: gather probe: {probe_ms * 1e3:.1f} us per pass of {m} gathers "
            f"-> {m / (probe_ms * 1e-3) / 1e9:.1f} G gathers/s")

    head = names[0]
    main_v = variants[head]
    wall = main_v["wall_s"]
    iters = main_v["iters"]
    gteps = m * iters / wall / 1e9
    per_iter_s = wall / iters
    b_iter = 8.0 * m + 12.0 * n + 4.0
    prof = [k for k in main_v["profile"] if k["launches"] and k["alg_bytes"] > 0]
    dom = max(prof, key=lambda k: k["total_ms"])
    dom_avg_s = dom["total_ms"] / dom["launches"] * 1e-3
    dom_bytes = dom["alg_bytes"] / dom["launches"]
    achieved = dom_bytes / dom_avg_s / 1e9
    src_hash = kernel_source_hash()
    traffic, traffic_stale = measured_traffic(workload, head, dom["name"], src_hash, dom_bytes)
    vg = {k: round(m * v["iters"] / v["wall_s"] / 1e9, 3) for k, v in variants.items()}
    # out-of-core: the neighbour array crosses the host link once per iteration
    link = {k: round(4.0 * m * v["iters"] / v["wall_s"] / 1e9, 2) for k, v in variants.items() if "_ooc_" in k}

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
        "config": {"workload": f"{workload}, {args.iters} synchronous LP iterations per step, " +
                               (f"out-of-core (batches < 2^{args.buffer_pow} edges)" if "_ooc_" in head else "in-core"),
                   "n": n, "m": m, "method": "sort (method 0)" if head.startswith("sort") else "hash (method 1)",
                   "dispatch": "learned" if head.startswith("hash_learned") else "plain",
                   "autotuned": head.endswith("_tuned"),
                   "iterations_per_step": iters // args.steps},
        "roofline": {"bound": "hbm", "kernel": dom["name"], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic) if traffic else None, "alg_bytes_per_launch": dom_bytes,
                     "traffic_source": "profiles/hbm_traffic.json: rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE bytes per "
                                       "algorithmic byte of the kernel's launches in the counter run, times this "
                                       "run's algorithmic bytes per launch; entries keyed by workload, variant, "
                                       "kernel and kernel-source hash",
                     "kernel_source_hash": src_hash, "traffic_stale": traffic_stale,
                     "avg_launch_us": round(dom_avg_s * 1e6, 2)},
        "iteration_roofline": {"alg_bytes": b_iter, "achieved_GBps": round(b_iter / per_iter_s / 1e9, 1),
                               "frac": round(b_iter / per_iter_s / 1e9 / HBM_PEAK_GBS, 4)},
        **({"gather_ceiling": {
            "kernel": "k_gather_probe", "ms_per_pass": round(probe_ms, 4),
            "gathers_per_s": round(m / (probe_ms * 1e-3), 1), "alg_GBps": round(8.0 * m / (probe_ms * 1e-3) / 1e9, 1),
            "dominant_kernel_frac": round(achieved / (8.0 * m / (probe_ms * 1e-3) / 1e9), 4),
            "iteration_frac": round((b_iter / per_iter_s) / (8.0 * m / (probe_ms * 1e-3)), 4)}}
           if probe_ms else {}),
        "variants": vg,
        **({"ooc_link_GBps": link, "ooc_buffer_pow": args.buffer_pow} if link else {}),
        "device": {"name": dev["name"], "arch": dev["arch"], "cus": dev["compute_units"]},
        "preprocess_ms": round(main_v["info"]["preprocess_ms"], 1),
    }
    if args.batch_graphs:
        out["batched"] = time_batched(ctx, args.batch_graphs, args.batch_nmax, args.batch_check, args.steps,
                                      args.iters)
        log(f"bench: batched: {out['batched']}")
    for a, b in (("hash_learned_tuned", "hash_tuned"), ("hash_learned", "hash_plain")):
        if a in vg and b in vg:   # both autotuned, or both with the default configurations
            out["learned_vs_plain"] = round(vg[a] / vg[b], 3)
            out["learned_vs_plain_pair"] = [a, b]
            break
    if train_info:
        out["learned_network"] = train_info
    if "tuning" in main_v:
        out["tuning"] = main_v["tuning"]
        out["tuned_streams"] = main_v["tuned_streams"]
        if "learned_iterations" in main_v:   # bit i: iteration i routed by the network
            out["learned_iterations"] = format(main_v["learned_iterations"] & 0x3FF, "010b")[::-1]

    if not args.no_cpu:
        out.update(cpu_legs(off, nbr, gname, args.iters, variants))
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
