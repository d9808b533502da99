"""GPU parity at the full benchmark sizes, through the C-ABI, against the
oracle on the same inputs:

- RMAT-26 (the scale the learned-vs-plain target is quoted on: n = 2^26,
  m = 2,103,849,724, just below 2^31, so int32 offsets): learned dispatch
  (autotuned) and plain
  dispatch, labels and change counts after every iteration;
- RMAT-27 (m = 4.2e9): learned dispatch in core, and out of core with
  2^24-edge batches streamed from pinned host memory on overlapped streams
  (outofcore.cuh:38-97, async.cuh:66-107): both equal the oracle after
  iterations 1 and 2, and each other's final labels after 10;
- LFR-8M (n = 8e6): both methods and learned dispatch, every iteration;
- the batched configuration: 4096 planted-partition graphs with n up to
  200,000 (m = 1.4e10) in one hipGraph-captured run with the class kernels on
  three streams, and again autotuned; 24 sampled graphs equal independent
  oracle runs (labels and iteration counts).

The networks are trained on the GPU from the distinct-label counts of a
logged plain run (RMAT-22 for the RMAT tests, the graph itself for LFR); the
labels may not depend on them.  Every step prints what it runs first.
"""
import gc
import sys
import time

import numpy as np
import pytest

import oracle
from galp_amd import _abi

pytestmark = pytest.mark.gpu

LEARNED = _abi.DISPATCH_LEARNED


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def trained_network(ctx, off, nbr, niter=10, max_samples=1_000_000):
    with _abi.Graph(ctx, off, nbr) as g:
        big = int((np.diff(off) > 64).sum())
        g.log_enable((niter + 1) * max(big, 1))
        g.run(niter, want_labels=False)
        x, y = g.log_read()
    order = np.lexsort((y, x[:, 3], x[:, 2], x[:, 1], x[:, 0]))
    x, y = x[order], y[order]
    if len(x) > max_samples:
        idx = np.random.default_rng(0).choice(len(x), max_samples, replace=False)
        x, y = x[idx], y[idx]
    p, loss = _abi.mlp_train(ctx, 32, _abi.mlp_init(32, 1), x, y, epochs=4, batch=512, lr=3e-3, seed=2)
    say(f"[scale] network trained on {len(x)} samples, loss {loss:.4f}")
    return p


@pytest.fixture(scope="module")
def rmat_net(ctx):
    t = time.perf_counter()
    off, nbr = _abi.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=22)
    p = trained_network(ctx, off, nbr)
    say(f"[scale] RMAT-22 network in {time.perf_counter() - t:.1f} s")
    return p


def oracle_iterations(off, nbr, niter):
    """Labels and change counts of the oracle's first niter iterations (or
    up to its stop)."""
    n = len(off) - 1
    prev = np.arange(n, dtype=np.int32)
    labs, counts = [], []
    t = time.perf_counter()
    for _ in range(niter):
        prev, c = oracle.lp_iterate(off, nbr, prev, nlabels=n)
        labs.append(prev)
        counts.append(c)
        if c <= n * 1e-5:
            break
    say(f"[scale] oracle: {len(labs)} iterations in {time.perf_counter() - t:.1f} s")
    return labs, counts


def step_against(g, labs, counts, tag):
    g.init()
    for it, (ref, cnt) in enumerate(zip(labs, counts)):
        st = g.step()
        got = g.labels()
        assert st["updated"] == cnt, f"{tag}: iteration {it + 1} change count"
        assert np.array_equal(got, ref), f"{tag}: iteration {it + 1} labels differ from the oracle"
    assert st["stopped"] == (counts[-1] <= len(labs[0]) * 1e-5)


def test_rmat26_learned_and_plain_every_iteration(ctx, rmat_net):
    t = time.perf_counter()
    off, nbr = _abi.gen_rmat(26, 16, 0.57, 0.19, 0.19, seed=26)
    say(f"[scale] RMAT-26 n={len(off) - 1} m={len(nbr)} generated in {time.perf_counter() - t:.1f} s")
    labs, counts = oracle_iterations(off, nbr, 10)
    say("[scale] RMAT-26 learned dispatch, autotuned")
    with _abi.Graph(ctx, off, nbr, dispatch=LEARNED) as g:
        g.set_mlp(32, rmat_net, 14.0)
        g.autotune(reps=1)
        step_against(g, labs, counts, "RMAT-26 learned")
        assert g.info()["offsets_are_64"] == int(len(nbr) >= 2 ** 31)   # m = 2,103,849,724 < 2^31
    say("[scale] RMAT-26 plain dispatch")
    with _abi.Graph(ctx, off, nbr) as g:
        step_against(g, labs, counts, "RMAT-26 plain")
    del off, nbr, labs
    gc.collect()


def test_rmat27_incore_and_out_of_core(ctx, rmat_net):
    t = time.perf_counter()
    off, nbr = _abi.gen_rmat(27, 16, 0.57, 0.19, 0.19, seed=27)
    say(f"[scale] RMAT-27 n={len(off) - 1} m={len(nbr)} generated in {time.perf_counter() - t:.1f} s")
    labs, counts = oracle_iterations(off, nbr, 2)
    finals = {}
    for policy in (_abi.POLICY_INCORE, _abi.POLICY_OOC_ASYNC):
        say(f"[scale] RMAT-27 learned dispatch, policy {policy} (-b 24)")
        t = time.perf_counter()
        with _abi.Graph(ctx, off, nbr, dispatch=LEARNED, policy=policy, buffer_pow=24) as g:
            g.set_mlp(32, rmat_net, 14.0)
            g.init()
            for it in range(10):
                st = g.step()
                if it < len(labs):
                    assert st["updated"] == counts[it], f"policy {policy}: iteration {it + 1} change count"
                    assert np.array_equal(g.labels(), labs[it]), f"policy {policy}: iteration {it + 1} labels"
                if st["stopped"]:
                    break
            finals[policy] = (g.labels(), it + 1)
        say(f"[scale] RMAT-27 policy {policy}: {finals[policy][1]} iterations, {time.perf_counter() - t:.1f} s")
    a, b = finals[_abi.POLICY_INCORE], finals[_abi.POLICY_OOC_ASYNC]
    assert a[1] == b[1] and np.array_equal(a[0], b[0]), "in-core and out-of-core final labels differ"
    del off, nbr, labs, finals
    gc.collect()


def test_lfr8m_both_methods_and_learned(ctx):
    t = time.perf_counter()
    off, nbr = _abi.gen_lfr(8_000_000, 2.5, 20.0, 2000, 1.5, 50, 50000, 0.3, seed=8)
    say(f"[scale] LFR-8M n={len(off) - 1} m={len(nbr)} generated in {time.perf_counter() - t:.1f} s")
    labs, counts = oracle_iterations(off, nbr, 10)
    for method in (_abi.METHOD_HASH, _abi.METHOD_SORT):
        say(f"[scale] LFR-8M method {method}")
        with _abi.Graph(ctx, off, nbr, method=method) as g:
            step_against(g, labs, counts, f"LFR-8M method {method}")
    net = trained_network(ctx, off, nbr)
    say("[scale] LFR-8M learned dispatch with a network trained on its own logged counts")
    with _abi.Graph(ctx, off, nbr, dispatch=LEARNED) as g:
        g.set_mlp(32, net, 14.0)
        step_against(g, labs, counts, "LFR-8M learned")


def test_batch4096_captured_three_streams(ctx):
    t = time.perf_counter()
    off, nbr, vbase, spec = _abi.gen_batch_planted(4096, n_range=(1000, 200000), comm_range=(32, 1024),
                                                   deg_range=(8, 48), inter_frac=0.25, seed=4096)
    say(f"[scale] batch: 4096 graphs n={len(off) - 1} m={len(nbr)} generated in {time.perf_counter() - t:.1f} s")
    sample = sorted(set(np.linspace(0, 4095, 24).astype(int).tolist()))
    results = {}
    for tuned in (False, True):
        say(f"[scale] batch: captured 10-iteration run, autotuned={tuned}")
        with _abi.Graph(ctx, off, nbr, reorder=0, graph_capture=1) as g:
            g.set_batch(vbase)
            if tuned:
                g.autotune(reps=1)
                say(f"[scale] batch: autotuned, {g.tuned_streams} stream(s)")
            for rep in range(2):   # the captured graph is replayed: every replay must agree
                labels, _ = g.run(10)
                iters = g.batch_iterations()
                if rep:
                    assert np.array_equal(labels, results[tuned][0]) and np.array_equal(iters, results[tuned][1])
                results[tuned] = (labels, iters)
    assert np.array_equal(results[False][0], results[True][0])
    labels, iters = results[False]
    for k in sample:
        a, b = int(vbase[k]), int(vbase[k + 1])
        o = off[a:b + 1] - off[a]
        nb = nbr[off[a]:off[b]] - np.int32(a)
        ref, _, cnt = oracle.lp_run(o, nb, niter=10, keep_iters=False)
        assert np.array_equal(labels[a:b] - a, ref), f"graph {k} labels"
        assert iters[k] == len(cnt), f"graph {k} iterations"
    say(f"[scale] batch: {len(sample)} sampled graphs equal independent oracle runs")
    del off, nbr, results
    gc.collect()
