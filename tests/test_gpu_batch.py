"""Batched runs (many graphs as one disjoint union) and the hipGraph-captured
iteration loop, through the C-ABI.  Every graph of a batch must match an
independent oracle run of it (labels shifted by its vertex base, and its own
iteration count); captured runs are replayed several times with eager work
between the replays and each replay is checked."""
import sys

import numpy as np
import pytest

import oracle
from galp_amd import _abi

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_batch(seed, count):
    """Planted-partition graphs of uneven size plus small RMATs (hubs)."""
    rng = np.random.default_rng(seed)
    parts = []
    for k in range(count):
        if k % 5 == 4:
            off, nbr = _abi.gen_rmat(12 + k % 3, 16, seed=100 + k)
        else:
            n = int(rng.integers(1000, 6000))
            comm = int(rng.integers(32, 512))
            intra = int(rng.integers(4, 24))
            off, nbr = _abi.gen_planted_partition(n, comm, intra, max(1, intra // 4), 1000 + k)
        parts.append((off, nbr))
    vbase = np.zeros(count + 1, np.int64)
    ebase = 0
    offs, nbrs = [np.zeros(1, np.int64)], []
    for k, (off, nbr) in enumerate(parts):
        n = len(off) - 1
        vbase[k + 1] = vbase[k] + n
        offs.append(off[1:] + ebase)
        nbrs.append(nbr + np.int32(vbase[k]))
        ebase += len(nbr)
    return parts, vbase, np.concatenate(offs), np.concatenate(nbrs).astype(np.int32)


def check_batch(labels, iters, parts, vbase, niter):
    for k, (off, nbr) in enumerate(parts):
        final, per, counts = oracle.lp_run(off, nbr, niter=niter)
        got = labels[vbase[k]:vbase[k + 1]] - vbase[k]
        assert np.array_equal(got, final), f"graph {k}"
        assert iters[k] == len(counts), (k, iters[k], len(counts))


@pytest.mark.parametrize("method", [_abi.METHOD_HASH, _abi.METHOD_SORT])
@pytest.mark.parametrize("reorder", [1, 0])
def test_batched_equals_independent_runs(ctx, method, reorder):
    parts, vbase, off, nbr = make_batch(1, 24)
    say(f"[batch] method={method} reorder={reorder} graphs=24 n={len(off) - 1} m={len(nbr)}")
    with _abi.Graph(ctx, off, nbr, method=method, reorder=reorder) as g:
        g.set_batch(vbase)
        labels, res = g.run(10)
        iters = g.batch_iterations()
        check_batch(labels, iters, parts, vbase, 10)
        assert res["iterations"] == iters.max()
        labels2, _ = g.run(10)   # a second run starts from scratch
        assert np.array_equal(labels, labels2)


def test_batched_autotuned_captured(ctx):
    # the batched configuration of the benchmark: autotuned class kernels and
    # the captured loop; per-graph labels stay exact
    parts, vbase, off, nbr = make_batch(4, 20)
    say(f"[batch] autotuned captured graphs=20 n={len(off) - 1} m={len(nbr)}")
    with _abi.Graph(ctx, off, nbr, reorder=0, graph_capture=1) as g:
        g.set_batch(vbase)
        chosen = g.autotune(reps=2)
        assert len(chosen) == 12 and all(0 <= c <= 2 for c in chosen)
        for _ in range(2):
            labels, res = g.run(10)
            check_batch(labels, g.batch_iterations(), parts, vbase, 10)


def test_batch_rejects_crossing_edges(ctx):
    off = np.array([0, 1, 2, 3, 4], np.int64)
    nbr = np.array([1, 0, 3, 0], np.int32)   # vertex 3 -> 0 crosses graphs {0,1} | {2,3}
    with _abi.Graph(ctx, off, nbr) as g:
        with pytest.raises(_abi.GalpError) as e:
            g.set_batch([0, 2, 4])
        assert e.value.code == -1
        g.set_batch([0, 4])
        labels, _ = g.run(5)


@pytest.mark.parametrize("method", [_abi.METHOD_HASH, _abi.METHOD_SORT])
def test_captured_loop_replays(ctx, method):
    off, nbr = _abi.gen_rmat(15, 16, seed=11)
    final, per, counts = oracle.lp_run(off, nbr, niter=10)
    other_off, other_nbr = _abi.gen_planted_partition(4096, 64, 8, 2, 3)
    ofinal, _, _ = oracle.lp_run(other_off, other_nbr, niter=10)
    with _abi.Graph(ctx, off, nbr, method=method, graph_capture=1) as g, \
            _abi.Graph(ctx, other_off, other_nbr, method=method) as eager:
        for rep in range(4):
            labels, res = g.run(10)
            say(f"[capture] method={method} replay {rep}: {res['iterations']} iterations")
            assert res["iterations"] == len(counts), rep
            assert np.array_equal(labels, final), rep
            elab, _ = eager.run(10)          # eager work between the replays
            assert np.array_equal(elab, ofinal)
            st = g.step()                    # eager steps on the captured graph too
            assert st["iteration"] == len(counts) + 1
        labels5, _ = g.run(5)                # another niter: a new capture
        f5, _, _ = oracle.lp_run(off, nbr, niter=5)
        assert np.array_equal(labels5, f5)


def test_captured_batched_and_learned(ctx):
    parts, vbase, off, nbr = make_batch(2, 12)
    with _abi.Graph(ctx, off, nbr, graph_capture=1) as g:
        g.set_batch(vbase)
        for rep in range(3):
            labels, _ = g.run(10)
            check_batch(labels, g.batch_iterations(), parts, vbase, 10)
    off, nbr = _abi.gen_rmat(15, 16, seed=12)
    final, _, _ = oracle.lp_run(off, nbr, niter=10)
    p = np.zeros(_abi.mlp_num_params(8), np.float32)   # everything to the candidate path
    with _abi.Graph(ctx, off, nbr, dispatch=_abi.DISPATCH_LEARNED, graph_capture=1) as g:
        g.set_mlp(8, p, 4.0)
        for rep in range(3):
            labels, _ = g.run(10)
            assert np.array_equal(labels, final), rep
