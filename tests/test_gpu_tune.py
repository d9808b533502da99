"""The autotuner: every kernel configuration of every degree class gives the
oracle's labels; autotuning picks valid options, keeps the graph
initialised, and round-trips its choices through the cache file."""
import os
import sys

import numpy as np
import pytest

import oracle
from galp_amd import _abi
from test_gpu_parity import check_per_iteration, hub_graph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("option", [0, 1, 2])
@pytest.mark.parametrize("reorder", [1, 0])
@pytest.mark.parametrize("method", [_abi.METHOD_HASH, _abi.METHOD_SORT])
def test_every_option_same_labels(ctx, option, reorder, method):
    off, nbr = hub_graph(5)
    # extra rows of 600..4000 edges so B1K..B4K have vertices
    rng = np.random.default_rng(option)
    n = len(off) - 1
    rows = [nbr[off[v]:off[v + 1]].tolist() for v in range(n)]
    for v in range(10, 60):
        rows[v] += rng.integers(10, n, size=int(rng.integers(600, 4000))).tolist()
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in rows])
    nbr = np.array([u for r in rows for u in r], np.int32)
    print(f"[tune] method {method} option {option} reorder {reorder}", flush=True)
    with _abi.Graph(ctx, off, nbr, reorder=reorder, method=method) as g:
        g.set_tuning([option] * 12)
        g.init()
        prev = np.arange(n, dtype=np.int32)
        for it in range(4):
            ref, cnt = oracle.lp_iterate(off, nbr, prev)
            st = g.step()
            assert np.array_equal(g.labels(), ref), it
            assert st["updated"] == cnt
            prev = ref


@pytest.mark.parametrize("method", [_abi.METHOD_HASH, _abi.METHOD_SORT])
def test_autotune_and_cache(ctx, tmp_path, method):
    off, nbr = _abi.gen_rmat(16, 16, seed=21)
    final, per, counts = oracle.lp_run(off, nbr, niter=10)
    cache = str(tmp_path / "tune.txt")
    print(f"[tune] autotune method {method}", flush=True)
    with _abi.Graph(ctx, off, nbr, method=method) as g:
        chosen = g.autotune(reps=2, cache_path=cache)
        assert all(0 <= c <= 2 for c in chosen)
        labels, res = g.run(10)
        assert np.array_equal(labels, final)
    lines = open(cache).read().split("\n")
    assert len([x for x in lines if x.strip()]) >= 5
    with _abi.Graph(ctx, off, nbr, method=method) as g2:
        chosen2 = g2.autotune(reps=2, cache_path=cache)   # every class hits the cache
        assert np.array_equal(chosen, chosen2)
        labels, _ = g2.run(10)
        assert np.array_equal(labels, final)
    assert open(cache).read().split("\n") == lines      # nothing new appended


def test_autotune_learned_graph(ctx):
    off, nbr = _abi.gen_rmat(15, 16, seed=22)
    final, _, _ = oracle.lp_run(off, nbr, niter=10)
    with _abi.Graph(ctx, off, nbr, dispatch=_abi.DISPATCH_LEARNED) as g:
        g.set_mlp(8, np.zeros(_abi.mlp_num_params(8), np.float32), 4.0)
        g.autotune(reps=1)
        labels, _ = g.run(10)
        assert np.array_equal(labels, final)


def test_autotune_cache_keeps_methods_apart(ctx, tmp_path):
    # a cache written by one method is not reused by the other
    off, nbr = _abi.gen_rmat(14, 16, seed=23)
    cache = str(tmp_path / "tune.txt")
    with _abi.Graph(ctx, off, nbr, method=_abi.METHOD_HASH) as g:
        g.autotune(reps=1, cache_path=cache)
    n_hash = len([x for x in open(cache).read().split("\n") if x.strip()])
    with _abi.Graph(ctx, off, nbr, method=_abi.METHOD_SORT) as g:
        g.autotune(reps=1, cache_path=cache)
    lines = [x.split() for x in open(cache).read().split("\n") if x.strip()]
    assert len(lines) == 2 * n_hash
    assert {ln[1] for ln in lines} == {str(_abi.METHOD_HASH), str(_abi.METHOD_SORT)}


def test_autotune_refuses_other_policies(ctx):
    off, nbr = _abi.gen_rmat(12, 16, seed=24)
    with _abi.Graph(ctx, off, nbr, method=_abi.METHOD_SORT, policy=_abi.POLICY_LI) as g:
        with pytest.raises(_abi.GalpError) as e:
            g.autotune(reps=1)
        assert e.value.code == -4


@pytest.mark.parametrize("method", [_abi.METHOD_HASH, _abi.METHOD_SORT])
def test_autotune_times_streams(ctx, tmp_path, method):
    """The autotuner also picks one stream or three for whole iterations; either
    choice, from the cache or timed, gives the oracle's labels."""
    off, nbr = _abi.gen_rmat(16, 16, seed=25)
    final, _, _ = oracle.lp_run(off, nbr, niter=10)
    cache = str(tmp_path / "tune.txt")
    with _abi.Graph(ctx, off, nbr, method=method) as g:
        g.autotune(reps=2, cache_path=cache)
        assert g.tuned_streams in (1, 3)
        labels, _ = g.run(10)
        assert np.array_equal(labels, final)
    rows = [x.split() for x in open(cache).read().split("\n") if x.strip()]
    streams = [r for r in rows if r[2] == "12"]
    assert len(streams) == 1 and streams[0][5] in ("1", "3")
    # force the other choice through the cache: the labels do not change
    other = "3" if streams[0][5] == "1" else "1"
    with open(cache, "w") as f:
        f.write("\n".join(" ".join(r[:5] + [other]) if r[2] == "12" else " ".join(r) for r in rows) + "\n")
    with _abi.Graph(ctx, off, nbr, method=method) as g:
        g.autotune(reps=1, cache_path=cache)
        assert g.tuned_streams == int(other)
        labels, _ = g.run(10)
        assert np.array_equal(labels, final)


def test_autotune_keeps_profile_totals(ctx):
    """The autotuner times classes with per-kernel events but leaves the
    caller's profile totals as they were."""
    off, nbr = _abi.gen_rmat(15, 16, seed=26)
    final, _, _ = oracle.lp_run(off, nbr, niter=10)
    with _abi.Graph(ctx, off, nbr, profile=1) as g:
        g.run(10, want_labels=False)
        before = {k["name"]: (k["launches"], k["total_ms"]) for k in g.profile()}
        assert sum(v[0] for v in before.values()) > 0
        g.autotune(reps=1)
        after = {k["name"]: (k["launches"], k["total_ms"]) for k in g.profile()}
        assert after == before
        labels, _ = g.run(10)
        assert np.array_equal(labels, final)
