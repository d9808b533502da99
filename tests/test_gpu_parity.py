"""Parity of the HIP LP path against the CPU oracle, through the C-ABI.

Bit-exact labels and change counts after every iteration, for every method,
variant and option that changes which kernels run, on hand goldens, on
seeded synthetic graphs covering every degree class, and on the full-size
RMAT-22 benchmark graph.
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from galp_amd import _abi

pytestmark = pytest.mark.gpu

METHODS = [_abi.METHOD_HASH, _abi.METHOD_SORT]
REORDER = [1, 0]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def check_per_iteration(ctx, off, nbr, niter, tag, **opts):
    """Steps the GPU path and the oracle side by side from labels[i] = i."""
    say(f"[parity] {tag} n={len(off) - 1} m={len(nbr)} opts={opts}")
    prev = np.arange(len(off) - 1, dtype=np.int32)
    with _abi.Graph(ctx, off, nbr, **opts) as g:
        g.init()
        for it in range(niter):
            ref, cnt = oracle.lp_iterate(off, nbr, prev)
            st = g.step()
            got = g.labels()
            if not np.array_equal(got, ref):
                bad = np.nonzero(got != ref)[0]
                raise AssertionError(f"{tag} iteration {it + 1}: {len(bad)} labels differ, first at vertex "
                                     f"{bad[0]}: gpu {got[bad[0]]} oracle {ref[bad[0]]}")
            assert st["updated"] == cnt, (tag, it + 1, st["updated"], cnt)
            assert st["stopped"] == (cnt <= (len(off) - 1) * 1e-5)
            prev = ref
            if st["stopped"]:
                break
    return prev


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("reorder", REORDER)
def test_sample_golden(ctx, method, reorder):
    g = load("sample_labels.json")
    off = np.array(g["offsets"], np.int64)
    nbr = np.array(g["neighbors"], np.int32)
    with _abi.Graph(ctx, off, nbr, method=method, reorder=reorder) as dg:
        dg.init()
        for it in g["iterations"]:
            st = dg.step()
            assert dg.labels().tolist() == it["labels"]
            assert st["updated"] == it["updated"]


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("reorder", REORDER)
def test_micro_single_steps(ctx, method, reorder):
    for case in load("micro_cases.json")["single_step"]:
        off = np.array(case["offsets"], np.int64)
        nbr = np.array(case["neighbors"], np.int32)
        with _abi.Graph(ctx, off, nbr, method=method, reorder=reorder) as dg:
            dg.set_labels(np.array(case["prev"], np.int32))
            st = dg.step()
            assert dg.labels().tolist() == case["next"], case["name"]
            assert st["updated"] == case["updated"], case["name"]


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("reorder", REORDER)
def test_micro_runs_device_loop(ctx, method, reorder):
    for case in load("micro_cases.json")["runs"]:
        off = np.array(case["offsets"], np.int64)
        nbr = np.array(case["neighbors"], np.int32)
        k = case.get("stops_after", len(case["iterations"]))
        niter = 10 if "stops_after" in case else k
        with _abi.Graph(ctx, off, nbr, method=method, reorder=reorder) as dg:
            labels, res = dg.run(niter)
            assert res["iterations"] == k, case["name"]
            assert labels.tolist() == case["iterations"][k - 1]["labels"], case["name"]


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("reorder", REORDER)
def test_planted_partition_16k(ctx, method, reorder):
    off, nbr = _abi.gen_planted_partition(16384, 256, 8, 2, 1)
    check_per_iteration(ctx, off, nbr, 10, "pp16k", method=method, reorder=reorder)


@pytest.mark.parametrize("method", METHODS)
def test_rmat16(ctx, method):
    off, nbr = _abi.gen_rmat(16, 16, seed=16)
    check_per_iteration(ctx, off, nbr, 10, "rmat16", method=method)


def hub_graph(seed=0):
    """Hubs of degree 5k..40k (global-table tiles), every smaller class, and
    repeated labels so counts tie and differ."""
    rng = np.random.default_rng(seed)
    n = 60000
    rows = [[] for _ in range(n)]
    for h, d in ((0, 40000), (1, 9000), (2, 5000)):
        for u in rng.choice(np.arange(3, n), size=d, replace=False):
            rows[h].append(int(u))
            rows[int(u)].append(h)
    for v in range(3, n):
        k = int(rng.integers(0, 12))
        for u in rng.integers(3, n, size=k):
            rows[v].append(int(u))
            rows[int(u)].append(v)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in rows])
    nbr = np.array([u for r in rows for u in r], np.int32)
    return off, nbr


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("reorder", REORDER)
def test_hubs_and_duplicates(ctx, method, reorder):
    off, nbr = hub_graph()
    deg = np.diff(off)
    assert deg.max() > 4096 and (deg > 512).sum() >= 3
    check_per_iteration(ctx, off, nbr, 6, "hubs", method=method, reorder=reorder)


@pytest.mark.parametrize("method", METHODS)
def test_hub_from_clustered_labels(ctx, method):
    # A hub whose neighbours carry few distinct labels with many ties.
    off, nbr = hub_graph(1)
    n = len(off) - 1
    rng = np.random.default_rng(4)
    prev = rng.integers(0, 7, size=n).astype(np.int32)
    ref, cnt = oracle.lp_iterate(off, nbr, prev)
    with _abi.Graph(ctx, off, nbr, method=method) as g:
        g.set_labels(prev)
        st = g.step()
        assert np.array_equal(g.labels(), ref)
        assert st["updated"] == cnt


def partition0_labels(n, parts, count, rng):
    """`count` distinct vertex ids whose hub partition is 0 of `parts` (the
    multiplicative hash the hub path partitions labels by), so that one
    partition of a hub receives all of them."""
    ids = np.arange(n, dtype=np.uint64)
    h = (ids * np.uint64(0x85EBCA77)) & np.uint64(0xFFFFFFFF)
    ok = np.nonzero(((h * np.uint64(parts)) >> np.uint64(32)) == 0)[0]
    return rng.choice(ok, size=count, replace=False).astype(np.int32)


@pytest.mark.parametrize("method", METHODS)
def test_hub_partition_overflow(ctx, method):
    # A hub of 6000 edges whose neighbours carry 5000 distinct labels, all in
    # one label partition (of 3 for the hash method, of 12 for the sort
    # method), more than a merge table holds: the merges must take that
    # partition in rounds and still give the oracle's labels.
    n, deg = 100_000, 6000
    rng = np.random.default_rng(11)
    rows = [[] for _ in range(n)]
    for u in range(1, deg + 1):
        rows[0].append(u)
        rows[u].append(0)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in rows])
    nbr = np.array([u for r in rows for u in r], np.int32)
    pool = partition0_labels(n, 12, 5000, rng)
    prev = np.arange(n, dtype=np.int32)
    prev[1:5001] = pool
    prev[5001:deg + 1] = rng.choice(pool, size=deg - 5000)   # repeats: counts of 1 and 2
    ref, cnt = oracle.lp_iterate(off, nbr, prev)
    with _abi.Graph(ctx, off, nbr, method=method, reorder=0) as g:
        g.set_labels(prev)
        st = g.step()
        assert np.array_equal(g.labels(), ref)
        assert st["updated"] == cnt


@pytest.mark.parametrize("method", METHODS)
def test_hub_beyond_partition_cap(ctx, method):
    # A star of 17.5 M leaves: the hub's partitions are capped at 4096, so in
    # the first iteration (every label distinct) each holds ~4270 labels,
    # more than a merge table: merged in rounds.  Once refused at upload.
    leaves = 17_500_000
    n = leaves + 1
    off = np.empty(n + 1, np.int64)
    off[0], off[1] = 0, leaves
    off[2:] = leaves + np.arange(1, leaves + 1, dtype=np.int64)
    nbr = np.empty(2 * leaves, np.int32)
    nbr[:leaves] = np.arange(1, n, dtype=np.int32)
    nbr[leaves:] = 0
    check_per_iteration(ctx, off, nbr, 2, f"star{leaves}", method=method)


@pytest.mark.parametrize("method", METHODS)
def test_device_loop_equals_steps(ctx, method):
    off, nbr = _abi.gen_rmat(14, 16, seed=3)
    final, per, counts = oracle.lp_run(off, nbr, niter=10)
    with _abi.Graph(ctx, off, nbr, method=method) as g:
        labels, res = g.run(10)
    assert res["iterations"] == len(per)
    assert np.array_equal(labels, final)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("policy", [_abi.POLICY_OOC_SYNC, _abi.POLICY_OOC_ASYNC])
@pytest.mark.parametrize("reorder", REORDER)
def test_out_of_core(ctx, method, policy, reorder):
    """Policies 1 and 2 stream the neighbour array in batches of < 2^b edges;
    2^14 puts the 40000-edge hub in a batch of its own and cuts the rest into
    many batches.  Labels equal the oracle's (and so the in-core ones)."""
    off, nbr = hub_graph(3)
    check_per_iteration(ctx, off, nbr, 5, f"ooc-p{policy}", method=method, policy=policy, reorder=reorder,
                        buffer_pow=14)


@pytest.mark.parametrize("method", METHODS)
def test_out_of_core_batch_sizes(ctx, method):
    off, nbr = _abi.gen_rmat(15, 16, seed=7)
    final, per, counts = oracle.lp_run(off, nbr, niter=10)
    for b in (10, 13, 17, 24):
        for policy in (_abi.POLICY_OOC_SYNC, _abi.POLICY_OOC_ASYNC):
            with _abi.Graph(ctx, off, nbr, method=method, policy=policy, buffer_pow=b) as g:
                labels, res = g.run(10)
            assert res["iterations"] == len(per), (b, policy)
            assert np.array_equal(labels, final), (b, policy)


@pytest.mark.parametrize("variant", [0, 2, 3])
def test_hash_variants(ctx, variant):
    """-l 0 (no degree classes), 2 (rows > 512 on the tiled path) and 3
    (rows > 1024 tiled) change only which kernels run."""
    off, nbr = hub_graph(4)
    check_per_iteration(ctx, off, nbr, 4, f"l{variant}", method=_abi.METHOD_HASH, variant=variant)
    off, nbr = _abi.gen_rmat(14, 16, seed=9)
    check_per_iteration(ctx, off, nbr, 4, f"l{variant}-rmat14", method=_abi.METHOD_HASH, variant=variant, reorder=0)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("policy", [_abi.POLICY_INCORE, _abi.POLICY_OOC_ASYNC])
def test_int64_offsets(ctx, method, policy):
    """The int64-offset kernels (used when m >= 2^31) on graphs small enough
    for the oracle, by forcing int64 offsets."""
    off, nbr = hub_graph(6)
    check_per_iteration(ctx, off, nbr, 4, "off64-hubs", method=method, policy=policy, offsets64=1, buffer_pow=14)
    off, nbr = _abi.gen_rmat(15, 16, seed=13)
    check_per_iteration(ctx, off, nbr, 4, "off64-rmat15", method=method, policy=policy, offsets64=1, reorder=0)


def test_int64_offsets_learned_batched_captured(ctx):
    off, nbr = _abi.gen_rmat(15, 16, seed=14)
    final, per, counts = oracle.lp_run(off, nbr, niter=10)
    p = np.zeros(_abi.mlp_num_params(16), np.float32)
    with _abi.Graph(ctx, off, nbr, dispatch=_abi.DISPATCH_LEARNED, offsets64=1, graph_capture=1) as g:
        g.set_mlp(16, p, 6.0)
        for rep in range(2):
            labels, res = g.run(10)
            assert np.array_equal(labels, final), rep


def test_sort_li_policy(ctx):
    """Policy 4 (li): method 0 with every class up to 4096 in the
    block-per-vertex kernel gives the same labels."""
    off, nbr = hub_graph(2)
    check_per_iteration(ctx, off, nbr, 4, "li-hubs", method=_abi.METHOD_SORT, policy=_abi.POLICY_LI)
    off, nbr = _abi.gen_rmat(14, 16, seed=5)
    check_per_iteration(ctx, off, nbr, 4, "li-rmat14", method=_abi.METHOD_SORT, policy=_abi.POLICY_LI, reorder=0)


@pytest.mark.parametrize("method", METHODS)
def test_empty_and_edgeless(ctx, method):
    for off, nbr in ((np.zeros(1, np.int64), np.zeros(0, np.int32)),
                     (np.zeros(6, np.int64), np.zeros(0, np.int32))):
        with _abi.Graph(ctx, off, nbr, method=method) as g:
            labels, res = g.run(10)
            assert labels.tolist() == list(range(len(off) - 1))
            assert res["iterations"] == 1   # 0 changes <= n*1e-5 stops at once


def test_errors(ctx):
    off = np.array([0, 1, 2], np.int64)
    with pytest.raises(_abi.GalpError) as e:
        _abi.Graph(ctx, off, np.array([1, 5], np.int32))
    assert e.value.code == -1
    with pytest.raises(_abi.GalpError) as e:
        _abi.Graph(ctx, np.array([0, 2, 1], np.int64), np.array([1, 0], np.int32))
    assert e.value.code == -1
    ok_nbr = np.array([1, 0], np.int32)
    for policy, method in ((3, 1), (4, 1), (5, 1), (3, 0), (5, 0)):
        with pytest.raises(_abi.GalpError) as e:
            _abi.Graph(ctx, off, ok_nbr, policy=policy, method=method)
        assert e.value.code == -4
    with pytest.raises(_abi.GalpError) as e:
        _abi.Graph(ctx, off, ok_nbr, mem_budget=64)
    assert e.value.code == -2
    with _abi.Graph(ctx, off, ok_nbr) as g:
        with pytest.raises(_abi.GalpError) as e:
            g.step()
        assert e.value.code == -6
        with pytest.raises(_abi.GalpError) as e:
            g.set_labels(np.array([0, 2], np.int32))
        assert e.value.code == -1


@pytest.mark.parametrize("method", METHODS)
def test_rmat22_full_size(ctx, method):
    """The benchmark graph itself (RMAT scale 22, m = 128.3M): every iteration
    bit-exact against the oracle."""
    off, nbr = _abi.gen_rmat(22, 16, seed=22)
    check_per_iteration(ctx, off, nbr, 10, "rmat22", method=method)


@pytest.mark.parametrize("method", METHODS)
def test_rmat20_concurrent_classes(ctx, method):
    """RMAT scale 20 (m = 33.5M, above the size from which the class kernels of
    an iteration run on three streams): every iteration bit-exact against the
    oracle, and a captured 10-iteration run gives the same final labels."""
    off, nbr = _abi.gen_rmat(20, 16, seed=20)
    assert len(nbr) >= 1 << 24
    final = check_per_iteration(ctx, off, nbr, 10, "rmat20-streams", method=method)
    with _abi.Graph(ctx, off, nbr, method=method, graph_capture=1) as g:
        for _ in range(3):   # capture, then replays
            labels, _ = g.run(10)
            assert np.array_equal(labels, final)


def test_gather_probe_checksum(ctx):
    """The gather probe reads every neighbour's current label: with the input
    numbering kept, its checksum is the sum of labels[nbr] (labels[i] = i
    after init, then the labels of a few iterations)."""
    off, nbr = _abi.gen_rmat(16, 16, seed=31)
    with _abi.Graph(ctx, off, nbr, reorder=0) as g:
        g.init()
        ms, cs = g.gather_probe(reps=2)
        assert ms > 0 and cs == int(nbr.astype(np.uint64).sum())
        for _ in range(3):
            g.step()
        lab = g.labels()
        _, cs = g.gather_probe(reps=1)
        assert cs == int(lab[nbr].astype(np.uint64).sum())


@pytest.mark.parametrize("method", METHODS)
def test_lfr_1m_per_iteration(ctx, method):
    """An LFR-style graph (1M vertices, power-law degrees up to 2000 and
    communities of 50..50000, mixing 0.3): every iteration bit-exact against
    the oracle.  Most of its edges sit in rows of 65..2000, the LDS and
    register-sort classes, with many distinct labels per row."""
    off, nbr = _abi.gen_lfr(1_000_000, 2.5, 20.0, 2000, 1.5, 50, 50000, 0.3, seed=9)
    check_per_iteration(ctx, off, nbr, 10, "lfr1m", method=method)
