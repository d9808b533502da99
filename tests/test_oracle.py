"""Pins the CPU oracle (oracle/lp_oracle.c) before anything is checked against it.

Each check is independent of the oracle's own code: hand-derived goldens, a
brute force of the closed form argmax (count, last position) written as a
different computation, properties the rule implies (renumbering
equivariance, degree-0 invariance) and finite differences for the MLP.
"""
import json
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def brute_next(off, nbr, prev):
    """Closed form of the tie rule: for each vertex, the label maximising
    (number of neighbours holding it, last CSR position holding it).
    Computed from an explicit table, not from a running scan."""
    n = len(off) - 1
    out = np.array(prev, dtype=np.int64).copy()
    for v in range(n):
        row = [int(prev[u]) for u in nbr[off[v]:off[v + 1]]]
        if not row:
            continue
        table = {}
        for pos, lab in enumerate(row):
            c, _ = table.get(lab, (0, -1))
            table[lab] = (c + 1, pos)
        out[v] = max(table.items(), key=lambda kv: kv[1])[0]
    return out


def first_wins_next(off, nbr, prev):
    """The `max_count < c` comparator (a plausible mistake): ties keep the
    label that reached the count first."""
    n = len(off) - 1
    out = np.array(prev, dtype=np.int64).copy()
    for v in range(n):
        cnt, best, bc = {}, None, 0
        for u in nbr[off[v]:off[v + 1]]:
            lab = int(prev[u])
            cnt[lab] = cnt.get(lab, 0) + 1
            if bc < cnt[lab]:
                bc, best = cnt[lab], lab
        if best is not None:
            out[v] = best
    return out


def random_graph(rng, n, p_edge, max_label=None):
    rows = []
    for v in range(n):
        d = rng.binomial(n, p_edge)
        rows.append(rng.integers(0, n, size=d))   # duplicates and any order allowed
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in rows])
    nbr = np.concatenate(rows).astype(np.int32) if off[-1] else np.zeros(0, np.int32)
    return off, nbr


def test_sample_golden_per_iteration():
    g = load("sample_labels.json")
    off = np.array(g["offsets"], np.int64)
    nbr = np.array(g["neighbors"], np.int32)
    labels = np.arange(len(off) - 1, dtype=np.int32)
    for it in g["iterations"]:
        labels, c = oracle.lp_iterate(off, nbr, labels)
        assert labels.tolist() == it["labels"]
        assert c == it["updated"]


def test_sample_golden_distinguishes_first_wins():
    # The hand golden fails the `<` comparator: it pins `<=` (hybrid_cpu_worker.cuh:183).
    g = load("sample_labels.json")
    off = np.array(g["offsets"], np.int64)
    nbr = np.array(g["neighbors"], np.int32)
    wrong = first_wins_next(off, nbr, np.arange(7))
    assert wrong.tolist() != g["iterations"][0]["labels"]


def test_micro_single_steps():
    for case in load("micro_cases.json")["single_step"]:
        off = np.array(case["offsets"], np.int64)
        nbr = np.array(case["neighbors"], np.int32)
        nxt, c = oracle.lp_iterate(off, nbr, np.array(case["prev"], np.int32))
        assert nxt.tolist() == case["next"], case["name"]
        assert c == case["updated"], case["name"]


def test_micro_runs_and_stop_rule():
    for case in load("micro_cases.json")["runs"]:
        off = np.array(case["offsets"], np.int64)
        nbr = np.array(case["neighbors"], np.int32)
        k = len(case["iterations"])
        niter = 10 if "stops_after" in case else k
        final, per, counts = oracle.lp_run(off, nbr, niter=niter)
        if "stops_after" in case:
            assert len(per) == case["stops_after"]
        for i, it in enumerate(case["iterations"]):
            assert per[i].tolist() == it["labels"], (case["name"], i)
            assert counts[i] == it["updated"], (case["name"], i)
        assert final.tolist() == per[-1].tolist()


@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_closed_form_on_random_graphs(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 60))
    off, nbr = random_graph(rng, n, float(rng.uniform(0.0, 0.3)))
    prev = rng.integers(0, max(1, n // 3), size=n).astype(np.int32)   # many ties
    nxt, c = oracle.lp_iterate(off, nbr, prev)
    ref = brute_next(off, nbr, prev)
    assert np.array_equal(nxt, ref)
    assert c == int(np.sum(ref != prev))


def test_oracle_threads_agree():
    rng = np.random.default_rng(11)
    off, nbr = random_graph(rng, 400, 0.05)
    prev = rng.integers(0, 50, size=400).astype(np.int32)
    a, ca = oracle.lp_iterate(off, nbr, prev, nthreads=1)
    b, cb = oracle.lp_iterate(off, nbr, prev, nthreads=4)
    assert np.array_equal(a, b) and ca == cb


@pytest.mark.parametrize("seed", range(3))
def test_renumbering_equivariance(seed):
    # labels'(pi v) = pi(labels v) when every row keeps its neighbour order.
    rng = np.random.default_rng(100 + seed)
    n = 80
    off, nbr = random_graph(rng, n, 0.08)
    pi = rng.permutation(n).astype(np.int32)
    inv = np.argsort(pi)
    deg = np.diff(off)
    off2 = np.zeros(n + 1, np.int64)
    off2[1:] = np.cumsum(deg[inv])
    nbr2 = np.concatenate([pi[nbr[off[inv[k]]:off[inv[k] + 1]]] for k in range(n)]).astype(np.int32)
    prev = rng.integers(0, 10, size=n).astype(np.int32)
    prev2 = np.empty(n, np.int32)
    prev2[pi] = pi[prev]
    a, _ = oracle.lp_iterate(off, nbr, prev)
    b, _ = oracle.lp_iterate(off2, nbr2, prev2)
    assert np.array_equal(b[pi], pi[a])


def test_degree_zero_keeps_label():
    off = np.array([0, 0, 1, 2, 2], np.int64)
    nbr = np.array([2, 1], np.int32)
    prev = np.array([3, 0, 1, 2], np.int32)
    nxt, c = oracle.lp_iterate(off, nbr, prev)
    assert nxt[0] == 3 and nxt[3] == 2
    assert nxt.tolist() == [3, 1, 0, 2] and c == 2


def test_all_distinct_row_takes_last_label():
    rng = np.random.default_rng(5)
    n = 300
    off = np.zeros(n + 1, np.int64)
    rows = [rng.choice(n, size=rng.integers(1, 30), replace=False) for _ in range(n)]
    off[1:] = np.cumsum([len(r) for r in rows])
    nbr = np.concatenate(rows).astype(np.int32)
    nxt, _ = oracle.lp_iterate(off, nbr, np.arange(n, dtype=np.int32))
    assert np.array_equal(nxt, np.array([r[-1] for r in rows]))


def test_distinct_counts_and_features():
    rng = np.random.default_rng(9)
    off, nbr = random_graph(rng, 120, 0.05)
    prev = rng.integers(0, 30, size=120).astype(np.int32)
    d = oracle.distinct_counts(off, nbr, prev)
    for v in range(120):
        assert d[v] == len(set(prev[nbr[off[v]:off[v + 1]]].tolist()))
    deg, mean, mx = oracle.vertex_features(off, nbr)
    degs = np.diff(off)
    assert np.array_equal(deg, degs.astype(np.float32))
    for v in range(120):
        nd = degs[nbr[off[v]:off[v + 1]]]
        if len(nd):
            assert mean[v] == np.float32(float(nd.sum()) / len(nd))
            assert mx[v] == nd.max()
        else:
            assert mean[v] == 0 and mx[v] == 0


def test_batch_boundaries():
    rng = np.random.default_rng(3)
    deg = rng.integers(0, 40, size=500)
    off = np.zeros(501, np.int64)
    off[1:] = np.cumsum(deg)
    for B in (16, 64, 100, 1000):
        bbs = oracle.batch_boundaries(off, B)
        assert bbs[0] == 0 and bbs[-1] == 500
        assert np.all(np.diff(bbs) > 0)
        for a, b in zip(bbs[:-1], bbs[1:]):
            e = off[b] - off[a]
            if e >= B:
                assert b - a == 1            # only a lone vertex of degree >= B
            elif b < 500:
                # greedy cut of outofcore.cuh:38-60: one more vertex reaches B
                assert off[b + 1] - off[a] >= B
    # a hub of degree >= B gets a batch of its own instead of hanging
    off = np.array([0, 3, 103, 106], np.int64)
    bbs = oracle.batch_boundaries(off, 10)
    assert bbs.tolist() == [0, 1, 2, 3]


def test_mlp_finite_differences():
    H = 8
    rng = np.random.default_rng(0)
    P = oracle.mlp_num_params(H)
    assert P == H * 4 + H + H * H + H + H + 1
    p = rng.normal(0, 0.5, size=P)
    x = rng.normal(0, 1, size=(16, 4))
    y = rng.normal(0, 1, size=16)
    loss, yhat, grad = oracle.mlp_loss_grad(H, p, x, y)
    assert np.isclose(loss, 0.5 * np.mean((yhat - y) ** 2))
    eps = 1e-6
    num = np.empty(P)
    for i in range(P):
        q = p.copy(); q[i] += eps
        lp, _, _ = oracle.mlp_loss_grad(H, q, x, y)
        q[i] -= 2 * eps
        lm, _, _ = oracle.mlp_loss_grad(H, q, x, y)
        num[i] = (lp - lm) / (2 * eps)
    assert np.allclose(grad, num, rtol=1e-5, atol=1e-8)


def test_mlp_forward_closed_form():
    # A network whose hidden layers are identities on positive inputs computes
    # w . x + b exactly.
    H = 4
    P = oracle.mlp_num_params(H)
    p = np.zeros(P)
    W1 = np.eye(H, 4)
    W2 = np.eye(H)
    w3 = np.array([1.0, 2.0, 3.0, 4.0])
    p[0:16] = W1.ravel()
    p[20:36] = W2.ravel()
    p[40:44] = w3
    p[44] = 0.5
    x = np.array([[1.0, 2.0, 3.0, 4.0], [0.5, 0.25, 0.0, 1.0]])
    _, yhat, _ = oracle.mlp_loss_grad(H, p, x)
    assert np.allclose(yhat, x @ w3 + 0.5)
