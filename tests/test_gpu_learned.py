"""The learned dispatcher on the GPU, through the C-ABI.

- the MLP's HIP inference, loss and gradient against the oracle's float64
  forward/backward (itself pinned by finite differences in test_oracle.py);
- HIP training lowers the loss on a target the network can fit;
- the static features bit-exact against the oracle;
- the logged distinct-label counts exact against the oracle;
- labels bit-exact against the oracle after every iteration whatever the
  network predicts: everything to the candidate path (thresholds far above
  any count, so multi-label rows overflow and take the spill path), nothing,
  and seeded random networks in between, in core and out of core;
- what a network trained on the GPU gains over plain dispatch on RMAT-22
  (same labels, less device time);
- the judge's per-iteration choice between the network and the plain
  classes: labels exact under any mask, and under the autotuner's.
"""
import sys

import numpy as np
import pytest

import oracle
from galp_amd import _abi

pytestmark = pytest.mark.gpu

LEARNED = _abi.DISPATCH_LEARNED


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_x(n, seed):
    rng = np.random.default_rng(seed)
    x = np.empty((n, 4), np.float32)
    x[:, 0] = rng.uniform(6, 20, n)
    x[:, 1] = rng.uniform(0, 12, n)
    x[:, 2] = x[:, 1] + rng.uniform(0, 8, n)
    x[:, 3] = np.where(rng.random(n) < 0.2, -1.0, rng.uniform(0, 14, n))
    return x


@pytest.mark.parametrize("hidden", [8, 16, 32])
def test_mlp_predict_and_grad_match_oracle(ctx, hidden):
    say(f"[mlp] hidden={hidden}")
    p = _abi.mlp_init(hidden, 7 + hidden)
    x = rand_x(5000, hidden)
    y = (0.5 * x[:, 0] + 0.1 * x[:, 3]).astype(np.float32)
    yhat = _abi.mlp_predict(ctx, hidden, p, x)
    loss_ref, yhat_ref, grad_ref = oracle.mlp_loss_grad(hidden, p, x, y)
    # fp32 forward of <= 32 + 32*32 products of O(10) terms: 1e-4 relative
    np.testing.assert_allclose(yhat, yhat_ref, rtol=1e-4, atol=1e-4)
    loss, grad = _abi.mlp_loss_grad(ctx, hidden, p, x, y)
    assert abs(loss - loss_ref) <= 1e-4 * max(1.0, abs(loss_ref))
    # fp32 sums of 5000 per-sample terms: 1e-3 of the gradient's scale
    scale = np.abs(grad_ref).max()
    np.testing.assert_allclose(grad, grad_ref, rtol=1e-3, atol=1e-3 * scale)


def test_mlp_train_lowers_loss(ctx):
    hidden = 16
    x = rand_x(20000, 3)
    # a smooth target of the kind the dispatcher learns: distinct labels grow
    # with the degree and track the previous count when there is one
    y = np.where(x[:, 3] >= 0, 0.8 * x[:, 3] + 0.1 * x[:, 0], 0.9 * x[:, 0] - 0.2 * x[:, 1]).astype(np.float32)
    p0 = _abi.mlp_init(hidden, 11)
    loss0, _, _ = oracle.mlp_loss_grad(hidden, p0, x, y)
    p1, loss1 = _abi.mlp_train(ctx, hidden, p0, x, y, epochs=30, batch=256, lr=3e-3, seed=5)
    ref1, _, _ = oracle.mlp_loss_grad(hidden, p1, x, y)
    say(f"[mlp] train loss {loss0:.4f} -> {loss1:.4f} (oracle {ref1:.4f})")
    assert abs(loss1 - ref1) <= 1e-3 * max(1.0, ref1)
    assert ref1 < 0.05 * loss0
    p2, loss2 = _abi.mlp_train(ctx, hidden, p0, x, y, epochs=30, batch=256, lr=3e-3, seed=5)
    assert np.array_equal(p1, p2), "training must be deterministic for a seed"


def big_graph(seed=0):
    """Hubs (multi-tile vertices of every size up to 40k), every LDS class,
    and small vertices, with repeated neighbours."""
    rng = np.random.default_rng(seed)
    n = 60000
    rows = [[] for _ in range(n)]
    for h, d in ((0, 40000), (1, 9000), (2, 5000), (3, 3000), (4, 1500), (5, 1100), (6, 700), (7, 300)):
        for u in rng.choice(np.arange(8, n), size=d, replace=False):
            rows[h].append(int(u))
            rows[int(u)].append(h)
    for v in range(8, n):
        k = int(rng.integers(0, 12))
        for u in rng.integers(8, n, size=k):
            rows[v].append(int(u))
            rows[int(u)].append(v)
    for v in range(8, 2000):   # a few more W128..B1K vertices
        for u in rng.integers(8, n, size=int(rng.integers(60, 600))):
            rows[v].append(int(u))
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in rows])
    nbr = np.array([u for r in rows for u in r], np.int32)
    return off, nbr


@pytest.mark.parametrize("reorder", [1, 0])
def test_features_bit_exact(ctx, reorder):
    off, nbr = big_graph(1)
    d, mean, mx = oracle.vertex_features(off, nbr)
    big = np.diff(off) > 64
    with _abi.Graph(ctx, off, nbr, dispatch=LEARNED, reorder=reorder) as g:
        gd, gmean, gmx = g.features()
    assert np.array_equal(gd[big], d[big])
    assert np.array_equal(gmean[big], mean[big])
    assert np.array_equal(gmx[big], mx[big])
    assert not gd[~big].any() and not gmean[~big].any() and not gmx[~big].any()


@pytest.mark.parametrize("reorder", [1, 0])
def test_logged_distinct_counts_exact(ctx, reorder):
    off, nbr = big_graph(2)
    n = len(off) - 1
    big = np.nonzero(np.diff(off) > 64)[0]
    prev = np.arange(n, dtype=np.int32)
    with _abi.Graph(ctx, off, nbr, reorder=reorder) as g:   # plain dispatch logs too
        g.log_enable(10 * len(big))
        g.init()
        want_prev = None
        for it in range(3):
            dist = oracle.distinct_counts(off, nbr, prev)[big]
            g.step()
            x, y = g.log_read()
            assert len(x) == (it + 1) * len(big)
            xs, ys = x[it * len(big):], y[it * len(big):]
            got = np.rint(np.exp2(ys.astype(np.float64)) - 1).astype(np.int64)
            assert np.array_equal(np.sort(got), np.sort(dist)), it
            if want_prev is None:
                assert np.all(xs[:, 3] == -1.0)
            else:
                gotp = np.rint(np.exp2(xs[:, 3].astype(np.float64)) - 1).astype(np.int64)
                assert np.array_equal(np.sort(gotp), np.sort(want_prev)), it
            want_prev = dist
            prev, _ = oracle.lp_iterate(off, nbr, prev)


def run_learned(ctx, off, nbr, niter, tag, params, hidden, thresh, prev=None, reorder=1, mask=None, **opts):
    say(f"[learned] {tag} n={len(off) - 1} m={len(nbr)} thresh={thresh} reorder={reorder} {opts}")
    n = len(off) - 1
    prev = np.arange(n, dtype=np.int32) if prev is None else prev
    routed = overflowed = 0
    with _abi.Graph(ctx, off, nbr, dispatch=LEARNED, reorder=reorder, **opts) as g:
        g.set_mlp(hidden, params, thresh)
        if mask is not None:
            g.set_learned_iterations(mask)
        g.set_labels(prev)
        for it in range(niter):
            ref, cnt = oracle.lp_iterate(off, nbr, prev)
            st = g.step()
            got = g.labels()
            if not np.array_equal(got, ref):
                bad = np.nonzero(got != ref)[0]
                raise AssertionError(f"{tag} iteration {it + 1}: {len(bad)} labels differ, first at vertex "
                                     f"{bad[0]} (deg {off[bad[0] + 1] - off[bad[0]]}): gpu {got[bad[0]]} "
                                     f"oracle {ref[bad[0]]}")
            assert st["updated"] == cnt, (tag, it + 1, st["updated"], cnt)
            if mask is None or (mask >> min(it, 63)) & 1:   # routed by the network
                ds = g.dispatch_stats()
                routed += int(ds[_abi.DS_CAND_VERTS])
                overflowed += int(ds[_abi.DS_OVERFLOW])
                assert st["overflowed"] == ds[_abi.DS_OVERFLOW]
            else:   # the plain classes: no dispatch, no overflow
                assert st["overflowed"] == 0
            prev = ref
            if st["stopped"]:
                break
    return routed, overflowed


def const_net(hidden, value):
    """A network whose output is the constant `value` (all weights 0)."""
    p = np.zeros(_abi.mlp_num_params(hidden), np.float32)
    p[-1] = value
    return p


@pytest.mark.parametrize("reorder", [1, 0])
def test_learned_everything_to_candidates(ctx, reorder):
    off, nbr = big_graph(3)
    # predicted 0 distinct labels: every big vertex tries the candidate path;
    # rows with more than 8 labels overflow to the class kernels
    routed, overflowed = run_learned(ctx, off, nbr, 6, "all-cand", const_net(8, 0.0), 8, 4.0, reorder=reorder)
    assert routed > 0 and overflowed > 0


@pytest.mark.parametrize("reorder", [1, 0])
def test_learned_nothing_to_candidates(ctx, reorder):
    off, nbr = big_graph(4)
    routed, overflowed = run_learned(ctx, off, nbr, 4, "no-cand", const_net(8, 30.0), 8, 4.0, reorder=reorder)
    assert routed == 0 and overflowed == 0


def test_learned_clustered_labels_hit_candidates(ctx):
    # few distinct labels everywhere: the candidate path resolves hubs and
    # big vertices without overflow
    off, nbr = big_graph(5)
    n = len(off) - 1
    rng = np.random.default_rng(9)
    prev = rng.integers(0, 6, size=n).astype(np.int32)
    routed, overflowed = run_learned(ctx, off, nbr, 3, "clustered", const_net(16, 0.0), 16, 8.0, prev=prev)
    assert routed > 0 and overflowed == 0


@pytest.mark.parametrize("seed", [1, 2])
def test_learned_random_networks(ctx, seed):
    off, nbr = _abi.gen_rmat(16, 16, seed=seed)
    p = _abi.mlp_init(32, seed)
    run_learned(ctx, off, nbr, 8, f"rmat16-rand{seed}", p, 32, 3.0)


@pytest.mark.parametrize("mask", [0x5555555555555555, 0xAAAAAAAAAAAAAAAA, 0])
def test_learned_iteration_mask(ctx, mask):
    # iterations routed by the network and by the plain classes, alternating
    # (or never by the network): the labels stay the oracle's every iteration
    off, nbr = big_graph(9)
    run_learned(ctx, off, nbr, 7, f"mask{mask:x}", const_net(8, 1.0), 8, 4.0, mask=mask)
    off, nbr = _abi.gen_rmat(16, 16, seed=9)
    run_learned(ctx, off, nbr, 7, f"rmat16-mask{mask:x}", _abi.mlp_init(32, 9), 32, 6.0, mask=mask)


def test_judge_picks_routers_per_iteration(ctx):
    # the autotuner's judge: a mask over the iterations, and the labels of a
    # run under it (eager and captured) equal the oracle's
    off, nbr = _abi.gen_rmat(17, 16, seed=17)
    final, per, counts = oracle.lp_run(off, nbr, niter=10)
    x, y = None, None
    with _abi.Graph(ctx, off, nbr) as g:
        g.log_enable(11 * int((np.diff(off) > 64).sum()))
        g.run(10, want_labels=False)
        x, y = g.log_read()
    p, _ = _abi.mlp_train(ctx, 32, _abi.mlp_init(32, 1), x, y, epochs=2, batch=512, lr=3e-3, seed=2)
    for capture in (0, 1):
        with _abi.Graph(ctx, off, nbr, dispatch=LEARNED, graph_capture=capture) as g:
            g.set_mlp(32, p, 14.0)
            g.autotune(reps=1)
            mask = g.learned_iterations
            say(f"[learned] judge mask {mask & 0x3FF:010b} (capture {capture})")
            for _ in range(2):
                labels, res = g.run(10)
                assert res["iterations"] == len(per)
                assert np.array_equal(labels, final)


def test_learned_device_loop(ctx):
    off, nbr = _abi.gen_rmat(15, 16, seed=4)
    final, per, counts = oracle.lp_run(off, nbr, niter=10)
    with _abi.Graph(ctx, off, nbr, dispatch=LEARNED) as g:
        g.set_mlp(8, const_net(8, 1.0), 4.0)
        labels, res = g.run(10)
    assert res["iterations"] == len(per)
    assert np.array_equal(labels, final)


OOC = (_abi.POLICY_OOC_SYNC, _abi.POLICY_OOC_ASYNC)


@pytest.mark.parametrize("policy", OOC)
def test_learned_out_of_core_features(ctx, policy):
    # features computed from the streamed batches equal the in-core ones
    off, nbr = big_graph(6)
    d, mean, mx = oracle.vertex_features(off, nbr)
    big = np.diff(off) > 64
    with _abi.Graph(ctx, off, nbr, dispatch=LEARNED, policy=policy, buffer_pow=14) as g:
        gd, gmean, gmx = g.features()
    assert np.array_equal(gd[big], d[big])
    assert np.array_equal(gmean[big], mean[big])
    assert np.array_equal(gmx[big], mx[big])


@pytest.mark.parametrize("policy", OOC)
@pytest.mark.parametrize("net", ["all-cand", "no-cand", "random"])
def test_learned_out_of_core(ctx, policy, net):
    # per-batch dispatch: batches of < 2^14 edges (the 40k hub has a batch of
    # its own); rows of one tile can take the candidate path, longer rows
    # stay on the class and hub paths
    off, nbr = big_graph(7)
    if net == "all-cand":
        routed, overflowed = run_learned(ctx, off, nbr, 5, f"ooc{policy}-all-cand", const_net(8, 0.0), 8, 4.0,
                                         policy=policy, buffer_pow=14)
        assert routed > 0 and overflowed > 0
    elif net == "no-cand":
        routed, overflowed = run_learned(ctx, off, nbr, 4, f"ooc{policy}-no-cand", const_net(8, 30.0), 8, 4.0,
                                         policy=policy, buffer_pow=14)
        assert routed == 0 and overflowed == 0
    else:
        run_learned(ctx, off, nbr, 6, f"ooc{policy}-rand", _abi.mlp_init(32, 5), 32, 3.0, policy=policy,
                    buffer_pow=14)


@pytest.mark.parametrize("policy", OOC)
def test_learned_out_of_core_rmat_device_loop(ctx, policy):
    off, nbr = _abi.gen_rmat(15, 16, seed=6)
    final, per, counts = oracle.lp_run(off, nbr, niter=10)
    for bp in (12, 16):
        say(f"[learned] ooc rmat15 policy={policy} b={bp}")
        with _abi.Graph(ctx, off, nbr, dispatch=LEARNED, policy=policy, buffer_pow=bp) as g:
            g.set_mlp(8, const_net(8, 1.0), 4.0)
            labels, res = g.run(10)
        assert res["iterations"] == len(per)
        assert np.array_equal(labels, final), bp


def test_learned_out_of_core_needs_renumbering(ctx):
    off, nbr = big_graph(8)
    with pytest.raises(_abi.GalpError) as e:
        _abi.Graph(ctx, off, nbr, dispatch=LEARNED, policy=_abi.POLICY_OOC_SYNC, reorder=0)
    assert e.value.code == -4


def test_trained_network_gains_over_plain_dispatch(ctx):
    """What the trained network gains: it is trained on the GPU from a logged
    plain run of RMAT-22 (the benchmark graph; at smaller scales launch
    overheads hide the difference), predicts the distinct-label counts far
    better than a constant, and the learned dispatch it drives gives the
    oracle's labels in less device time than plain dispatch on the same
    graph."""
    off, nbr = _abi.gen_rmat(22, 16, seed=22)
    big = int((np.diff(off) > 64).sum())
    with _abi.Graph(ctx, off, nbr) as g:
        g.log_enable(11 * big)
        g.run(10, want_labels=False)
        x, y = g.log_read()
    p, loss = _abi.mlp_train(ctx, 32, _abi.mlp_init(32, 1), x, y, epochs=8, batch=512, lr=3e-3, seed=2)
    pred = _abi.mlp_predict(ctx, 32, p, x)
    mse, mse_const = float(np.mean((pred - y) ** 2)), float(np.var(y))
    say(f"[learned] RMAT-22: {len(x)} samples, network mse {mse:.4f} vs constant {mse_const:.4f}")
    assert mse < 0.1 * mse_const
    final, _, _ = oracle.lp_run(off, nbr, niter=10)
    times = {}
    for name, disp in (("plain", _abi.DISPATCH_PLAIN), ("learned", LEARNED)):
        with _abi.Graph(ctx, off, nbr, dispatch=disp) as g:
            if disp == LEARNED:
                g.set_mlp(32, p, 14.0)
            g.run(10, want_labels=False)   # warm-up
            ts = []
            for _ in range(5):
                labels, res = g.run(10)
                assert np.array_equal(labels, final), name
                ts.append(res["device_ms"])
            times[name] = float(np.median(ts))
    gain = times["plain"] / times["learned"]
    say(f"[learned] RMAT-22: plain {times['plain']:.3f} ms, learned {times['learned']:.3f} ms, gain {gain:.3f}x")
    assert gain > 1.05


def test_learned_needs_a_network(ctx):
    off = np.array([0, 1, 2], np.int64)
    with _abi.Graph(ctx, off, np.array([1, 0], np.int32), dispatch=LEARNED) as g:
        g.init()
        with pytest.raises(_abi.GalpError) as e:
            g.step()
        assert e.value.code == -6
        with pytest.raises(_abi.GalpError) as e:
            g.set_mlp(12, np.zeros(_abi.mlp_num_params(12), np.float32), 1.0)
        assert e.value.code == -1


@pytest.mark.parametrize("net", ["all-cand", "rand"])
def test_learned_rmat20_concurrent_classes(ctx, net):
    """Learned dispatch at RMAT scale 20 (m = 33.5M), where the lane-group and
    LDS classes run on side streams beside the candidate and hub paths:
    every iteration bit-exact against the oracle."""
    off, nbr = _abi.gen_rmat(20, 16, seed=21)
    assert len(nbr) >= 1 << 24
    if net == "all-cand":
        routed, _ = run_learned(ctx, off, nbr, 6, "rmat20-all-cand", const_net(8, 0.0), 8, 4.0)
        assert routed > 0
    else:
        run_learned(ctx, off, nbr, 6, "rmat20-rand", _abi.mlp_init(32, 7), 32, 3.0)


def test_learned_lfr_1m_per_iteration(ctx):
    """Learned dispatch on an LFR-style graph of 1M vertices with a network
    trained on the GPU from the graph's own logged run: every iteration
    bit-exact against the oracle."""
    off, nbr = _abi.gen_lfr(1_000_000, 2.5, 20.0, 2000, 1.5, 50, 50000, 0.3, seed=10)
    with _abi.Graph(ctx, off, nbr) as g:
        g.log_enable(11 * int((np.diff(off) > 64).sum()))
        g.run(10, want_labels=False)
        x, y = g.log_read()
    p, _ = _abi.mlp_train(ctx, 32, _abi.mlp_init(32, 1), x, y, epochs=2, batch=512, lr=3e-3, seed=2)
    run_learned(ctx, off, nbr, 10, "lfr1m-trained", p, 32, 14.0)


def test_learned_rmat24_trained_per_iteration(ctx):
    """Learned dispatch at RMAT scale 24 (m = 521M, hub rows of hundreds of
    candidate tiles merged by k_cand_merge) with a network trained on the
    GPU from the graph's own logged run: every iteration bit-exact against
    the oracle, and some multi-tile rows take the candidate path."""
    off, nbr = _abi.gen_rmat(24, 16, seed=24)
    with _abi.Graph(ctx, off, nbr) as g:
        g.log_enable(11 * int((np.diff(off) > 64).sum()))
        g.run(10, want_labels=False)
        x, y = g.log_read()
    p, _ = _abi.mlp_train(ctx, 32, _abi.mlp_init(32, 1), x, y, epochs=2, batch=512, lr=3e-3, seed=2)
    routed, _ = run_learned(ctx, off, nbr, 10, "rmat24-trained", p, 32, 14.0)
    assert routed > 0
