"""Host-side graph input of libgalp (no device calls): loaders, generators,
accuracy metrics, and the C-ABI's exported symbols."""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from galp_amd import _abi


def test_every_declared_symbol_is_exported():
    lib = _abi.lib()
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            with open(os.path.join(ROOT, "include", h)) as f:
                names |= set(re.findall(r"GALP_API\s+[\w\s\*]*?\b(galp_\w+)\s*\(", f.read()))
    assert len(names) >= 20
    for name in sorted(names):
        assert hasattr(lib, name), name


def test_txt_loader_matches_reference_rows():
    # graph.h:166-188: row u gets v and row v gets u, in line order.
    off, nbr = _abi.load_graph(os.path.join(GOLDEN, "sample.txt"))
    import json
    with open(os.path.join(GOLDEN, "sample_labels.json")) as f:
        g = json.load(f)
    assert off.tolist() == g["offsets"]
    assert nbr.tolist() == g["neighbors"]


def test_bin_roundtrip_and_rows(tmp_path):
    off, nbr = _abi.gen_planted_partition(2000, 100, 4, 1, 7)
    p = tmp_path / "g.bin"
    _abi.save_bin(str(p), off, nbr)
    raw = np.fromfile(p, dtype=np.int32).reshape(-1, 2)
    assert len(raw) == len(nbr)
    assert np.array_equal(raw[:, 1], nbr)
    off2, nbr2 = _abi.load_graph(str(p))
    assert np.array_equal(off, off2) and np.array_equal(nbr, nbr2)


def test_bin_keeps_ids_across_gaps(tmp_path):
    # vertex 1 has no edges: it keeps an empty row (the reference's loader
    # would shift vertex 2 into row 1).
    pairs = np.array([[0, 2], [2, 0]], np.int32)
    p = tmp_path / "gap.bin"
    pairs.tofile(p)
    off, nbr = _abi.load_graph(str(p))
    assert off.tolist() == [0, 1, 1, 2] and nbr.tolist() == [2, 0]


def test_bin_odd_length_rejected(tmp_path):
    p = tmp_path / "bad.bin"
    np.array([0, 1, 1], np.int32).tofile(p)
    with pytest.raises(_abi.GalpError) as e:
        _abi.load_graph(str(p))
    assert e.value.code == -5


def test_missing_file():
    with pytest.raises(_abi.GalpError) as e:
        _abi.load_graph("/nonexistent/graph.bin")
    assert e.value.code == -5


def check_simple_symmetric(off, nbr):
    n = len(off) - 1
    src = np.repeat(np.arange(n), np.diff(off))
    assert np.all(nbr >= 0) and np.all(nbr < n)
    assert not np.any(src == nbr), "self loop"
    for v in range(0, n, max(1, n // 200)):
        row = nbr[off[v]:off[v + 1]]
        assert np.all(np.diff(row) > 0), "rows sorted and unique"
    a = src.astype(np.int64) * n + nbr
    b = nbr.astype(np.int64) * n + src
    assert np.array_equal(np.sort(a), np.sort(b)), "symmetric"


def test_planted_partition_shape_and_determinism():
    off, nbr = _abi.gen_planted_partition(16384, 256, 8, 2, 1)
    off2, nbr2 = _abi.gen_planted_partition(16384, 256, 8, 2, 1)
    assert np.array_equal(off, off2) and np.array_equal(nbr, nbr2)
    check_simple_symmetric(off, nbr)
    n = 16384
    src = np.repeat(np.arange(n), np.diff(off))
    intra = np.mean(src // 256 == nbr // 256)
    assert 0.75 < intra < 0.85          # 16 intra vs 4 inter expected
    assert 14 < len(nbr) / n < 21       # ~16 + ~4, minus duplicates
    off3, _ = _abi.gen_planted_partition(16384, 256, 8, 2, 2)
    assert not np.array_equal(off, off3)


def test_batch_generator_is_the_disjoint_union():
    # every slice is the single-graph generator's output for its spec row
    count = 24
    off, nbr, vbase, spec = _abi.gen_batch_planted(count, n_range=(500, 3000), comm_range=(32, 256),
                                                   deg_range=(8, 48), inter_frac=0.25, seed=7)
    assert vbase[0] == 0 and vbase[-1] == len(off) - 1 and np.all(np.diff(vbase) == spec[:, 0])
    check_simple_symmetric(off, nbr)
    for k in range(count):
        n, cs, kin, kout, sd = (int(x) for x in spec[k])
        assert 500 <= n <= 3000 and 32 <= cs <= 256 and 4 <= kin <= 24 and kout >= 1
        o, nb = _abi.gen_planted_partition(n, cs, kin, kout, sd & ((1 << 64) - 1))
        a, b = int(vbase[k]), int(vbase[k + 1])
        assert np.array_equal(off[a:b + 1] - off[a], o), k
        assert np.array_equal(nbr[off[a]:off[b]] - a, nb), k   # no edge leaves its graph
    # degree ~ d + d/4 with d ~ U[8, 48] (35 on average), less the repeated
    # draws that small communities cause
    assert 18 < len(nbr) / (len(off) - 1) < 42
    off2, nbr2, _, _ = _abi.gen_batch_planted(count, n_range=(500, 3000), comm_range=(32, 256), seed=7)
    assert np.array_equal(off, off2) and np.array_equal(nbr, nbr2)


def test_rmat_shape():
    off, nbr = _abi.gen_rmat(12, 16, seed=5)
    n = 4096
    assert len(off) == n + 1
    check_simple_symmetric(off, nbr)
    deg = np.diff(off)
    assert deg.max() > 20 * deg.mean()  # heavy tail
    assert 0.4 * 2 * 16 * n < len(nbr) < 2 * 16 * n
    off2, nbr2 = _abi.gen_rmat(12, 16, seed=5)
    assert np.array_equal(nbr, nbr2)


def test_lfr_shape():
    off, nbr, comm = _abi.gen_lfr(20000, avg_deg=20.0, max_deg=500, cmin=20, cmax=2000, mu=0.3, seed=8,
                                  with_communities=True)
    check_simple_symmetric(off, nbr)
    deg = np.diff(off)
    assert 12 < deg.mean() < 24 and deg.max() <= 500
    n = 20000
    src = np.repeat(np.arange(n), deg)
    mix = np.mean(comm[src] != comm[nbr])
    assert 0.2 < mix < 0.45
    sizes = np.bincount(comm)
    assert sizes.min() >= 20


def test_accuracy_closed_forms(tmp_path):
    truth = tmp_path / "truth.txt"
    truth.write_text("0 1 2\n3 4 5\n")
    nmi, fm, ari = _abi.accuracy(str(truth), np.array([7, 7, 7, 9, 9, 9], np.int32))
    assert np.isclose(nmi, 1.0) and np.isclose(fm, 1.0) and np.isclose(ari, 1.0)
    # one big cluster: NMI 0, F = 2*3/(3+6) = 2/3, ARI 0
    nmi, fm, ari = _abi.accuracy(str(truth), np.zeros(6, np.int32))
    assert np.isclose(nmi, 0.0, atol=1e-12) and np.isclose(fm, 2 / 3) and np.isclose(ari, 0.0, atol=1e-12)
    # Hubert-Arabie ARI on a 2x2 contingency [[2,1],[1,2]]: sum C(nij,2)=2,
    # a=b=[3,3] -> 3+3, expected 6*6/15 = 2.4, ARI = (2-2.4)/(6-2.4) = -1/9
    nmi, fm, ari = _abi.accuracy(str(truth), np.array([0, 0, 1, 1, 1, 0], np.int32))
    assert np.isclose(ari, -1.0 / 9.0)
