"""Host-only pieces that need no GPU: the galp_convert program (edge lists
and METIS files to .bin with both directions; utils/edgelist2bin.cpp,
utils/metis2bin.cpp), the serial CPU path behind the program's serial key
(main.cu:200-206) against the oracle, and the std::map-structured CPU
baseline (oracle/lp_baseline.cpp) against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, ROOT
from galp_amd import _abi

CONVERT = os.path.join(ROOT, "galp_amd", "bin", "galp_convert")
GALP = os.path.join(ROOT, "galp_amd", "bin", "galp")


def convert(tmp_path, text, *flags):
    src = tmp_path / "in.txt"
    src.write_text(text)
    out = tmp_path / "out.bin"
    r = subprocess.run([CONVERT, *flags, "-o", str(out), str(src)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return np.fromfile(out, dtype=np.int32).reshape(-1, 2), r.stdout


def test_convert_normalises_and_writes_both_directions(tmp_path):
    # SNAP comment, tab and space separators, ids 10/20/30/50 -> 0/1/2/3,
    # rows sorted (edgelist2bin.cpp:41-58), both directions written
    pairs, out = convert(tmp_path, "# comment\n30\t10\n10 20\n50 20\n")
    assert out.split() == ["4", "6"]
    assert pairs.tolist() == [[0, 1], [0, 2], [1, 0], [1, 3], [2, 0], [3, 1]]


def test_convert_plain_variant_keeps_ids_and_line_order(tmp_path):
    # --keep-ids --file-order: edgelist2bin.cpp:68-96 (row u gets v, row v
    # gets u, in line order; n = max id + 1)
    pairs, out = convert(tmp_path, "2 0\n0 1\n", "--keep-ids", "--file-order")
    assert out.split() == ["3", "4"]
    assert pairs.tolist() == [[0, 2], [0, 1], [1, 0], [2, 0]]


def test_convert_sample_matches_the_text_loader(tmp_path):
    with open(os.path.join(GOLDEN, "sample.txt")) as f:
        text = f.read()
    pairs, _ = convert(tmp_path, text, "--keep-ids", "--file-order")
    off, nbr = _abi.load_graph(os.path.join(GOLDEN, "sample.txt"))
    rows = np.repeat(np.arange(len(off) - 1), np.diff(off))
    assert np.array_equal(pairs[:, 0], rows) and np.array_equal(pairs[:, 1], nbr)


def test_convert_metis(tmp_path):
    # METIS: header, then the neighbours of vertex i (ids from 1) on line i;
    # vertex 3 is isolated and dropped, 4 -> 2 (metis2bin.cpp:19-44)
    pairs, out = convert(tmp_path, "4 2\n2 4\n1\n\n1\n", "--metis")
    assert out.split() == ["3", "4"]
    assert pairs.tolist() == [[0, 1], [0, 2], [1, 0], [2, 0]]


def test_convert_round_trip_through_the_loader(tmp_path):
    off, nbr = _abi.gen_planted_partition(500, 50, 4, 1, 3)
    lines = "".join(f"{v} {u}\n" for v in range(len(off) - 1) for u in nbr[off[v]:off[v + 1]] if v < u)
    src = tmp_path / "pp.txt"
    src.write_text(lines)
    out = tmp_path / "pp.bin"
    subprocess.run([CONVERT, "-o", str(out), str(src)], check=True, capture_output=True, timeout=60)
    off2, nbr2 = _abi.load_graph(str(out))
    # generated rows are sorted ascending with no gaps: the converted graph is identical
    assert np.array_equal(off, off2) and np.array_equal(nbr, nbr2)


def test_serial_key_runs_the_cpu_path(tmp_path):
    # method*5+policy = 11: the serial CPU path, no GPU needed
    labels = tmp_path / "l.txt"
    r = subprocess.run([GALP, "-i", "3", "--labels", str(labels), "2", "1", os.path.join(GOLDEN, "sample.txt")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "----------serial\t ----------" in r.stdout
    fields = r.stderr.strip().split("\n")[-1].split("\t")
    assert fields[:2] == ["serial", " "] and fields[2] == "sample.txt" and fields[4] == "0"
    import json
    with open(os.path.join(GOLDEN, "sample_labels.json")) as f:
        want = json.load(f)["iterations"][2]["labels"]
    assert [int(x) for x in labels.read_text().split()] == want


@pytest.mark.parametrize("graph", ["pp", "rmat", "lfr"])
def test_cpu_lp_equals_oracle(graph):
    if graph == "pp":
        off, nbr = _abi.gen_planted_partition(16384, 256, 8, 2, 1)
    elif graph == "rmat":
        off, nbr = _abi.gen_rmat(14, 16, 0.57, 0.19, 0.19, seed=14)
    else:
        off, nbr = _abi.gen_lfr(20000, 2.5, 20.0, 500, 1.5, 50, 2000, 0.3, seed=8)
    ref, per, counts = oracle.lp_run(off, nbr, niter=10, keep_iters=False)
    got, iters, _ = _abi.cpu_lp_run(off, nbr, 10)
    assert iters == len(counts)
    assert np.array_equal(got, ref)


def test_map_baseline_equals_oracle():
    off, nbr = _abi.gen_rmat(13, 16, 0.57, 0.19, 0.19, seed=13)
    n = len(off) - 1
    prev = np.arange(n, dtype=np.int32)
    for _ in range(4):
        ref, c = oracle.lp_iterate(off, nbr, prev)
        got, c2 = oracle.map_iterate(off, nbr, prev, nthreads=3)
        assert np.array_equal(got, ref) and c == c2
        vl = np.arange(1, n, 7, dtype=np.int64)     # a sample of vertices, as bench.py times
        part, _ = oracle.map_iterate(off, nbr, prev, vlist=vl, nthreads=1)
        assert np.array_equal(part[vl], ref[vl])
        prev = ref
