"""The `galp [options] method policy data [test]` program (main.cu:80-207)
end to end on the GPU: every method/policy key the one-GPU build runs, the
-l variants and learned dispatch give the golden labels of the reference's
sample graph; the stderr line keeps the reference's tab-separated layout;
the keys outside the build exit 1 with a message; a .bin written by
galp_convert runs through the program with the oracle's labels."""
import json
import os
import subprocess
import sys

import pytest

import numpy as np

import oracle
from conftest import GOLDEN, ROOT
from galp_amd import _abi

pytestmark = pytest.mark.gpu

GALP = os.path.join(ROOT, "galp_amd", "bin", "galp")
SAMPLE = os.path.join(GOLDEN, "sample.txt")


def run(args, tmp_path):
    out = tmp_path / "labels.txt"
    print(f"[cli] {' '.join(args)}", flush=True)
    p = subprocess.run([GALP, "--labels", str(out)] + args, capture_output=True, text=True, timeout=120)
    labels = [int(x) for x in out.read_text().split()] if p.returncode == 0 else None
    return p, labels


def golden(niter):
    with open(os.path.join(GOLDEN, "sample_labels.json")) as f:
        g = json.load(f)
    its = g["iterations"]
    return its[min(niter, len(its)) - 1]["labels"]


@pytest.mark.parametrize("key", [("0", "0"), ("0", "1"), ("0", "2"), ("0", "4"), ("1", "0"), ("1", "1"), ("1", "2")])
def test_every_key(tmp_path, key):
    p, labels = run(["-i", "3", "-b", "10", key[0], key[1], SAMPLE], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    assert labels == golden(3)
    assert "Data is ready on the host (7, 18)" in p.stdout
    fields = p.stderr.strip().split("\n")[-1].split("\t")
    assert len(fields) == 10, fields          # proc_name carries its own tab
    assert fields[2] == "sample.txt" and fields[3] == "3"
    assert fields[0] == ("dpp" if key[0] == "0" else "lfht1")


@pytest.mark.parametrize("variant", ["0", "2", "3"])
def test_hash_variants(tmp_path, variant):
    p, labels = run(["-i", "3", "-l", variant, "1", "0", SAMPLE], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    assert labels == golden(3)
    assert p.stderr.split("\t")[0] == "lfht" + variant


@pytest.mark.parametrize("policy", ["0", "1", "2"])
def test_learned_dispatch(tmp_path, policy):
    # in-core and both out-of-core policies (batches of < 2^10 edges)
    p, labels = run(["-i", "3", "-b", "10", "--dispatch", "learned", "1", policy, SAMPLE], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    assert labels == golden(3)


@pytest.mark.parametrize("args", [["0", "3"], ["1", "3"], ["1", "4"], ["-g", "2", "1", "0"], ["2", "0"]])
def test_keys_outside_the_build(tmp_path, args):
    p, _ = run(args + [SAMPLE], tmp_path)
    assert p.returncode == 1
    assert p.stdout.strip(), "a message says why"


@pytest.mark.parametrize("key", [("1", "0"), ("0", "0"), ("1", "2")])
def test_converted_bin_through_the_program(tmp_path, key):
    # an upper-triangle SNAP-style edge list with comments and an id gap ->
    # galp_convert -> .bin (both directions) -> galp, against the oracle on
    # the CSR the program loads
    pairs = []
    off, nbr = _abi.gen_planted_partition(3000, 60, 5, 1, 11)
    for v in range(len(off) - 1):
        for u in nbr[off[v]:off[v + 1]]:
            if v < u:
                pairs.append((2 * v, 2 * int(u)))   # even ids only: gaps the converter closes
    txt = tmp_path / "g.txt"
    with open(txt, "w") as f:
        f.write("# a comment line\n")
        for u, v in pairs:
            f.write(f"{u}\t{v}\n")
    binf = tmp_path / "g.bin"
    r = subprocess.run([os.path.join(ROOT, "galp_amd", "bin", "galp_convert"), "-o", str(binf), str(txt)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o2, n2 = _abi.load_graph(str(binf))
    assert len(o2) - 1 == len(off) - 1 and len(n2) == 2 * len(pairs)
    p, labels = run(["-i", "10", "-b", "12", key[0], key[1], str(binf)], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    ref, _, _ = oracle.lp_run(o2, n2, niter=10, keep_iters=False)
    assert np.array_equal(np.array(labels, np.int32), ref)
