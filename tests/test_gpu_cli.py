"""The `galp [options] method policy data [test]` program (main.cu:80-207)
end to end on the GPU: every method/policy key the one-GPU build runs, the
-l variants and learned dispatch give the golden labels of the reference's
sample graph; the stderr line keeps the reference's tab-separated layout;
the keys outside the build exit 1 with a message."""
import json
import os
import subprocess
import sys

import pytest

from conftest import GOLDEN, ROOT

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


@pytest.mark.parametrize("args", [["0", "3"], ["1", "3"], ["1", "4"], ["-g", "2", "1", "0"], ["2", "1"]])
def test_keys_outside_the_build(tmp_path, args):
    p, _ = run(args + [SAMPLE], tmp_path)
    assert p.returncode == 1
    assert p.stdout.strip(), "a message says why"
