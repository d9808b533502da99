"""bench.py's contract pieces that need no GPU: the refusal of more than one
GPU and the lookup of measured HBM traffic for the roofline line, which only
returns a figure measured on the current kernel sources."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

RMAT22 = "RMAT scale 22, edge factor 16, a=0.57 b=0.19 c=0.19, seed 22"


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_measured_traffic_lookup(tmp_path, monkeypatch):
    f = tmp_path / "traffic.json"
    entry = {"workload": RMAT22, "variant": "hash_learned_tuned", "kernel": "cand", "bytes_per_launch": 900,
             "algorithmic_bytes_per_launch": 600, "source_hash": "abc"}
    f.write_text(json.dumps({"entries": [entry]}))
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(f))
    assert bench.measured_traffic(RMAT22, "hash_learned_tuned", "cand", "abc") == (900.0, False)
    # rescaled to a run whose launches carry twice the algorithmic bytes
    assert bench.measured_traffic(RMAT22, "hash_learned_tuned", "cand", "abc", 1200) == (1800.0, False)
    # measured on other kernel sources: withheld and flagged
    assert bench.measured_traffic(RMAT22, "hash_learned_tuned", "cand", "def") == (None, True)
    assert bench.measured_traffic(RMAT22, "hash_plain", "no-such-kernel", "abc") == (None, False)
    assert bench.measured_traffic("another workload", "hash_learned_tuned", "cand", "abc") == (None, False)


def test_committed_traffic_entries_carry_a_source_hash():
    with open(bench.TRAFFIC_FILE) as f:
        entries = json.load(f)["entries"]
    for e in entries:
        assert e["bytes_per_launch"] >= e["algorithmic_bytes_per_launch"] > 0
        assert len(e.get("source_hash", "")) == 16, e


def test_kernel_source_hash_is_stable():
    assert bench.kernel_source_hash() == bench.kernel_source_hash()
    assert len(bench.kernel_source_hash()) == 16
