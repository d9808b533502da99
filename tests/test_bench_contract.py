"""bench.py's contract pieces that need no GPU: the refusal of more than one
GPU and the lookup of measured HBM traffic for the roofline line."""
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


def test_measured_traffic_lookup():
    with open(bench.TRAFFIC_FILE) as f:
        entries = json.load(f)["entries"]
    cand = [e for e in entries if e["workload"] == RMAT22 and e["variant"] == "hash_learned_tuned"
            and e["kernel"] == "cand"]
    assert len(cand) == 1
    got = bench.measured_traffic(RMAT22, "hash_learned_tuned", "cand")
    assert got == float(cand[0]["bytes_per_launch"])
    assert got > cand[0]["algorithmic_bytes_per_launch"] > 0   # traffic is at least the algorithmic bytes
    assert bench.measured_traffic(RMAT22, "hash_plain", "no-such-kernel") is None
    assert bench.measured_traffic("another workload", "hash_learned_tuned", "cand") is None
