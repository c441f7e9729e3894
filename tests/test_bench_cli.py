"""bench.py's host-side contract: one GPU only, and the seeded batched-ensemble workload."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_ensemble_members_are_seeded():
    import bench

    a, b = bench.c4_members(), bench.c4_members()
    assert a == b and len(a) == 64
    assert all(m["Nxx"] in bench.C4_GRIDS for m in a)
    assert all(0.3 <= m["m1"] <= 0.7 and 0.3 <= m["m2"] <= 0.7 and 0.6 <= m["d"] <= 2.0 for m in a)
    assert len({m["Nxx"] for m in a}) == 4
