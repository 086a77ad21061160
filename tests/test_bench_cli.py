"""The benchmark's contract on the CPU: more than one GPU is refused with a
message, and the algorithmic bytes per kernel are the formulas the roofline
figures divide by."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "--gpus 2 is not supported" in r.stderr


def test_algorithmic_bytes():
    sys.path.insert(0, ROOT)
    import bench

    n, u, o = 1 << 26, 4_875_439, 961_046
    b = bench.algorithmic_bytes(n, u, o)
    assert b["morton"] == 20 * n  # 16 B point in, 4 B code out
    assert b["binning"] == 8 * n  # one key read + one key write per pass
    assert b["histogram"] == 4 * n
    assert b["unique"] == 4 * n + 4 * u
    assert b["radix_tree"] == 4 * u + 15 * (u - 1)
    assert b["scan"] == 8 * (u - 1)
    assert b["octree"] == 21 * (u - 1) + 64 * o
    # The whole build as the bench line reports it: 36 B/key for the sort among them.
    build = b["morton"] + 4 * b["binning"] + b["unique"] + b["radix_tree"] + b["scan"] + b["octree"]
    assert build == 4_073_126_112
