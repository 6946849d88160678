"""bench.py's contract pieces that need no GPU."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_traffic_figure_comes_from_the_committed_profile():
    sys.path.insert(0, ROOT)
    import bench
    t = bench.profiled_traffic()
    # 98 streams of 2^28 bits, read or written once (+ popcount slots), as
    # measured: within 1 % of the algorithmic bytes
    alg = 98 * (1 << 28) // 8
    assert t is not None and abs(t / alg - 1) < 0.01
