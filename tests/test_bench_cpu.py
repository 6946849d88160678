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


def test_cpu_baseline_reports_both_modes_and_the_host():
    """The CPU baseline runs the oracle's executor in the bsim-faithful and
    the best-CPU mode on the given workload and names the host, the threads
    and the committed oracle/reference time ratio."""
    sys.path.insert(0, ROOT)
    import bench
    from bitstreamzoo_amd import workloads
    w = workloads.adder32(1 << 16, seed=5)
    cb = bench.cpu_baseline(w)
    assert cb["kind"] == "port" and cb["unit"] == "Gbit-ops/s" and cb["value"] > 0
    assert cb["best_cpu"]["value"] > 0 and cb["cores"] >= 1 and cb["cpu_model"]
    assert cb["port_over_reference_time"]["faithful"] is not None
