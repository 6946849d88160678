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


def test_dagprof_joins_a_kernel_trace_with_the_plan(tmp_path):
    """The trace analysis sums per-kernel average durations, prices each
    kernel at its planned bytes and separates the kernels whose chunks ran
    again after their look-back (two kernels sharing generated code share
    their dispatches' average)."""
    import json
    from bitstreamzoo_amd import dagprof
    plan = {"bitlength": 64 * 1000, "n_chunks": 4, "rerun_chunks": [0, 3, 0],
            "kernels": [{"name": "bsz_k0", "entry": "bsz_fk_a", "inputs": 2, "outputs": 1},
                        {"name": "bsz_k1", "entry": "bsz_fk_b", "inputs": 1, "outputs": 1},
                        {"name": "bsz_k2", "entry": "bsz_fk_a", "inputs": 2, "outputs": 1}]}
    pj = tmp_path / "plan.json"
    pj.write_text(json.dumps(plan))
    head = "Kind,Kernel_Name,Start_Timestamp,End_Timestamp\n"
    rows = ["KERNEL_DISPATCH,bsz_fk_a,0,1000", "KERNEL_DISPATCH,bsz_fk_a,2000,5000",
            "KERNEL_DISPATCH,bsz_fk_b,6000,10000", "KERNEL_DISPATCH,k_other,0,99999"]
    tr = tmp_path / "trace.csv"
    tr.write_text(head + "\n".join(rows) + "\n")
    r = dagprof.analyze(str(tr), str(pj))
    sb = 8 * 1000
    assert r["kernels"] == 3
    assert r["kernel_ms_per_run"] == round((2000 + 4000 + 2000) / 1e6, 1)
    assert r["planned_bytes"] == (3 + 2 + 3) * sb
    assert r["kernels_with_reruns"]["count"] == 1 and r["kernels_with_reruns"]["chunks_rerun"] == 3
    assert r["kernels_without_reruns"]["GBps"] == round(6 * sb / 4000, 1)
    assert r["rate_GBps_quantiles"]["50"] == round(3 * sb / 2000, 1)


def test_balance_spreads_programs_over_streams():
    """Longest first, each to the least-loaded stream: the loads differ by
    at most the largest single item, and every item gets a stream."""
    import numpy as np
    from bitstreamzoo_amd import tuning
    rng = np.random.default_rng(3)
    costs = list(rng.lognormal(0, 1.5, size=257))
    for n in (1, 3, 8):
        b = tuning.balance(costs, n)
        assert len(b) == len(costs) and set(b) <= set(range(n))
        load = [sum(c for c, k in zip(costs, b) if k == j) for j in range(n)]
        assert max(load) - min(load) <= max(costs) + 1e-9
    assert tuning.balance([5.0, 1.0, 1.0, 1.0, 1.0, 1.0], 2) == [0, 1, 1, 1, 1, 1]


def test_stream_batch_forks_and_joins_every_stream():
    """The captured batch forks every other stream from the capturing one
    before any plan runs and joins each back at the end; `serial` chains the
    streams instead, each starting after the previous group.  Checked on
    stand-in contexts and plans that record the calls."""
    from bitstreamzoo_amd import tuning

    calls = []

    class FakePlan:
        def __init__(self, name):
            self.name = name

        def set_executor(self, e):
            pass

        def prepare(self):
            pass

        def run(self):
            calls.append(("run", self.name))

    class FakeGraph:
        def launch(self, n):
            pass

    class FakeCtx:
        def __init__(self, name):
            self.name = name

        def sync(self):
            pass

        def order_after(self, other):
            calls.append(("order", other.name, self.name))

        def capture(self, fn):
            calls.clear()
            fn()
            return FakeGraph()

        def timer_start(self):
            pass

        def timer_stop(self):
            return 6.0

    c0, c1, c2 = FakeCtx("c0"), FakeCtx("c1"), FakeCtx("c2")
    groups = [(c0, [FakePlan("p0")]), (c1, [FakePlan("p1"), FakePlan("p2")]), (c2, [FakePlan("p3")])]
    assert tuning.time_runs_streams(c0, groups, 3, 1) == 2.0
    assert calls == [("order", "c0", "c1"), ("order", "c0", "c2"), ("run", "p0"), ("run", "p1"), ("run", "p2"),
                     ("run", "p3"), ("order", "c1", "c0"), ("order", "c2", "c0")]
    tuning.time_runs_streams(c0, groups, 3, 1, serial=True)
    assert calls == [("run", "p0"), ("order", "c0", "c1"), ("run", "p1"), ("run", "p2"), ("order", "c1", "c2"),
                     ("run", "p3"), ("order", "c1", "c0"), ("order", "c2", "c0")]
