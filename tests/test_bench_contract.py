"""bench.py's inputs that are committed files: the HBM-traffic summary it
reads for roofline.traffic must name the workload the bench runs (the
match is exact, and a mismatch silently reports traffic = null), and hold
both phases' byte counts, for each of its two workloads."""
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    sys.path.insert(0, ROOT)
    import bench
    return bench


@pytest.mark.parametrize("kind", ["c4", "c5"])
def test_traffic_summary_matches_the_bench_workload(kind):
    bench = _bench()
    from rajzel_amd import workloads as W
    sv = W.c4_gradient() if kind == "c4" else W.c5_gradient()
    workload = bench.workload_name(sv, kind)
    path = bench.traffic_path(sv)
    rec = json.load(open(path))
    assert rec["workload"] == workload
    for phase in ("forward", "adjoint"):
        b = bench.measured_traffic(path, phase, workload)
        assert b is not None and 1e10 < b < 1e13, (phase, b)
    # calibration of the counters by the 1 GiB copy
    assert 1.9 < rec["fetch_scale"] < 2.1 and 0.9 < rec["write_scale"] < 1.1


def test_gpus_above_one_exits_with_a_message():
    import subprocess
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode != 0
    assert "one GPU" in (out.stdout + out.stderr) or "1 GPU" in (out.stdout + out.stderr)
