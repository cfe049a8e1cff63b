"""bench.py's inputs that are committed files: the HBM-traffic summary it
reads for roofline.traffic must name the workload the bench runs (the
match is exact, and a mismatch silently reports traffic = null), and hold
both phases' byte counts."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TRAFFIC = os.path.join(ROOT, "profiles", "hbm_traffic_fatt_gradient_1001x3001_200shots.json")


def _bench():
    sys.path.insert(0, ROOT)
    import bench
    return bench


def test_traffic_summary_matches_the_bench_workload():
    bench = _bench()
    from rajzel_amd import workloads as W
    sv = W.c4_gradient()
    workload = (f"FATT gradient evaluation, {sv.ny}x{sv.nx} grid, DH={sv.dh} m, "
                f"{sv.nshots} shots, {sv.ntr} receivers, 12-layer model")
    rec = json.load(open(TRAFFIC))
    assert rec["workload"] == workload
    for phase in ("forward", "adjoint"):
        b = bench.measured_traffic(TRAFFIC, phase, workload)
        assert b is not None and 1e10 < b < 1e12, (phase, b)
    # calibration of the counters by the 1 GiB copy
    assert 1.9 < rec["fetch_scale"] < 2.1 and 0.9 < rec["write_scale"] < 1.1


def test_gpus_above_one_exits_with_a_message():
    import subprocess
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode != 0
    assert "one GPU" in (out.stdout + out.stderr) or "1 GPU" in (out.stdout + out.stderr)
