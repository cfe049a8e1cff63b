#!/usr/bin/env python3
"""Throughput of the batched FATT gradient evaluation on one MI355X.

A step is one objective + gradient evaluation (the reference's grad_obj):
for every shot the Eikonal forward, residuals, the adjoint and the gradient
assembly, summed over shots.  Workload: 1001 x 3001 grid (DH = 5 m), 200
surface shots, 1500 receivers at row 2; current model = 12-layer model,
picks = times in the model with 20 Gaussian anomalies + N(0, 1 ms) noise
(synthetic, seeded).  Inputs are resident in HBM before the timed region.
``--workload c5``: the 2000 x 6000 grid (DH = 2.5 m), 400 shots, 3000
receivers, current model the 1-D gradient 1500 + 0.6 z, picks from it plus
40 Gaussian anomalies.

Printed: one JSON line with the shot-gradients/s of the batched launch,
the plain one-shot-per-launch form on the same GPU, the HBM roofline of the
dominant kernel (phase times from HIP events on the work stream) and the
reference CPU path (its own grad_obj, compiled from its sources) on the
host cores over a bounded sample of the same evaluation, and over the
200 x 600 / 120-shot evaluation in full.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "shot-gradients/s on 1001x3001 grid: batched vs per-shot GPU; % HBM roofline"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def model_dict(sv):
    return dict(nx=sv.nx, ny=sv.ny, dh=sv.dh, time=sv.time, ttnorm=sv.ttnorm, eps_adj=sv.eps_adj,
                src_ix=sv.src_ix, src_iy=sv.src_iy, rec_ix=sv.rec_ix, rec_iy=sv.rec_iy)


def subset(sv, n):
    from rajzel_amd import workloads as W
    s = W.Survey(name=sv.name, nx=sv.nx, ny=sv.ny, dh=sv.dh, vp=sv.vp, src_x_m=sv.src_x_m[:n],
                 src_z_m=sv.src_z_m[:n], rec_x_m=sv.rec_x_m, rec_z_m=sv.rec_z_m, time=sv.time,
                 ttnorm=sv.ttnorm, eps_adj=sv.eps_adj)
    s.tobs = None if sv.tobs is None else sv.tobs[:n]
    return s


def algorithmic_bytes(nx, ny, it_f, it_a):
    """Bytes per launch by phase: 12 B per cell per directional sweep (read S
    or T, read and write the swept field), 4 sweeps per outer iteration, 4 B
    per cell to initialise T and lambda, 8 B per cell per shot for the
    gradient (read lambda and S)."""
    cells = float(nx) * float(ny)
    fwd = cells * float(np.sum(48 * np.asarray(it_f, np.float64) + 4))
    adj = cells * float(np.sum(48 * np.asarray(it_a, np.float64) + 4))
    grd = cells * 8.0 * len(it_f)
    return fwd, adj, grd


def gauss_filter_timing(nx, ny, fsx=81, fsy=41, reps=5, cpu_rows=84):
    """The stock Gaussian gradient filter (FILT_SIZE_GRAD x FILT_SIZE_GRAD1 =
    81 x 41, gauss_filt.c) on an ny x nx field: GPU time per call of
    rz_gauss_filt on a device buffer (its weight and input copies and their
    allocation included), and the reference's own C filter on the first
    cpu_rows rows (1 thread), scaled to the grid.  Rows 0..cpu_rows-21 of
    that sample depend on the same inputs as in the full grid, so they are
    compared bit for bit with the GPU result."""
    import ctypes as C
    from rajzel_amd import native as N
    rng = np.random.default_rng(12)
    A = rng.normal(0.0, 1e-3, (ny, nx)).astype(np.float32)
    hip = C.CDLL("libamdhip64.so.7")
    n = C.c_size_t(A.nbytes)
    d = C.c_void_p()
    if hip.hipMalloc(C.byref(d), n) != 0:
        raise RuntimeError("hipMalloc failed")
    try:
        times = []
        for _ in range(reps + 1):
            if hip.hipMemcpy(d, A.ctypes.data_as(C.c_void_p), n, 1) != 0:
                raise RuntimeError("hipMemcpy failed")
            x, y = C.c_int(fsx), C.c_int(fsy)
            t0 = time.perf_counter()
            rc = N.lib().rz_gauss_filt(0, None, d, nx, ny, C.byref(x), C.byref(y), 1)
            times.append(time.perf_counter() - t0)
            if rc != N.RZ_OK:
                raise RuntimeError(N.lib().rz_last_error(None).decode())
        out = np.empty_like(A)
        if hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d, n, 2) != 0:
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(d)
    gpu_ms = float(np.median(times[1:])) * 1e3
    res = {"grid": f"{ny}x{nx}", "filter": f"{fsx}x{fsy}", "gpu_ms": gpu_ms,
           "mflop_per_call": 2.0 * nx * ny * fsx * fsy / 1e6}
    try:
        from oracle.bindings import Reference
        ref = Reference()
        sub = A[:cpu_rows].copy()
        ref.setup(nx, cpu_rows, 5.0)
        t0 = time.perf_counter()
        r, _, _ = ref.gauss_filt(sub, fsx, fsy)
        dt = time.perf_counter() - t0
        keep = cpu_rows - fsy // 2
        res.update(cpu_reference_ms_scaled=dt * 1e3 * ny / cpu_rows, cpu_sample_rows=cpu_rows,
                   bitwise_vs_reference_rows=bool(np.array_equal(r[:keep].view(np.uint32),
                                                                 out[:keep].view(np.uint32))))
    except OSError as e:
        log(f"reference CPU build unavailable for the filter ({e})")
    return res


def lbfgs_timing(nx, ny, iterate, nlbfgs=20, reps=3):
    """One LBFGS() call (LBFGS.c:12-226) with a full history of nlbfgs slots
    (the stock NLBFGS = 20) on an ny x nx grid: the GPU two-loop
    (rz_lbfgs_step: fields in, direction and the three history vectors out,
    all included) and the reference's own LBFGS() on one host core (its
    history files included, one float per fread/fwrite call), over the same
    iterates, with the directions compared bit for bit.  iterate(k) gives
    the gradient and model of iteration k (row-major float32 fields)."""
    import tempfile
    from rajzel_amd import native as N
    ptr = lambda it: 1 if it <= 1 else (it - 2) % nlbfgs + 1     # fatt.c:108-119
    its = [iterate(k) for k in range(1, nlbfgs + 2 + reps)]
    dev = N.LBFGS(nx, ny, nlbfgs)
    times, dirs = [], []
    for it, (g, vp) in enumerate(its, start=1):
        t0 = time.perf_counter()
        out = dev.step(it, ptr(it), g, -g, vp)
        if it > nlbfgs + 1:                                     # every slot filled
            times.append(time.perf_counter() - t0)
            dirs.append(out["hgrad"])
    dev.close()
    res = {"grid": f"{ny}x{nx}", "nlbfgs": nlbfgs, "dot_products_per_call": 3 * nlbfgs + 2,
           "calls_timed": len(times), "gpu_ms": float(np.median(times)) * 1e3}
    try:
        from oracle.bindings import Reference
        ref = Reference()
        ref.setup(nx, ny, 5.0)
        st = ref.lbfgs_state(nlbfgs)
        with tempfile.TemporaryDirectory(prefix="rzlb", dir="/tmp") as d:
            jac = os.path.join(d, "j")
            for it, (g, vp) in enumerate(its[:nlbfgs + 2], start=1):
                t0 = time.perf_counter()
                H = ref.lbfgs(st, jac, it, ptr(it), g, -g, vp)
                dt = time.perf_counter() - t0
        res["cpu_reference_ms"] = dt * 1e3
        res["cpu_cores"] = 1
        res["bitwise_vs_reference"] = bool(np.array_equal(H.view(np.uint32), dirs[0].view(np.uint32)))
    except OSError as e:
        log(f"reference CPU build unavailable for LBFGS ({e})")
    return res


def concurrency_scan(sv, rows_per_lane, counts=(1, 8, 64, 128, 200)):
    """Per-outer-iteration time of the forward and adjoint kernels with n
    shots in flight (n = 1 ... 200, one workgroup per shot): the difference
    of runs capped at 3 and at 1 outer iterations, halved.  Flat times mean
    a per-CU limit (instruction issue, latency); times growing with n mean a
    shared one (HBM, L2, fabric)."""
    from rajzel_amd import native as N
    rows = []
    for n in counts:
        res = {}
        for it in (1, 3):
            md = dict(model_dict(subset(sv, n)), max_iter=it)
            with N.Context([md], rows_per_lane=rows_per_lane) as ctx:
                ctx.set_slowness(0, sv.S)
                ctx.set_picks(0, np.full((n, sv.ntr), 0.5, np.float32))

                def run():
                    try:
                        ctx.gradient()
                    except N.NotConverged:
                        pass
                run()
                ctx.set_profiling(True)
                ctx.sync()
                for _ in range(2):
                    run()
                ctx.sync()
                res[it] = {k: v[0] / max(v[1], 1) for k, v in ctx.profile().items()}
                R = ctx.rows_per_lane
        rows.append({"shots": n, "rows_per_lane": R,
                     "forward_ms_per_iteration": (res[3]["forward"] - res[1]["forward"]) / 2,
                     "adjoint_ms_per_iteration": (res[3]["adjoint"] - res[1]["adjoint"]) / 2})
        log(f"[bench] concurrency {rows[-1]}")
    return rows


def copy_probe(mib):
    """One device-to-device hipMemcpy of ``mib`` MiB: the plain copy whose
    FETCH_SIZE / WRITE_SIZE readings calibrate the HBM-traffic counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")            # the runtime librajzel_amd.so already loaded
    n = C.c_size_t(mib << 20)
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), n) != 0:
            raise RuntimeError("hipMalloc failed in the copy probe")
    try:
        if hip.hipMemset(a, 1, n) != 0 or hip.hipMemcpy(b, a, n, 3) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("copy probe failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def workload_name(sv, kind):
    """The workload named in the bench line and in the traffic summaries
    (kind "c4": the 12-layer model, "c5": the 1-D gradient model)."""
    return (f"FATT gradient evaluation, {sv.ny}x{sv.nx} grid, DH={sv.dh} m, {sv.nshots} shots, "
            f"{sv.ntr} receivers, " + ("12-layer model" if kind == "c4" else "1-D gradient model"))


def traffic_path(sv):
    """The committed counter summary of this workload's evaluation."""
    return os.path.join(ROOT, "profiles", f"hbm_traffic_fatt_gradient_{sv.ny}x{sv.nx}_{sv.nshots}shots.json")


def measured_traffic(path, phase, workload):
    """HBM bytes of one evaluation's ``phase`` ("forward" or "adjoint": the
    one-workgroup and the split kernel of the phase together) from the
    committed counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy
    calibration), when it was taken on this same workload; None otherwise."""
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    k = rec.get("phases", {}).get(phase)
    if k is None or rec.get("workload") != workload:
        return None
    return float(k["hbm_bytes"])


def _cpu_worker(sv, shots, noise_rows, barrier, queue):
    """One process of the CPU baseline: the reference's forward in the true
    model for its shots plus the workload's noise gives the picks (untimed);
    then, after every process is ready, the reference's grad_obj of those
    shots (pick files written and read as the reference does) is timed."""
    from oracle.bindings import Reference
    ref = Reference()
    ref.setup(sv.nx, sv.ny, sv.dh, sv.time, sv.ttnorm, sv.eps_adj)
    St = W_slowness(sv.vp_true)
    tm = np.stack([ref.eikonal(St, float(sv.src_x_m[s]), float(sv.src_z_m[s]), sv.rec_ix, sv.rec_iy)[1]
                   for s in shots])
    tobs = (tm + noise_rows).astype(np.float32)
    barrier.wait()
    t0 = time.perf_counter()
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        out = ref.grad_obj(sv.S, sv.src_x_m[shots], sv.src_z_m[shots], sv.rec_ix, sv.rec_iy, tobs, workdir=d)
    queue.put((list(shots), time.perf_counter() - t0, out["tmod"], out["tres"], out["it_f"], out["it_a"]))


def W_slowness(vp):
    from rajzel_amd import workloads as W
    return W.slowness(vp)


def cpu_reference_rate(sv, shots, workers):
    """The reference's grad_obj (oracle/_ref, compiled from its sources) over
    `shots` of the workload, split into `workers` single-rank processes
    (consecutive shot ranges, as init_MPIshot.c does for MPI ranks).
    Returns (shot-gradients/s over the slowest process's wall time, the
    per-shot results).  Must run before this process touches the GPU: the
    workers are forked."""
    import multiprocessing as mp
    ctx = mp.get_context("fork")
    rng = np.random.default_rng(sv.meta.get("seed_noise", sv.meta.get("seed", 0) + 1000))
    noise = rng.normal(0.0, sv.noise_s, (sv.nshots, sv.ntr)) if sv.noise_s > 0 else np.zeros((sv.nshots, sv.ntr))
    parts = [p for p in np.array_split(np.asarray(shots), workers) if len(p)]
    barrier, queue = ctx.Barrier(len(parts)), ctx.Queue()
    procs = [ctx.Process(target=_cpu_worker, args=(sv, p, noise[p], barrier, queue)) for p in parts]
    for pr in procs:
        pr.start()
    res = [queue.get() for _ in procs]
    for pr in procs:
        pr.join()
    wall = max(r[1] for r in res)
    rows = {}
    for sh, _, tm, tr, itf, ita in res:
        for j, s in enumerate(sh):
            rows[s] = (tm[j], tr[j], int(itf[j]), int(ita[j]))
    return len(shots) / wall, wall, rows


def cpu_baseline(sv_c4, c4_shots, sv_c2, cores):
    """The reference C path on the host cores (BASELINE.md's plan): the C4
    sample on all cores and on one, and the C2 evaluation in full on all
    cores and on a 40-shot sample on one."""
    c4_all, c4_wall, rows = cpu_reference_rate(sv_c4, c4_shots, cores)
    c4_one, c4_one_wall, _ = cpu_reference_rate(sv_c4, c4_shots[:2], 1)
    c2_all, c2_wall, _ = cpu_reference_rate(sv_c2, list(range(sv_c2.nshots)), cores)
    c2_one, c2_one_wall, _ = cpu_reference_rate(sv_c2, list(range(40)), 1)
    return {
        "value": c4_all, "unit": "shot-gradients/s", "cores": cores, "kind": "reference",
        "sample": f"{len(c4_shots)} of the {sv_c4.nshots} shots of the same {sv_c4.ny}x{sv_c4.nx} evaluation "
                  f"({sv_c4.ntr} receivers), forward+adjoint+gradient, reference grad_obj in {cores} single-rank "
                  f"processes over consecutive shot ranges, {c4_wall:.1f} s",
        "one_core": {"value": c4_one, "shots": 2, "seconds": c4_one_wall},
        "c2_full": {"value": c2_all, "cores": cores, "shots": sv_c2.nshots, "seconds": c2_wall,
                    "one_core_value": c2_one, "one_core_shots": 40, "one_core_seconds": c2_one_wall},
    }, rows


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("c4", "c5"), default="c4",
                    help="c4: 1001 x 3001, DH 5 m, 200 shots (the default); c5: 2000 x 6000, DH 2.5 m, 400 shots")
    ap.add_argument("--shots", type=int, default=0, help="shots of the workload (default: all of it)")
    ap.add_argument("--plain-shots", type=int, default=10,
                    help="shots timed in the one-shot-per-launch form (rate is per shot)")
    ap.add_argument("--cpu-shots", type=int, default=0,
                    help="shots of the reference CPU sample (default: two per host core)")
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="processes of the reference CPU baseline (default: the host cores, at most 16)")
    ap.add_argument("--rows-per-lane", type=int, default=-1,
                    help="rows of a band per lane (1, 2, 4, 8); 0 = by band count, -1 = autotuned")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-precond", action="store_true", help="skip the Gaussian gradient filter timing")
    ap.add_argument("--no-lbfgs", action="store_true", help="skip the l-BFGS two-loop timing")
    ap.add_argument("--split-shots", type=int, default=None,
                    help="split shots per mixed launch (-1 = every free CU, 0 = none); default: tuned")
    ap.add_argument("--no-split", action="store_true",
                    help="one workgroup per shot in every launch (RAJZEL_AMD_SPLIT=0)")
    ap.add_argument("--concurrency-scan", action="store_true",
                    help="only print per-iteration kernel times versus shots in flight")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="finish with one device-to-device copy of MIB MiB (counter calibration)")
    ap.add_argument("--traffic", default=None,
                    help="counter summary giving the HBM bytes per launch of the dominant kernel "
                         "(default: profiles/hbm_traffic_fatt_gradient_<ny>x<nx>_<shots>shots.json)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the gradient path shards by shot, and a multi-GPU run "
              f"(--gpus {args.gpus}) is not part of this build", file=sys.stderr)
        sys.exit(2)

    # The one JSON line goes to this process's stdout; everything else that
    # writes to file descriptor 1 (the reference's C routines print there,
    # and so do the CPU baseline's processes) goes to stderr.
    sys.stdout.flush()
    result = os.fdopen(os.dup(1), "w")
    os.dup2(2, 1)
    if args.no_split:
        os.environ["RAJZEL_AMD_SPLIT"] = "0"
    from rajzel_amd import formats as F
    from rajzel_amd import native as N
    from rajzel_amd import workloads as W

    t_setup = time.perf_counter()
    make = W.c4_gradient if args.workload == "c4" else W.c5_gradient
    sv = make(nshots=args.shots) if args.shots else make()
    S = sv.S
    S_true = W.slowness(sv.vp_true)
    log(f"[bench] workload {sv.name}: {sv.nshots} shots, {sv.ntr} receivers "
        f"({time.perf_counter() - t_setup:.1f} s to generate)")
    cpu = None
    if not args.no_cpu and not args.concurrency_scan:
        # the reference on the host cores, before this process touches the GPU
        cores = args.cpu_cores or min(len(os.sched_getaffinity(0)), 16)
        nsample = min(args.cpu_shots or 2 * cores, sv.nshots)
        picks_ = np.linspace(0, sv.nshots - 1, nsample).round().astype(int).tolist()
        try:
            cpu = cpu_baseline(sv, picks_, W.c2_gradient(), cores)
            log(f"[bench] cpu baseline: {cpu[0]['value']:.3f} shot-gradients/s on {cores} cores "
                f"(one core {cpu[0]['one_core']['value']:.3f}); C2 {cpu[0]['c2_full']['value']:.1f}/s")
        except OSError as e:  # the reference build needs libmpi at load time
            log(f"reference CPU build unavailable ({e})")
    if args.concurrency_scan:
        print(json.dumps({"workload": f"{sv.ny}x{sv.nx} grid, {sv.ntr} receivers, picks 0.5 s",
                          "concurrency": concurrency_scan(sv, args.rows_per_lane)}), file=result, flush=True)
        return

    ctx = N.Context([model_dict(sv)], rows_per_lane=args.rows_per_lane)
    # picks: forward in the true model + seeded noise (setup, untimed)
    ctx.set_slowness(0, S_true)
    ctx.forward()
    rng = np.random.default_rng(sv.meta["seed"] + 1000)
    # as written to and read back from the pick files (printf %e, strtof)
    sv.tobs = F.roundtrip_picks(ctx.tmod(0) + rng.normal(0.0, sv.noise_s, (sv.nshots, sv.ntr)))
    ctx.set_slowness(0, S)
    ctx.set_picks(0, sv.tobs)
    # One evaluation as part of the setup: it gives the context the shots'
    # outer-iteration counts that schedule the split shots of later ones (in
    # an inversion every evaluation after the first has them).
    ctx.gradient()
    # shots per CU: how many of the slowest shots run split in the mixed
    # launches, timed on this evaluation (every candidate gives the same bits)
    if args.no_split or args.split_shots is not None:
        split_shots, tune_ms = (-1 if args.split_shots is None else args.split_shots), None
        ctx.set_split_shots(split_shots)
    else:
        split_shots, tune_ms = ctx.tune_split()
    log(f"[bench] setup done, R={ctx.rows_per_lane}, {ctx.device_bytes / 2**30:.2f} GiB on device, "
        f"split {ctx.split_state()}, split shots {split_shots} ({tune_ms} ms)")

    for _ in range(args.warmup):
        ctx.gradient()
    # All work runs on the context's stream; torch is not loaded here (it
    # would bring a second HIP runtime into the process), so the timed
    # region is bracketed by synchronising that stream.
    ctx.set_profiling(True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.gradient()
    ctx.sync()
    t1 = time.perf_counter()
    prof = ctx.profile()
    ctx.set_profiling(False)
    ms = (t1 - t0) * 1e3 / args.steps
    value = sv.nshots / (ms / 1e3)
    log(f"[bench] batched: {ms:.2f} ms/evaluation, {value:.1f} shot-gradients/s; phases {prof}")

    itf, ita = ctx.iterations(0)
    fwd_b, adj_b, grd_b = algorithmic_bytes(sv.nx, sv.ny, itf, ita)
    phase_ms = {k: v[0] / max(v[1], 1) for k, v in prof.items()}
    phase_bytes = {"forward": fwd_b, "adjoint": adj_b, "gradient": grd_b}
    dom = max(("forward", "adjoint"), key=lambda k: phase_ms[k])
    achieved = phase_bytes[dom] / (phase_ms[dom] * 1e-3) / 1e9
    eval_gbs = (fwd_b + adj_b + grd_b) / (ms * 1e-3) / 1e9
    workload = workload_name(sv, args.workload)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": measured_traffic(args.traffic or traffic_path(sv), dom, workload),
                "kernel": f"{dom} phase: k_{dom} (+ k_{dom}_split in mixed launches)",
                "bytes_per_launch": phase_bytes[dom], "avg_launch_ms": phase_ms[dom]}

    out = {
        "metric": METRIC if args.workload == "c4" else METRIC.replace("1001x3001", f"{sv.ny}x{sv.nx}"),
        "value": value, "unit": "shot-gradients/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload,
                   "nx": sv.nx, "ny": sv.ny, "shots": sv.nshots, "receivers": sv.ntr,
                   "rows_per_lane": ctx.rows_per_lane, "split_shots": split_shots},
        "roofline": roofline,
        "evaluation_hbm_frac": eval_gbs / HBM_PEAK_GBS,
        "phase_ms": phase_ms,
        "iterations": {"forward_mean": float(np.mean(itf)), "adjoint_mean": float(np.mean(ita)),
                       "forward_max": int(np.max(itf)), "adjoint_max": int(np.max(ita))},
        "stop_test_passes": {k: int(np.sum(v)) for k, v in ctx.stop_passes(0).items() if not k.startswith("sum")},
    }

    if not args.no_plain:
        n = min(args.plain_shots, sv.nshots)
        with N.Context([model_dict(subset(sv, n))], rows_per_lane=ctx.rows_per_lane) as cp:
            cp.set_slowness(0, S)
            cp.set_picks(0, sv.tobs[:n])
            cp.gradient(per_shot=True)
            cp.sync()
            p0 = time.perf_counter()
            cp.gradient(per_shot=True)
            cp.sync()
            p1 = time.perf_counter()
            plain = n / (p1 - p0)
            same = bool(np.array_equal(cp.tmod(0).view(np.uint32), ctx.tmod(0)[:n].view(np.uint32)))
        out["plain_per_shot"] = {"value": plain, "unit": "shot-gradients/s", "shots_timed": n,
                                 "tmod_bitwise_vs_batched": same}
        out["speedup_vs_plain"] = value / plain
        log(f"[bench] plain one-shot-per-launch: {plain:.1f} shot-gradients/s")

    if cpu is not None:
        cb, rows = cpu
        gpu_tmod, gpu_tres = ctx.tmod(0), ctx.tres(0)
        shots = sorted(rows)
        out["cpu_baseline"] = cb
        out["parity_sample"] = {
            "shots": len(shots),
            "tmod_bitwise": all(np.array_equal(rows[s][0].view(np.uint32), gpu_tmod[s].view(np.uint32))
                                for s in shots),
            "tres_bitwise": all(np.array_equal(rows[s][1].view(np.uint32), gpu_tres[s].view(np.uint32))
                                for s in shots),
            "iterations_equal": all(rows[s][2] == itf[s] and rows[s][3] == ita[s] for s in shots),
        }
        log(f"[bench] parity of the CPU sample: {out['parity_sample']}")
    rows_used = ctx.rows_per_lane
    ctx.close()
    if not args.no_precond:
        out["precond_gauss_filt"] = gauss_filter_timing(sv.nx, sv.ny)
        log(f"[bench] gauss filter: {out['precond_gauss_filt']}")
    if not args.no_lbfgs:
        # l-BFGS iterates: gradients of this workload at models moved along a
        # smooth perturbation (untimed evaluations)
        X, Z = np.meshgrid(np.arange(sv.nx), np.arange(sv.ny))
        bump = (np.sin(X / 97.0) * np.cos(Z / 61.0)).astype(np.float32)
        with N.Context([model_dict(sv)], rows_per_lane=rows_used) as cl:
            cl.set_picks(0, sv.tobs)

            def iterate(k):
                vp = (sv.vp + np.float32(3.0 * k) * bump).astype(np.float32)
                cl.set_slowness(0, W.slowness(vp))
                cl.gradient()
                return cl.grad(0), vp
            out["lbfgs"] = lbfgs_timing(sv.nx, sv.ny, iterate)
        log(f"[bench] lbfgs: {out['lbfgs']}")
    if args.copy_probe:
        copy_probe(args.copy_probe)
    print(json.dumps(out), file=result, flush=True)


if __name__ == "__main__":
    main()
