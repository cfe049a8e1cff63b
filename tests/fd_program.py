"""Workspaces for the reference program (oracle/_ref/librajzel_cpu.so) and
the same program on the MI355X path (oracle/_ref/librajzel_gpu.so, the
fd.h entry points of rajzel_amd/csrc/rz_fd_surface.c), run by
oracle/_ref/prog_launcher as ``prog_launcher <lib> <par file> <workflow>``
from inside the workspace.  All paths in the parameter file are relative
and short: the reference keeps them in 74-byte buffers (fd.h:25).
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np

from rajzel_amd import formats as F
from rajzel_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
LAUNCHER = os.path.join(REF, "prog_launcher")
CPU_LIB = os.path.join(REF, "librajzel_cpu.so")
GPU_LIB = os.path.join(REF, "librajzel_gpu.so")


def available(gpu=False):
    libs = [LAUNCHER, CPU_LIB] + ([GPU_LIB] if gpu else [])
    return all(os.path.exists(p) for p in libs)


def par_values(nx, ny, dh, invmat, **kw):
    v = dict(INVMAT=invmat, NX=nx, NY=ny, DH=dh, TIME=6.0, SOURCE_FILE="src.dat", READMOD=1,
             MFILE="model/start", SNAP=0, IDX=1, IDY=1, SNAP_FORMAT=3, SNAP_FILE="tt/snap", SEISMO=1,
             REC_FILE="rec.dat", REFREC1=0.0, REFREC2=0.0, N_STREAMER=0, REC_INCR_X=0.0, REC_INCR_Y=0.0,
             PICKS_FILE="picks/mod", TTNORM=1e-10, LOG_FILE="log/rz.log", LOG=0, ITERMAX=3, EPS_ADJ=1e-9,
             JACOBIAN="jac/grad", DATA_DIR="picks/obs", TAPER=0, TAPERLENGTH=4, INVTYPE=2, GRADT1=1,
             GRADT2=2, GRADT3=nx - 1, GRADT4=nx, SWS_TAPER_GRAD_VERT=0, SWS_TAPER_GRAD_HOR=0,
             EXP_TAPER_GRAD_HOR=0.0, SWS_TAPER_GRAD_SOURCES=0, SWS_TAPER_CIRCULAR_PER_SHOT=0, SRTSHAPE=1,
             SRTRADIUS=20.0, FILTSIZE=1, SWS_TAPER_FILE=0, INV_MODELFILE="model/inv", nfstart=1, nf=1,
             nfstart_jac=1000, nf_jac=1, VPUPPERLIM=6000.0, VPLOWERLIM=300.0, GRAD_METHOD=2, NLBFGS=5,
             MODEL_FILTER=0, FILT_SIZE=5, LINESEARCH=1, EPS_SCALE=0.01, C1=1e-4, C2=0.9, STEPMAX=6,
             SCALEFAC=2.0, MISFIT_LOG_FILE="misfit.dat", MIN_ITER=0, GRAD_FILTER=0, FILT_SIZE_GRAD=0,
             VP0_1=1500.0, VP0_2=1600.0, DVP0=50.0, GRAD0_1=0.0, GRAD0_2=1.0, DGRAD0=0.5,
             GRIDSEARCH_FILE="gs.dat")
    v.update(kw)
    return v


def make_workspace(sv, vp_model, invmat, picks=None, spatfilter=0, stages=None, **kw):
    """A fresh directory with parameter, workflow, acquisition and model
    files (and pick files <DATA_DIR>_<shot>.dat when ``picks`` is given).
    ``stages``: workflow rows (dicts of read_par_inv.c's six columns), by
    default one stage with PRO = 0 and the given SPATFILTER."""
    ws = tempfile.mkdtemp(prefix="rzp", dir="/tmp")
    for d in ("model", "picks", "jac", "log", "tt"):
        os.makedirs(os.path.join(ws, d))
    F.write_par(os.path.join(ws, "rz.inp"), par_values(sv.nx, sv.ny, sv.dh, invmat, **kw))
    F.write_workflow(os.path.join(ws, "wf.inp"),
                     stages or [dict(PRO=0.0, SPATFILTER=spatfilter, WD_DAMP=1e-3, WD_DAMP1=1e-3,
                                     FILT_SIZE_GRAD=8, FILT_SIZE_GRAD1=6)])
    F.write_sources(os.path.join(ws, "src.dat"), sv.src_x_m, sv.src_z_m)
    F.write_receivers(os.path.join(ws, "rec.dat"), sv.rec_x_m, sv.rec_z_m)
    F.write_model(os.path.join(ws, "model", "start.vp"), vp_model)
    if picks is not None:
        for s in range(picks.shape[0]):
            F.write_picks(os.path.join(ws, "picks", f"obs_{s + 1}.dat"), picks[s])
    return ws


MPIEXEC = next((p for p in ("/opt/conda/bin/mpiexec", "/usr/bin/mpiexec") if os.path.exists(p)), None)


def run(ws, lib, timeout=300, env=None, ranks=1, log=None):
    """Run the program in ``ws`` (``env``: extra environment variables) as
    one process, or as ``ranks`` MPI ranks under mpiexec; returns the
    CompletedProcess.  With ``log`` (a file path) the program writes its
    output there as it runs, and stdout and stderr of the result both hold
    that file's text."""
    env = dict(os.environ, **(env or {}))
    cmd = [LAUNCHER, lib, "rz.inp", "wf.inp"]
    if ranks > 1:
        cmd = [MPIEXEC, "-n", str(ranks)] + cmd
    if log is None:
        return subprocess.run(cmd, cwd=ws, env=env, capture_output=True, text=True, timeout=timeout)
    with open(log, "w") as fh:
        r = subprocess.run(cmd, cwd=ws, env=env, stdout=fh, stderr=subprocess.STDOUT, timeout=timeout)
    with open(log) as fh:
        text = fh.read()
    return subprocess.CompletedProcess(cmd, r.returncode, text, text)


def outputs(ws):
    """Every output file of the run (path -> bytes), logs excluded: they
    carry wall-clock times."""
    out = {}
    for base, _, files in os.walk(ws):
        rel = os.path.relpath(base, ws)
        if rel.split(os.sep)[0] == "log":
            continue
        for f in files:
            p = os.path.join(base, f)
            with open(p, "rb") as fh:
                out[os.path.normpath(os.path.join(rel, f))] = fh.read()
    return out


def remove(ws):
    if ws and ws.startswith("/tmp/rzp"):
        shutil.rmtree(ws, ignore_errors=True)


def survey(nx=72, ny=40, nshots=5, seed=21):
    """Small acquisition: sources at 10 m depth, receivers on row 2."""
    sv = W.small_case(nx=nx, ny=ny, nshots=nshots, seed=seed, rec_row=2)
    return sv


def start_model(sv):
    """1-D gradient start model 1500 + 0.5 z."""
    z = (np.arange(sv.ny, dtype=np.float32) + 1) * np.float32(sv.dh)
    return np.repeat((1500.0 + 0.5 * z)[:, None], sv.nx, axis=1).astype(np.float32)


def time_grid_search(nx=600, ny=200, dh=10.0, nshots=40, seed=31):
    """Wall time of the grid-search mode (INVMAT=2, grid_search.c) of the
    reference program and of the MI355X program on identical workspaces:
    a 200 x 600 grid, 40 shots, 600 receivers on row 2, a 4 x 4 grid of
    (vp0, grad0) models; the MI355X program also with every pick file
    parsed on every read (RAJZEL_AMD_PICK_CACHE=0), as the reference does.
    Returns a dict with the times and whether the grid-search tables are
    byte-identical."""
    import time
    from oracle.bindings import Restatement
    sv = W.small_case(nx=nx, ny=ny, dh=dh, nshots=nshots, seed=seed, rec_row=2)
    S = (1.0 / sv.vp_true.astype(np.float64)).astype(np.float32)
    tm, _, _ = Restatement().forward(S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.dh)
    obs = F.roundtrip_picks(tm + np.random.default_rng(3).normal(0, 1e-3, tm.shape))
    grid = dict(VP0_1=1400.0, VP0_2=1600.0, DVP0=50.0, GRAD0_1=0.2, GRAD0_2=1.0, DGRAD0=0.2)
    res = {"grid": f"{ny}x{nx}", "shots": nshots, "receivers": len(sv.rec_x_m), "models": 16}
    spaces = []
    try:
        # the last run parses every pick file on every read, as the reference does
        for key, lib, flag in (("cpu_reference_s", CPU_LIB, "1"), ("gpu_s", GPU_LIB, "1"),
                               ("gpu_parse_every_read_s", GPU_LIB, "0")):
            ws = make_workspace(sv, start_model(sv), invmat=2, picks=obs, **grid)
            spaces.append(ws)
            t0 = time.perf_counter()
            r = run(ws, lib, timeout=600, env={"RAJZEL_AMD_PICK_CACHE": flag})
            res[key] = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        a, b, c = (outputs(ws)["gs.dat"] for ws in spaces)
        res["tables_identical"] = a == b == c
        res["speedup"] = res["cpu_reference_s"] / res["gpu_s"]
    finally:
        for ws in spaces:
            remove(ws)
    return res


def time_fatt_picks(itermax=2):
    """Wall time of the MI355X program's FATT inversion (INVMAT=1, L-BFGS
    with the Wolfe line search) on the 1001 x 3001 workload of 200 shots and
    1500 receivers, with the pick files parsed once per change (default)
    and on every evaluation as the reference does (RAJZEL_AMD_PICK_CACHE=0).
    Picks: the forward in the true model plus N(0, 1 ms), written as pick
    files.  Returns both times, the number of evaluations (grad_obj calls,
    counted from the program's output) and whether every output file of the
    two runs is byte-identical."""
    import time
    from rajzel_amd import native as N
    sv = W.c4_gradient()
    with N.Context([dict(nx=sv.nx, ny=sv.ny, dh=sv.dh, time=sv.time, ttnorm=sv.ttnorm, eps_adj=sv.eps_adj,
                         src_ix=sv.src_ix, src_iy=sv.src_iy, rec_ix=sv.rec_ix, rec_iy=sv.rec_iy)]) as ctx:
        ctx.set_slowness(0, W.slowness(sv.vp_true))
        ctx.forward()
        tm = ctx.tmod(0)
    obs = F.roundtrip_picks(tm + np.random.default_rng(sv.meta["seed"] + 1000).normal(0.0, sv.noise_s, tm.shape))
    res = {"grid": f"{sv.ny}x{sv.nx}", "shots": sv.nshots, "receivers": sv.ntr, "itermax": itermax}
    spaces, outs = [], []
    try:
        for key, flag in (("gpu_pick_cache_s", "1"), ("gpu_parse_every_evaluation_s", "0")):
            ws = make_workspace(sv, sv.vp, invmat=1, picks=obs, ITERMAX=itermax, TIME=sv.time, SEISMO=0)
            spaces.append(ws)
            t0 = time.perf_counter()
            r = run(ws, GPU_LIB, timeout=900, env={"RAJZEL_AMD_PICK_CACHE": flag})
            res[key] = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
            # one grad_obj per FATT iteration (fatt.c:144) and per Wolfe trial (wolfels.c:66-72)
            res.setdefault("evaluations", r.stdout.count("FATT ITERATION") + r.stdout.count("gradient evaluation"))
            outs.append(outputs(ws))
        res["outputs_identical"] = outs[0] == outs[1]
        res["saved_s_per_evaluation"] = ((res["gpu_parse_every_evaluation_s"] - res["gpu_pick_cache_s"])
                                         / max(1, res["evaluations"]))
    finally:
        for ws in spaces:
            remove(ws)
    return res


def time_fatt_line_search(itermax=5, linesearch=2, workload="c2"):
    """Wall time of the MI355X program's FATT inversion (INVMAT=1, L-BFGS)
    on the 200 x 600 workload of 120 shots and 600 receivers ("c2"), or the
    1001 x 3001 one with n of its shots and 1500 receivers ("c4_<n>", a
    shard that leaves CUs idle) (its truth plus noise as pick files, no
    pick muted) with the stock parabolic line
    search (LINESEARCH = 2; 1 = Wolfe), once evaluating each line-search
    trial alone (RAJZEL_AMD_SPECULATE=0) and once with its predicted
    successors in the same launch (default).  Returns both times, the
    program's grad_obj calls and how many were served from a trial
    evaluated ahead, and whether every output file is byte-identical."""
    import time
    from rajzel_amd import native as N
    sv = W.c2_gradient() if workload == "c2" else W.c4_gradient(nshots=int(workload.split("_")[1]))
    with N.Context([dict(nx=sv.nx, ny=sv.ny, dh=sv.dh, src_ix=sv.src_ix, src_iy=sv.src_iy, rec_ix=sv.rec_ix,
                         rec_iy=sv.rec_iy)]) as ctx:
        def fwd(S, _sv):
            ctx.set_slowness(0, S)
            ctx.forward()
            return ctx.tmod(0)
        W.synthesize_picks(sv, fwd)
    obs = F.roundtrip_picks(sv.tobs)
    res = {"grid": f"{sv.ny}x{sv.nx}", "shots": sv.nshots, "receivers": sv.ntr, "itermax": itermax,
           "linesearch": "parabolic" if linesearch == 2 else "wolfe"}
    spaces, outs = [], []
    try:
        for key, flag in (("one_trial_per_launch_s", "0"), ("trials_ahead_s", "1")):
            ws = make_workspace(sv, sv.vp, invmat=1, picks=obs, ITERMAX=itermax, SEISMO=0, LINESEARCH=linesearch,
                                NLBFGS=20)
            spaces.append(ws)
            t0 = time.perf_counter()
            r = run(ws, GPU_LIB, timeout=900, env={"RAJZEL_AMD_SPECULATE": flag, "RAJZEL_AMD_SPECULATE_STATS": "1"})
            res[key] = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
            line = [t for t in r.stderr.splitlines() if t.startswith("rajzel_amd line search:")][-1]
            nums = [int(w) for w in line.replace(",", " ").split() if w.isdigit()]
            res[key.replace("_s", "_calls")] = {"grad_obj": nums[0], "evaluated_ahead": nums[1], "served": nums[2]}
            outs.append(outputs(ws))
        res["outputs_identical"] = outs[0] == outs[1]
        res["speedup"] = res["one_trial_per_launch_s"] / res["trials_ahead_s"]
    finally:
        for ws in spaces:
            remove(ws)
    return res


def time_fatt_c4(itermax=20):
    """The 1001 x 3001 configuration's inversion: 200 shots, 1500 receivers,
    l-BFGS (NLBFGS = 20) with the Wolfe line search for `itermax`
    iterations, on the MI355X program; once with the l-BFGS recursion on the
    GPU (default) and once on the host (RAJZEL_AMD_LBFGS=0).  Picks: the
    forward in the true model plus N(0, 1 ms).  Returns the wall times, the
    grad_obj calls, the first and last misfit of the log and whether the two
    runs wrote the same files."""
    import time
    from rajzel_amd import native as N
    sv = W.c4_gradient()
    with N.Context([dict(nx=sv.nx, ny=sv.ny, dh=sv.dh, src_ix=sv.src_ix, src_iy=sv.src_iy, rec_ix=sv.rec_ix,
                         rec_iy=sv.rec_iy)]) as ctx:
        ctx.set_slowness(0, W.slowness(sv.vp_true))
        ctx.forward()
        tm = ctx.tmod(0)
    obs = F.roundtrip_picks(tm + np.random.default_rng(sv.meta["seed"] + 1000).normal(0.0, sv.noise_s, tm.shape))
    res = {"grid": f"{sv.ny}x{sv.nx}", "shots": sv.nshots, "receivers": sv.ntr, "itermax": itermax,
           "method": "l-BFGS, NLBFGS = 20, Wolfe line search"}
    spaces, outs = [], []
    try:
        for key, flag in (("gpu_lbfgs_s", "1"), ("host_lbfgs_s", "0")):
            ws = make_workspace(sv, sv.vp, invmat=1, picks=obs, ITERMAX=itermax, SEISMO=0, LINESEARCH=1,
                                NLBFGS=20, GRAD_METHOD=2)
            spaces.append(ws)
            t0 = time.perf_counter()
            r = run(ws, GPU_LIB, timeout=1100, env={"RAJZEL_AMD_LBFGS": flag, "RAJZEL_AMD_SPECULATE_STATS": "1"})
            res[key] = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
            line = [t for t in r.stderr.splitlines() if t.startswith("rajzel_amd line search:")][-1]
            res.setdefault("grad_obj_calls", int(line.split()[3]))
            outs.append(outputs(ws))
            print(f"[step] {key} {res[key]:.1f} s", flush=True)
        misfit = [l.split() for l in outs[0]["misfit.dat"].decode().splitlines() if l.strip()]
        res["misfit_first"], res["misfit_last"] = float(misfit[0][1]), float(misfit[-1][1])
        res["iterations_logged"] = len(misfit)
        res["outputs_identical"] = outs[0] == outs[1]
    finally:
        for ws in spaces:
            remove(ws)
    return res


def digests(ws):
    """SHA-256 of every output file of the run (logs excluded)."""
    import hashlib
    out = {}
    for base, _, files in os.walk(ws):
        rel = os.path.relpath(base, ws)
        if rel.split(os.sep)[0] == "log":
            continue
        for f in files:
            h = hashlib.sha256()
            with open(os.path.join(base, f), "rb") as fh:
                for chunk in iter(lambda: fh.read(1 << 24), b""):
                    h.update(chunk)
            out[os.path.normpath(os.path.join(rel, f))] = h.hexdigest()
    return out


def time_fatt_c5(itermax=30, batch=100, log_dir=None):
    """The 2000 x 6000 configuration's inversion: 400 shots, 3000 receivers,
    DH = 2.5 m, l-BFGS (NLBFGS = 20) with the Wolfe line search for
    `itermax` iterations, the Gaussian gradient filter (SPATFILTER = 4, an
    81 x 41 cell window) and the horizontal gradient taper in precond, on
    the MI355X program; once with every shot's fields resident (default)
    and once under an HBM budget of `batch` shots (RAJZEL_AMD_HBM_MIB), so
    that each evaluation runs in shot batches.  Picks: the forward in the
    true model plus N(0, 1 ms).  Returns the wall times, the grad_obj
    calls, the misfit log and whether the two runs wrote the same files.
    ``log_dir``: where the program's output of each run goes as it runs."""
    import time
    from rajzel_amd import native as N
    t0 = time.perf_counter()
    sv = W.c5_gradient()
    with N.Context([dict(nx=sv.nx, ny=sv.ny, dh=sv.dh, src_ix=sv.src_ix, src_iy=sv.src_iy, rec_ix=sv.rec_ix,
                         rec_iy=sv.rec_iy)]) as ctx:
        ctx.set_slowness(0, W.slowness(sv.vp_true))
        ctx.forward()
        tm = ctx.tmod(0)
        budget_mib = -(-ctx.slot_bytes * batch // (1 << 20))
        slot_mib = ctx.slot_bytes / 2**20
    obs = F.roundtrip_picks(tm + np.random.default_rng(sv.meta["seed"] + 1000).normal(0.0, sv.noise_s, tm.shape))
    print(f"[step] workload and picks {time.perf_counter() - t0:.1f} s", flush=True)
    res = {"grid": f"{sv.ny}x{sv.nx}", "dh": sv.dh, "shots": sv.nshots, "receivers": sv.ntr, "itermax": itermax,
           "method": "l-BFGS, NLBFGS = 20, Wolfe line search, SPATFILTER = 4 (81 x 41), horizontal taper",
           "field_mib_per_shot": slot_mib, "budget_mib": budget_mib}
    stage = [dict(PRO=0.0, SPATFILTER=4, WD_DAMP=1e-3, WD_DAMP1=1e-3, FILT_SIZE_GRAD=80, FILT_SIZE_GRAD1=40)]
    spaces, outs = [], []
    try:
        for key, env in (("resident_s", {}), ("batched_s", {"RAJZEL_AMD_HBM_MIB": str(budget_mib)})):
            ws = make_workspace(sv, sv.vp, invmat=1, picks=obs, stages=stage, ITERMAX=itermax, SEISMO=0,
                                LINESEARCH=1, NLBFGS=20, GRAD_METHOD=2, SWS_TAPER_GRAD_HOR=1, GRADT1=4, GRADT2=8,
                                EXP_TAPER_GRAD_HOR=0.0)
            spaces.append(ws)
            t0 = time.perf_counter()
            r = run(ws, GPU_LIB, timeout=1000, env=dict(env, RAJZEL_AMD_SPECULATE_STATS="1"),
                    log=os.path.join(log_dir, f"fatt_c5_{key[:-2]}.log") if log_dir else None)
            res[key] = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
            line = [t for t in r.stderr.splitlines() if t.startswith("rajzel_amd line search:")][-1]
            res.setdefault("grad_obj_calls", int(line.split()[3]))
            outs.append(digests(ws))
            if key == "resident_s":
                with open(os.path.join(ws, "misfit.dat")) as fh:
                    res["misfit_log"] = [float(l.split()[1]) for l in fh if l.strip()]
            print(f"[step] {key} {res[key]:.1f} s", flush=True)
            remove(ws)
        res["iterations_logged"] = len(res["misfit_log"])
        res["files"] = len(outs[0])
        res["outputs_identical"] = outs[0] == outs[1]
    finally:
        for ws in spaces:
            remove(ws)
    return res


if __name__ == "__main__":
    import json
    import sys
    if sys.argv[1:] == ["--time-grid-search"]:
        print(json.dumps(time_grid_search()), flush=True)
    elif sys.argv[1:] == ["--time-fatt-picks"]:
        print(json.dumps(time_fatt_picks()), flush=True)
    elif sys.argv[1:] == ["--time-fatt-c4"]:
        print(json.dumps(time_fatt_c4()), flush=True)
    elif sys.argv[1:2] == ["--time-fatt-c5"]:
        print(json.dumps(time_fatt_c5(log_dir=sys.argv[2] if len(sys.argv) > 2 else None)), flush=True)
    elif sys.argv[1:2] == ["--time-line-search"]:
        ls = int(sys.argv[2]) if len(sys.argv) > 2 else 2
        wl = sys.argv[3] if len(sys.argv) > 3 else "c2"
        print(json.dumps(time_fatt_line_search(itermax=5 if wl == "c2" else 6, linesearch=ls, workload=wl)),
              flush=True)
    else:
        sys.exit("usage: python -m tests.fd_program --time-grid-search | --time-fatt-picks | "
                 "--time-line-search [2|1] [c2|c4_<shots>] | --time-fatt-c4 | --time-fatt-c5 [log dir]")
