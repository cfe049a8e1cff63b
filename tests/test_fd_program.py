"""The reference program driving the MI355X path through the fd.h entry
points (rajzel_amd/csrc/rz_fd_surface.c, which also replaces gauss_filt.c,
taper_grad.c, grid_search.c, LBFGS.c, zero_LBFGS.c, zero_LBFGS1.c and
calc_mat_change_wolfe.c): rajzel.c, fatt.c, PCG.c,
wolfels.c, parabolicls.c, precond.c, readmod.c and writemod.c run unchanged, and every
file they write must equal, byte for byte, what the unchanged reference
program writes from the same inputs.

Both programs are the reference's sources compiled by oracle/Makefile
(oracle/_ref/librajzel_cpu.so, oracle/_ref/librajzel_gpu.so) and started
by oracle/_ref/prog_launcher; tests/fd_program.py builds the workspaces.
"""
import os
import subprocess

import numpy as np
import pytest

from rajzel_amd import formats as F
from tests import fd_program as P
from tests.helpers import bits_equal

SURFACE = ("eikonal", "eikonal_adj", "ass_grad", "grad_obj", "forward")


def _need(gpu=False):
    if not P.available(gpu):
        pytest.skip("reference program build (oracle/_ref) not present")


def _defined(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}


def test_gpu_program_links_the_fd_surface():
    _need(gpu=True)
    cpu, gpu = _defined(P.CPU_LIB), _defined(P.GPU_LIB)
    for name in SURFACE + ("sweep", "sweep_adj", "main", "fatt", "LBFGS", "wolfels"):
        assert name in cpu, name
    for name in SURFACE + ("main", "fatt", "LBFGS", "PCG", "wolfels", "parabolicls", "grid_search"):
        assert name in gpu, name
    # the replaced sources are gone: their sweeps live in the kernels
    assert "sweep" not in gpu and "sweep_adj" not in gpu


def test_reference_program_forward_equals_restatement(ora):
    _need()
    sv = P.survey()
    ws = P.make_workspace(sv, sv.vp_true, invmat=0)
    try:
        r = P.run(ws, P.CPU_LIB)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        picks = np.stack([F.read_picks(os.path.join(ws, "picks", f"mod_{s + 1}.dat"), sv.ntr)
                          for s in range(sv.nshots)])
        S = (1.0 / sv.vp_true.astype(np.float64)).astype(np.float32)
        tm, _, _ = ora.forward(S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.dh)
        assert bits_equal(picks, F.roundtrip_picks(tm))
    finally:
        P.remove(ws)


def _observed(sv, muted=True):
    """Picks of the reference forward in the true model, N(0, 1 ms) noise,
    one muted pick (unless muted=False); as read back from pick files."""
    from oracle.bindings import Restatement
    S = (1.0 / sv.vp_true.astype(np.float64)).astype(np.float32)
    tm, _, _ = Restatement().forward(S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.dh)
    obs = F.roundtrip_picks(tm + np.random.default_rng(3).normal(0, 1e-3, tm.shape))
    if muted:
        obs[1, 4] = -1.0
    return obs


def _compare(make, env=None, runs=None, ranks=1):
    """Run the unchanged and the MI355X program on identical workspaces
    (env: extra environment of the MI355X run; runs: a list that receives
    its CompletedProcess; ranks: MPI ranks of both runs)."""
    _need(gpu=True)
    ws_c, ws_g = make(), make()
    try:
        print(f"[case] reference program in {ws_c}, MI355X program in {ws_g}, {ranks} rank(s)", flush=True)
        rc = P.run(ws_c, P.CPU_LIB, ranks=ranks)
        assert rc.returncode == 0, rc.stdout[-2000:] + rc.stderr[-2000:]
        rg = P.run(ws_g, P.GPU_LIB, env=env, ranks=ranks)
        assert rg.returncode == 0, rg.stdout[-2000:] + rg.stderr[-2000:]
        if runs is not None:
            runs.append(rg)
        oc, og = P.outputs(ws_c), P.outputs(ws_g)
        assert sorted(oc) == sorted(og)
        differ = [k for k in oc if oc[k] != og[k]]
        assert not differ, f"files differ: {differ}"
        return oc
    finally:
        P.remove(ws_c)
        P.remove(ws_g)


@pytest.mark.gpu
def test_forward_program_with_snapshots():
    sv = P.survey()
    out = _compare(lambda: P.make_workspace(sv, sv.vp_true, invmat=0, SNAP=1))
    assert sum(k.startswith("picks/mod_") for k in out) == sv.nshots
    assert sum(k.startswith("tt/snap_shot_") for k in out) == sv.nshots


@pytest.mark.gpu
def test_programs_under_an_hbm_budget():
    """RAJZEL_AMD_HBM_MIB=1 leaves room for the fields of some of 16 shots:
    the forward with snapshots (the shots outside the last batch solved
    again one at a time) and three l-BFGS iterations run in shot batches and
    write the reference's files byte for byte."""
    from rajzel_amd import native as N
    sv = P.survey(nshots=16)
    with N.Context([dict(nx=sv.nx, ny=sv.ny, dh=sv.dh, src_ix=sv.src_ix, src_iy=sv.src_iy, rec_ix=sv.rec_ix,
                         rec_iy=sv.rec_iy)], hbm_budget=1 << 20) as ctx:
        assert 1 <= ctx.batch_shots < sv.nshots, (ctx.batch_shots, ctx.slot_bytes)
    env = {"RAJZEL_AMD_HBM_MIB": "1"}
    out = _compare(lambda: P.make_workspace(sv, sv.vp_true, invmat=0, SNAP=1), env=env)
    assert sum(k.startswith("tt/snap_shot_") for k in out) == sv.nshots
    obs = _observed(sv)
    out = _compare(lambda: P.make_workspace(sv, P.start_model(sv), invmat=1, picks=obs, ITERMAX=3), env=env)
    assert len([m for m in out["misfit.dat"].decode().split("\n") if m.strip()]) == 3


@pytest.mark.gpu
def test_programs_on_a_transposed_grid():
    """A 4500 x 18 grid (more rows than a workgroup holds, solved
    transposed): the forward with snapshots, and the grid search over 9
    (vp0, grad0) models in one launch, write the reference's files byte for
    byte."""
    sv = P.survey(nx=18, ny=4500, nshots=3, seed=23)
    out = _compare(lambda: P.make_workspace(sv, sv.vp_true, invmat=0, SNAP=1))
    assert sum(k.startswith("tt/snap_shot_") for k in out) == sv.nshots
    obs = _observed(sv)
    grid = dict(VP0_1=1400.0, VP0_2=1650.0, DVP0=100.0, GRAD0_1=0.3, GRAD0_2=0.8, DGRAD0=0.2)
    out = _compare(lambda: P.make_workspace(sv, P.start_model(sv), invmat=2, picks=obs, **grid))
    assert len([m for m in out["gs.dat"].decode().split("\n") if m.strip()]) == 10   # header + 9 models


@pytest.mark.gpu
def test_fatt_lbfgs_wolfe():
    sv = P.survey()
    obs = _observed(sv)
    out = _compare(lambda: P.make_workspace(sv, P.start_model(sv), invmat=1, picks=obs, ITERMAX=3))
    misfit = out["misfit.dat"].decode().split("\n")
    assert len([m for m in misfit if m.strip()]) == 3


@pytest.mark.gpu
@pytest.mark.parametrize("kw", [dict(ITERMAX=5, NLBFGS=2), dict(ITERMAX=3, IDX=2, IDY=3)])
def test_fatt_lbfgs_history(kw):
    """The shim's LBFGS: on the GPU with slots that wrap around (NLBFGS = 2
    over 5 iterations), on the host with strided history (IDX, IDY); the
    history files, models and misfit log equal the reference's."""
    sv = P.survey()
    obs = _observed(sv)
    out = _compare(lambda: P.make_workspace(sv, P.start_model(sv), invmat=1, picks=obs, **kw))
    for f in ("jac/grad_p.old", "jac/grad_p_vp.old", "jac/grad_c.old"):
        assert f in out
    n = ((sv.nx + kw.get("IDX", 1) - 1) // kw.get("IDX", 1)) * ((sv.ny + kw.get("IDY", 1) - 1) // kw.get("IDY", 1))
    assert len(out["jac/grad_c.old"]) == 4 * n


@pytest.mark.gpu
def test_fatt_lbfgs_stages():
    """Two workflow stages: the first ends after three iterations (PRO = 1
    accepts any misfit change, fatt.c:219-251), which zeroes the l-BFGS
    history (zero_LBFGS, fatt.c:232) and starts the second at iteration 1
    with the pointer at slot 1; the GPU history must follow."""
    sv = P.survey()
    obs = _observed(sv)
    row = dict(SPATFILTER=0, WD_DAMP=1e-3, WD_DAMP1=1e-3, FILT_SIZE_GRAD=8, FILT_SIZE_GRAD1=6)
    out = _compare(lambda: P.make_workspace(sv, P.start_model(sv), invmat=1, picks=obs, ITERMAX=4, NLBFGS=3,
                                            stages=[dict(row, PRO=1.0), dict(row, PRO=0.0)]))
    assert "model/inv_vp_stage_1.bin" in out
    assert len([m for m in out["misfit.dat"].decode().split("\n") if m.strip()]) > 3   # the second stage ran


@pytest.mark.gpu
@pytest.mark.parametrize("ls,method,eps", [(2, 2, 0.01), (1, 2, 1e-5), (2, 1, 0.01)],
                         ids=["parabolic-lbfgs", "wolfe-lbfgs", "parabolic-pcg"])
def test_fatt_line_search_trials_ahead(ls, method, eps):
    """The batched line search: with no muted pick and idle CUs, grad_obj
    evaluates each line-search trial together with the predicted next one
    and serves the next call from it when the search asks for that model.
    Every file equals the reference program's, and calls are served (the
    Wolfe case starts from a step too short, so that its search expands)."""
    sv = P.survey()
    obs = _observed(sv, muted=False)
    runs = []
    _compare(lambda: P.make_workspace(sv, P.start_model(sv), invmat=1, picks=obs, ITERMAX=4, LINESEARCH=ls,
                                      GRAD_METHOD=method, EPS_SCALE=eps),
             env={"RAJZEL_AMD_SPECULATE_STATS": "1"}, runs=runs)
    line = [t for t in runs[0].stderr.splitlines() if t.startswith("rajzel_amd line search:")]
    assert line, runs[0].stderr[-2000:]
    served = int(line[-1].split(",")[-1].split()[0])
    print(f"[result] {line[-1]}", flush=True)
    assert served > 0


@pytest.mark.gpu
@pytest.mark.parametrize("ranks", [2, 3])
def test_fatt_mpi_ranks(ranks):
    """The program as 2 and 3 MPI ranks on the one GPU, each evaluating its
    shot range [NSHOT1, NSHOT2) of init_MPIshot.c (5 shots: uneven ranges
    with 3 ranks) and summing gradients and misfits through MPI
    (sum_grad_MPI.c, grad_obj.c:58-61), l-BFGS on rank 0 and its direction
    broadcast (exchange_grad_MPI.c): every file equals the reference
    program's with the same ranks.  A forward run writes each rank's picks."""
    if P.MPIEXEC is None:
        pytest.skip("no mpiexec")
    sv = P.survey()
    obs = _observed(sv)
    out = _compare(lambda: P.make_workspace(sv, P.start_model(sv), invmat=1, picks=obs, ITERMAX=3),
                   ranks=ranks)
    assert len([m for m in out["misfit.dat"].decode().split("\n") if m.strip()]) == 3
    out = _compare(lambda: P.make_workspace(sv, sv.vp_true, invmat=0), ranks=ranks)
    assert sum(k.startswith("picks/mod_") for k in out) == sv.nshots


@pytest.mark.gpu
@pytest.mark.parametrize("spatfilter", [2, 3])
def test_fatt_smoothing_filters(spatfilter):
    """precond.c's SPATFILTER = 2 (smooth2.c, WD_DAMP/WD_DAMP1) and 3
    (smooth_grad.c, the median over FILT_SIZE_GRAD) through the shim's GPU
    versions: every file (models, history, smooth2's two files) equals the
    reference program's."""
    sv = P.survey()
    obs = _observed(sv)
    row = dict(PRO=0.0, SPATFILTER=spatfilter, WD_DAMP=2.0, WD_DAMP1=3.0, FILT_SIZE_GRAD=4, FILT_SIZE_GRAD1=6)
    out = _compare(lambda: P.make_workspace(sv, P.start_model(sv), invmat=1, picks=obs, ITERMAX=3, stages=[row]))
    if spatfilter == 2:
        assert "jac/grad_smooth2.old" in out and "jac/grad_wavenumber.old" in out


@pytest.mark.gpu
def test_fatt_pcg_parabolic_line_search():
    """PCG.c directions with the parabolic line search (parabolicls.c)."""
    sv = P.survey()
    obs = _observed(sv)
    out = _compare(lambda: P.make_workspace(sv, P.start_model(sv), invmat=1, picks=obs, ITERMAX=3, GRAD_METHOD=1,
                                            LINESEARCH=2))
    assert len([m for m in out["misfit.dat"].decode().split("\n") if m.strip()]) == 3


@pytest.mark.gpu
def test_fatt_wolfe_gauss_filter():
    """The Gaussian gradient filter of precond.c / fatt.c (SPATFILTER=4)."""
    sv = P.survey()
    obs = _observed(sv)
    _compare(lambda: P.make_workspace(sv, P.start_model(sv), invmat=1, picks=obs, ITERMAX=3, spatfilter=4))


@pytest.mark.gpu
@pytest.mark.parametrize("kw", [dict(SWS_TAPER_CIRCULAR_PER_SHOT=1, SRTRADIUS=30.0),
                                dict(SWS_TAPER_CIRCULAR_PER_SHOT=2, FILTSIZE=0)])
def test_fatt_per_shot_taper(kw):
    """The per-shot gradient tapers of ass_grad.c:31-37 in the inversion.
    Sources sit on row 1, so FILTSIZE 0 keeps taper_grad_shot1.c's windows
    on the grid (a larger one writes through a row pointer above the
    reference's matrix)."""
    sv = P.survey()
    obs = _observed(sv)
    _compare(lambda: P.make_workspace(sv, P.start_model(sv), invmat=1, picks=obs, ITERMAX=3, **kw))


@pytest.mark.gpu
@pytest.mark.parametrize("kw", [dict(SWS_TAPER_GRAD_VERT=1, GRADT1=5, GRADT2=12, GRADT3=55, GRADT4=68),
                                dict(SWS_TAPER_GRAD_HOR=1, GRADT1=3, GRADT2=6, EXP_TAPER_GRAD_HOR=0.5),
                                dict(SWS_TAPER_GRAD_SOURCES=1, SRTRADIUS=30.0),
                                dict(SWS_TAPER_FILE=1),
                                dict(SWS_TAPER_GRAD_VERT=1, SWS_TAPER_GRAD_HOR=1, SWS_TAPER_GRAD_SOURCES=1,
                                     SWS_TAPER_FILE=1, GRADT1=3, GRADT2=6, GRADT3=60, GRADT4=70, SRTRADIUS=25.0)],
                         ids=["vertical", "horizontal", "sources", "file", "all"])
def test_fatt_gradient_tapers(kw):
    """The gradient tapers of precond.c:43-53 (taper_grad.c) in the
    inversion, with the Gaussian filter's companion taper files: models,
    misfit log, L-BFGS files and taper_coeff_*.bin equal the reference's."""
    sv = P.survey()
    obs = _observed(sv)

    def make():
        ws = P.make_workspace(sv, P.start_model(sv), invmat=1, picks=obs, ITERMAX=3, **kw)
        if kw.get("SWS_TAPER_FILE"):   # read by taper_grad.c:272 from the working directory
            tb = np.random.default_rng(9).uniform(0.5, 1.5, sv.nx * sv.ny).astype(np.float32)
            tb.tofile(os.path.join(ws, "taper.bin"))
        return ws
    out = _compare(make)
    for flag, name in (("SWS_TAPER_GRAD_VERT", "taper_coeff_vert.bin"), ("SWS_TAPER_GRAD_HOR", "taper_coeff_hor.bin"),
                       ("SWS_TAPER_FILE", "taper_coeff_file.bin")):
        assert (name in out) == bool(kw.get(flag)), name


@pytest.mark.gpu
def test_grid_search_program():
    sv = P.survey(nx=60, ny=30, nshots=3, seed=23)
    obs = _observed(sv)
    out = _compare(lambda: P.make_workspace(sv, P.start_model(sv), invmat=2, picks=obs))
    assert out["gs.dat"]
