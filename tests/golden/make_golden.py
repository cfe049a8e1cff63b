#!/usr/bin/env python3
"""Write the golden vectors of tests/golden/*.npz from the reference itself.

Every expected output here comes from the reference's own C sources compiled
by oracle/Makefile (oracle/_ref/librzref.so, driven by oracle/ref_harness.c);
inputs are seeded synthetic data.  Run from the repository root:

    make -C oracle && python tests/golden/make_golden.py

Cases
  sweeps.npz      one reference sweep() in each of the 4 directions from a
                  random T on a 17x23 grid (borders and corners included)
  sweeps_adj.npz  one reference sweep_adj() in each direction, with receiver
                  cells (recflag) and cells at TT >= INF
  ass_grad.npz    ass_grad() + cp_grad_frame() on random lambda, S, grad
  eval_small.npz  grad_obj's shot loop, 37x53, 4 shots, one muted pick per
                  shot 1..2 and a non-zero initial Tres (carry)
                  (receivers alternate between row 2 and a line at depth)
  eval_bands.npz  the same on a 150x70 grid (several row bands on the GPU)
  eval_surface.npz  surface receivers only (row 2), as the reference's own
                  example acquisition
  eval_early_wide.npz, eval_early_wide6.npz, eval_early_tall.npz
                  grad_obj's shot loop on rough models (workloads.edge_case)
                  at TTNORM 1e-4 / 1e-6 (200x600) and 1e-6 (1001x29): the
                  forward and adjoint solves stop before a fixed point;
                  receivers duplicated on one cell, on the top and bottom
                  frame rows, on the frame columns and at depth
  c1_forward.npz  eikonal() of the 200x600 forward configuration: receiver
                  times, iterations and the full traveltime grid
  gauss_filt.npz  gauss_filt() of a random 37x53 field and of a gradient-like
                  48x130 field, with odd, even (made odd) and negative sizes
                  (no normalisation), and a size of 0
  taper.npz       grad_obj with the per-shot gradient tapers of ass_grad.c:
                  31-37 (SWS_TAPER_CIRCULAR_PER_SHOT 1: SRTRADIUS 35 and 80 m;
                  2: FILTSIZE 1 and 0) on workloads.taper_case (windows that
                  run off the grid sideways) with single-shot ass_grad of
                  shots 0 and 3, and on 3 shots of the 200x600 C2 geometry
                  (radial taper at shots 57-59; windows of FILTSIZE 0), the
                  latter as a SHA-256 digest of the gradient and four rows
  grad_taper.npz  taper_grad() of precond.c:43-53 on a seeded 34x60 field:
                  vertical (5 GRADT sets) and horizontal (5 GRADT/EXP sets)
                  tapers with the taper_coeff files they write, the
                  source/receiver taper (SRTRADIUS 35, 80 m) and the file
                  taper (sws 4, 6); the stock horizontal taper on 1001x3001
                  and the source/receiver taper on the C2 geometry as
                  digests and rows
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.bindings import Reference  # noqa: E402
from rajzel_amd import workloads as W  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def sweeps(ref):
    rng = np.random.default_rng(101)
    ny, nx, dh = 17, 23, 10.0
    S = (1.0 / rng.uniform(1500, 4000, (ny, nx))).astype(np.float32)
    T0 = rng.uniform(0.0, 0.05, (ny, nx)).astype(np.float32)
    ref.setup(nx, ny, dh, 6.0, 1e-10, 1e-9)
    out = {"S": S, "T0": T0, "dh": np.float32(dh)}
    for sx, sy in [(1, 1), (-1, 1), (-1, -1), (1, -1)]:
        out[f"T_{sx:+d}{sy:+d}"] = ref.sweep(S, T0, sx, sy)
    np.savez_compressed(os.path.join(OUT, "sweeps.npz"), **out)


def sweeps_adj(ref):
    rng = np.random.default_rng(202)
    ny, nx, dh = 19, 21, 5.0
    S = (1.0 / rng.uniform(1500, 4000, (ny, nx))).astype(np.float32)
    ref.setup(nx, ny, dh, 6.0, 1e-10, 1e-9)
    T, _, _ = ref.eikonal(S, 10 * dh, 1 * dh)
    T = T.copy()
    T[12:14, 15:17] = np.float32(2.0e7)       # >= INF
    rec_ix = np.array([3, 7, 7, 12, 0, 20, 9], np.int32)   # one duplicate cell, frame cells
    rec_iy = np.array([1, 1, 1, 4, 3, 5, 18], np.int32)
    flag = ref.init_recflag(rec_ix, rec_iy)
    lam0 = rng.uniform(-0.01, 0.01, (ny, nx)).astype(np.float32)
    out = {"S": S, "T": T, "rec_ix": rec_ix, "rec_iy": rec_iy, "flag": flag, "lam0": lam0,
           "dh": np.float32(dh), "eps_adj": np.float32(1e-9)}
    for sx, sy in [(1, 1), (1, -1), (-1, 1), (-1, -1)]:
        out[f"lam_{sx:+d}{sy:+d}"] = ref.sweep_adj(lam0, T, flag, sx, sy)
    np.savez_compressed(os.path.join(OUT, "sweeps_adj.npz"), **out)


def ass_grad(ref):
    rng = np.random.default_rng(303)
    ny, nx = 13, 16
    S = (1.0 / rng.uniform(1500, 4000, (ny, nx))).astype(np.float32)
    lam = rng.normal(0, 1e-2, (ny, nx)).astype(np.float32)
    g0 = rng.normal(0, 1e-9, (ny, nx)).astype(np.float32)
    ref.setup(nx, ny, 10.0)
    gs, g = ref.ass_grad(S, lam, g0)
    np.savez_compressed(os.path.join(OUT, "ass_grad.npz"), S=S, lam=lam, grad0=g0, grad_shot=gs, grad=g)


def evaluation(ref, name, nx, ny, nshots, seed, mute, rec_row):
    sv = W.small_case(nx=nx, ny=ny, nshots=nshots, seed=seed, rec_row=rec_row)
    ref.setup(nx, ny, sv.dh, sv.time, sv.ttnorm, sv.eps_adj)
    # picks: reference forward in the true model + noise, as written to and read back from pick files
    St = W.slowness(sv.vp_true)
    tm = np.stack([ref.eikonal(St, float(sv.src_x_m[s]), float(sv.src_z_m[s]), sv.rec_ix, sv.rec_iy)[1]
                   for s in range(nshots)])
    rng = np.random.default_rng(seed + 1)
    tobs = (tm + rng.normal(0, 0.002, tm.shape)).astype(np.float32)
    for (s, i) in mute:
        tobs[s, i] = -1.0
    tres0 = np.linspace(-1e-3, 1e-3, sv.ntr).astype(np.float32)
    out = ref.grad_obj(sv.S, sv.src_x_m, sv.src_z_m, sv.rec_ix, sv.rec_iy, tobs, tres0=tres0)
    np.savez_compressed(
        os.path.join(OUT, f"{name}.npz"),
        nx=nx, ny=ny, dh=np.float32(sv.dh), time=np.float32(sv.time), ttnorm=np.float32(sv.ttnorm),
        eps_adj=np.float32(sv.eps_adj), S=sv.S, src_x_m=sv.src_x_m, src_z_m=sv.src_z_m,
        src_ix=sv.src_ix, src_iy=sv.src_iy, rec_ix=sv.rec_ix, rec_iy=sv.rec_iy, tres0=tres0,
        tobs=out["tobs"], grad=out["grad"], T_last=out["T"], lam_last=out["lam"], tmod=out["tmod"],
        tres=out["tres"], tres_carry=out["tres_carry"], l2_shot=out["l2_shot"], L2=np.float64(out["L2"]),
        it_f=out["it_f"], it_a=out["it_a"])


def evaluation_survey(ref, name, sv, mute, seed):
    """grad_obj on a workload Survey (its TTNORM), picks from the reference's
    forward in the survey's true model plus N(0, noise), muted picks."""
    ref.setup(sv.nx, sv.ny, sv.dh, sv.time, sv.ttnorm, sv.eps_adj)
    St = W.slowness(sv.vp_true)
    tm = np.stack([ref.eikonal(St, float(sv.src_x_m[s]), float(sv.src_z_m[s]), sv.rec_ix, sv.rec_iy)[1]
                   for s in range(sv.nshots)])
    rng = np.random.default_rng(seed)
    tobs = (tm + rng.normal(0, sv.noise_s, tm.shape)).astype(np.float32)
    for (s, i) in mute:
        tobs[s, i] = -1.0
    tres0 = np.linspace(-1e-3, 1e-3, sv.ntr).astype(np.float32)
    out = ref.grad_obj(sv.S, sv.src_x_m, sv.src_z_m, sv.rec_ix, sv.rec_iy, tobs, tres0=tres0)
    np.savez_compressed(
        os.path.join(OUT, f"{name}.npz"),
        nx=sv.nx, ny=sv.ny, dh=np.float32(sv.dh), time=np.float32(sv.time), ttnorm=np.float32(sv.ttnorm),
        eps_adj=np.float32(sv.eps_adj), S=sv.S, src_x_m=sv.src_x_m, src_z_m=sv.src_z_m,
        src_ix=sv.src_ix, src_iy=sv.src_iy, rec_ix=sv.rec_ix, rec_iy=sv.rec_iy, tres0=tres0,
        tobs=out["tobs"], grad=out["grad"], T_last=out["T"], lam_last=out["lam"], tmod=out["tmod"],
        tres=out["tres"], tres_carry=out["tres_carry"], l2_shot=out["l2_shot"], L2=np.float64(out["L2"]),
        it_f=out["it_f"], it_a=out["it_a"])


def c1_forward(ref):
    sv = W.c1_forward()
    ref.setup(sv.nx, sv.ny, sv.dh, sv.time, sv.ttnorm, sv.eps_adj)
    T, tm, it = ref.eikonal(sv.S, float(sv.src_x_m[0]), float(sv.src_z_m[0]), sv.rec_ix, sv.rec_iy)
    np.savez_compressed(os.path.join(OUT, "c1_forward.npz"), S=sv.S, src_ix=sv.src_ix,
                        src_iy=sv.src_iy, rec_ix=sv.rec_ix, rec_iy=sv.rec_iy, dh=np.float32(sv.dh),
                        T=T, tmod=tm, iters=np.int32(it))


GAUSS_CASES = [("a", 37, 53, 9, 5), ("b", 37, 53, 10, 6), ("c", 48, 130, 81, 41), ("d", 37, 53, -7, 9),
               ("e", 37, 53, 0, 5)]


def gauss_filt(ref):
    rng = np.random.default_rng(404)
    out = {}
    for name, ny, nx, fx, fy in GAUSS_CASES:
        A = rng.normal(0.0, 1e-3, (ny, nx)).astype(np.float32)
        ref.setup(nx, ny, 10.0)
        B, ox, oy = ref.gauss_filt(A, fx, fy)
        out[f"{name}_in"] = A
        out[f"{name}_out"] = B
        out[f"{name}_sizes"] = np.array([fx, fy, ox, oy], np.int32)
    np.savez_compressed(os.path.join(OUT, "gauss_filt.npz"), **out)


TAPER_SMALL = [("r35", 1, 35.0, 1), ("r80", 1, 80.0, 1), ("w1", 2, 80.0, 1), ("w0", 2, 80.0, 0)]
# C2 geometry (sources at row 1, receivers at row 2): FILTSIZE 0 keeps every
# taper_grad_shot1.c window on the grid
TAPER_C2 = [("c2_r80", 1, 80.0, 1, 57), ("c2_w0", 2, 80.0, 0, 0)]


def _taper_picks(ref, sv, seed):
    St = W.slowness(sv.vp_true)
    tm = np.stack([ref.eikonal(St, float(sv.src_x_m[s]), float(sv.src_z_m[s]), sv.rec_ix, sv.rec_iy)[1]
                   for s in range(sv.nshots)])
    return (tm + np.random.default_rng(seed).normal(0, sv.noise_s, tm.shape)).astype(np.float32)


def tapers(ref):
    """grad_obj with the per-shot gradient tapers (ass_grad.c:31-37), and
    single-shot ass_grad with them."""
    import hashlib
    out = {}
    sv = W.taper_case()
    ref.setup(sv.nx, sv.ny, sv.dh, sv.time, sv.ttnorm, sv.eps_adj)
    tobs = _taper_picks(ref, sv, 6)
    for name, sws, radius, filt in TAPER_SMALL:
        ref.setup(sv.nx, sv.ny, sv.dh, sv.time, sv.ttnorm, sv.eps_adj)
        ref.set_taper(sws, radius, filt)
        o = ref.grad_obj(sv.S, sv.src_x_m, sv.src_z_m, sv.rec_ix, sv.rec_iy, tobs)
        out["small_tobs"] = o["tobs"]   # as read back from the pick files
        out[f"small_{name}_set"] = np.array([sws, radius, filt], np.float64)
        out[f"small_{name}_grad"] = o["grad"]
        out[f"small_{name}_L2"] = np.float64(o["L2"])
        # single-shot ass_grad of shots 0 and 3 on a random lambda and grad
        rng = np.random.default_rng(90 + sws * 10 + filt)
        lam = rng.normal(0, 1, (sv.ny, sv.nx)).astype(np.float32)
        g0 = rng.normal(0, 1, (sv.ny, sv.nx)).astype(np.float32)
        out[f"small_{name}_lam"], out[f"small_{name}_g0"] = lam, g0
        for shot in (0, 3):
            gs, g = ref.ass_grad_at(sv.S, lam, g0, sv.src_x_m, sv.src_z_m, sv.rec_ix, sv.rec_iy, shot)
            out[f"small_{name}_shot{shot}_gs"], out[f"small_{name}_shot{shot}_g"] = gs, g
    full = W.c2_gradient()
    ref.setup(full.nx, full.ny, full.dh, full.time, full.ttnorm, full.eps_adj)
    for name, sws, radius, filt, first in TAPER_C2:
        sub = W.c2_gradient()
        sel = slice(first, first + 3)
        sub.src_x_m, sub.src_z_m = full.src_x_m[sel].copy(), full.src_z_m[sel].copy()
        ref.setup(full.nx, full.ny, full.dh, full.time, full.ttnorm, full.eps_adj)
        tobs = _taper_picks(ref, sub, 7)
        ref.set_taper(sws, radius, filt)
        o = ref.grad_obj(sub.S, sub.src_x_m, sub.src_z_m, sub.rec_ix, sub.rec_iy, tobs)
        out[f"{name}_set"] = np.array([sws, radius, filt, first], np.float64)
        out[f"{name}_tobs"] = o["tobs"]
        out[f"{name}_d_grad"] = np.array(hashlib.sha256(np.ascontiguousarray(o["grad"]).tobytes()).hexdigest())
        out[f"{name}_grad_rows"] = o["grad"][[1, 2, 100, 198]]
        out[f"{name}_L2"] = np.float64(o["L2"])
        print(f"  {name}: L2 {o['L2']:.9e}", flush=True)
    np.savez_compressed(os.path.join(OUT, "taper.npz"), **out)


GT_COLS = [(22, 25, 40, 50), (0, 5, 50, 59), (3, 3, 10, 10), (-2, 4, 55, 70), (5, 12, 20, 30)]
GT_ROWS = [((22, 25, 0, 0), 0.0), ((2, 9, 0, 0), 1.0), ((0, 5, 0, 0), 0.5), ((3, 3, 0, 0), 2.0), ((-1, 6, 0, 0), 0.25)]


def grad_tapers(ref):
    """taper_grad() (precond.c:43-53) on seeded random fields: every sws
    with several geometries on 34x60, the stock horizontal taper on the
    1001x3001 grid and the source/receiver taper on the 200x600 C2
    geometry (the last two as digests and rows)."""
    import hashlib
    dig = lambda a: np.array(hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest())  # noqa: E731
    out = {}
    rng = np.random.default_rng(505)
    nx, ny = 60, 34
    A = rng.normal(0, 1, (ny, nx)).astype(np.float32)
    out["small_in"] = A
    ref.setup(nx, ny, 10.0)
    for i, g in enumerate(GT_COLS):
        B, c = ref.taper_grad(A, 1, gradt=g)
        out[f"v{i}_gradt"], out[f"v{i}_out"], out[f"v{i}_coeff"] = np.array(g, np.int32), B, c
    for i, (g, e) in enumerate(GT_ROWS):
        B, c = ref.taper_grad(A, 2, gradt=g, exp_hor=e)
        out[f"h{i}_gradt"], out[f"h{i}_exp"], out[f"h{i}_out"], out[f"h{i}_coeff"] = \
            np.array(g, np.int32), np.float32(e), B, c
    sv = W.taper_case()
    for i, R in enumerate((35.0, 80.0)):
        B, _ = ref.taper_grad(A, 3, srtradius=R, src_x_m=sv.src_x_m, src_z_m=sv.src_z_m, rec_ix=sv.rec_ix,
                              rec_iy=sv.rec_iy)
        out[f"r{i}_radius"], out[f"r{i}_out"] = np.float32(R), B
    tf = rng.uniform(0, 2, nx * ny).astype(np.float32)
    for sws in (4, 6):
        B, c = ref.taper_grad(A, sws, taper_file=tf)
        out[f"f{sws}_out"], out[f"f{sws}_coeff"] = B, c
    out["taper_file"] = tf
    # the stock horizontal taper (GRADT 22, 25, EXP 0) on the 1001x3001 grid
    big = np.random.default_rng(506).normal(0, 1e-3, (1001, 3001)).astype(np.float32)
    ref.setup(3001, 1001, 5.0)
    B, c = ref.taper_grad(big, 2, gradt=(22, 25, 490, 500), exp_hor=0.0)
    out["c4hor_d_out"], out["c4hor_rows"], out["c4hor_d_coeff"] = dig(B), B[[0, 21, 22, 24, 25, 1000]], dig(c)
    # sources and receivers of the C2 geometry, SRTRADIUS 80 m
    c2 = W.c2_gradient()
    mid = np.random.default_rng(507).normal(0, 1e-3, (c2.ny, c2.nx)).astype(np.float32)
    ref.setup(c2.nx, c2.ny, c2.dh)
    B, _ = ref.taper_grad(mid, 3, srtradius=80.0, src_x_m=c2.src_x_m, src_z_m=c2.src_z_m, rec_ix=c2.rec_ix,
                          rec_iy=c2.rec_iy)
    out["c2disc_d_out"], out["c2disc_rows"] = dig(B), B[[0, 1, 2, 5, 100, 199]]
    np.savez_compressed(os.path.join(OUT, "grad_taper.npz"), **out)


def main():
    if sys.argv[1:] == ["taper"]:
        tapers(Reference())
        return
    if sys.argv[1:] == ["grad_taper"]:
        grad_tapers(Reference())
        return
    ref = Reference()
    sweeps(ref)
    sweeps_adj(ref)
    ass_grad(ref)
    evaluation(ref, "eval_small", 53, 37, 4, 31, [(1, 5), (2, 5), (2, 11)], (2, 36))
    evaluation(ref, "eval_bands", 70, 150, 3, 41, [(0, 0), (2, 69)], (2, 120, 149))
    evaluation(ref, "eval_surface", 53, 37, 3, 51, [(1, 7)], 2)
    # solves that stop before a fixed point, with receiver edge cases
    evaluation_survey(ref, "eval_early_wide", W.edge_case(600, 200, ttnorm=1e-4), [(0, 3), (1, 3), (2, 40)], 71)
    evaluation_survey(ref, "eval_early_wide6", W.edge_case(600, 200, ttnorm=1e-6, seed=62), [(1, 0)], 72)
    evaluation_survey(ref, "eval_early_tall", W.edge_case(29, 1001, rec_step=3, ttnorm=1e-6, sigma=1.0),
                      [(2, 5)], 73)
    c1_forward(ref)
    gauss_filt(ref)
    tapers(ref)
    grad_tapers(ref)
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    main()
