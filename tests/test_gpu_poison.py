"""No kernel reads a field position before a kernel wrote it: with
RAJZEL_AMD_POISON=1 every field buffer of a context starts as NaN, so such a
read would carry NaN into the results.  Evaluations (split shots, one
workgroup per shot, HBM-budget batches, solves stopped before a fixed point
whose converged fields end in either buffer of the alternation) and the
single-shot entry points must still give the reference's bits."""
import os

import numpy as np
import pytest

from rajzel_amd import workloads as W
from tests.helpers import bits_equal, model_dict, ndiff, oracle_picks

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name,R,split", [("eval_early_wide", 2, True), ("eval_early_wide", 4, False),
                                          ("eval_early_tall", 4, True), ("eval_bands", 1, False)])
def test_poisoned_fields_golden(name, R, split, monkeypatch):
    from rajzel_amd import native as N
    monkeypatch.setenv("RAJZEL_AMD_POISON", "1")
    d = np.load(os.path.join(GOLD, f"{name}.npz"))
    print(f"[case] poisoned fields {name} R={R} split={split}", flush=True)
    m = dict(nx=int(d["nx"]), ny=int(d["ny"]), dh=float(d["dh"]), time=float(d["time"]),
             ttnorm=float(d["ttnorm"]), eps_adj=float(d["eps_adj"]), src_ix=d["src_ix"], src_iy=d["src_iy"],
             rec_ix=d["rec_ix"], rec_iy=d["rec_iy"], max_iter=100)
    with N.Context([m], rows_per_lane=R) as ctx:
        ctx.set_split(split)
        ctx.set_slowness(0, d["S"])
        ctx.set_picks(0, d["tobs"])
        ctx.set_tres_carry(0, d["tres0"])
        ctx.gradient()
        g = ctx.grad(0)
        assert bits_equal(g, d["grad"]), f"grad: {ndiff(g, d['grad'])} cells differ"
        assert np.array_equal(ctx.iterations(0)[0], d["it_f"]) and np.array_equal(ctx.iterations(0)[1], d["it_a"])
        ns = len(d["src_ix"])
        assert bits_equal(ctx.field(0, ns - 1, "T"), d["T_last"])
        assert bits_equal(ctx.field(0, ns - 1, "lambda"), d["lam_last"])


def test_poisoned_fields_budget_and_single_shot(ora, monkeypatch):
    from rajzel_amd import native as N
    monkeypatch.setenv("RAJZEL_AMD_POISON", "1")
    sv = W.edge_case(90, 150, nshots=5, seed=9, sigma=1.0, ttnorm=1e-4, rec_step=5)
    oracle_picks(ora, sv)
    print("[case] poisoned fields, HBM budget and single-shot entry points", flush=True)
    o = ora.grad_obj(sv.S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.tobs, sv.dh, sv.time, sv.ttnorm,
                     sv.eps_adj)
    with N.Context([model_dict(sv, max_iter=100)], rows_per_lane=1) as ctx:
        slot = ctx.slot_bytes
    with N.Context([model_dict(sv, max_iter=100)], rows_per_lane=1, hbm_budget=2 * slot) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.gradient()
        assert bits_equal(ctx.grad(0), o["grad"])
        T, tm, it = ctx.eikonal_shot(0, sv.src_ix[1], sv.src_iy[1])
        T_o, it_o = ora.eikonal(sv.S, sv.src_ix[1], sv.src_iy[1], sv.dh, sv.time, sv.ttnorm)
        assert it == it_o and bits_equal(T, T_o)
        lam, ita = ctx.eikonal_adj_shot(0, T_o, o["tres"][1])
        lam_o, ita_o = ora.eikonal_adj(T_o, sv.rec_ix, sv.rec_iy, o["tres"][1], sv.dh, sv.time, sv.ttnorm,
                                       sv.eps_adj)
        assert ita == ita_o and bits_equal(lam, lam_o), f"{ndiff(lam, lam_o)} cells differ"
