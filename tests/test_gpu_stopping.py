"""The stopping test of eikonal.c:44 / eikonal_adj.c:37 (norm1.c:17-31) on
the GPU, where it ends solves before a fixed point: iteration counts,
fields, residuals and the gradient against the CPU restatement (whose
norm1 is pinned to the reference build), bit for bit.

The kernels decide from their double sum of per-sweep decreases where a
rigorous bound allows, else from a tighter bound given the number of cells
that changed, and otherwise take norm1's float sum in the reference's
order; RAJZEL_AMD_ORDERED_NORM=1 makes every nonzero decision take the
ordered sum and =2 every one go through the count, so all three ways are
checked on the same inputs.  Split
shots (two workgroups per shot) and one-workgroup shots are both run.
"""
import numpy as np
import pytest

from rajzel_amd import workloads as W
from tests.helpers import bits_equal, model_dict, ndiff, oracle_picks

pytestmark = pytest.mark.gpu

CASES = [  # (NX, NY, sigma, TTNORM, R, rec_step)
    (600, 200, 2.0, 1e-3, 2, 9),
    (600, 200, 3.0, 1e-5, 4, 9),
    (90, 150, 1.0, 1e-4, 1, 5),
    (29, 1001, 1.0, 1e-5, 4, 3),
    (21, 1100, 2.0, 1e-7, 8, 3),
]


@pytest.mark.parametrize("ordered", [0, 1, 2])
@pytest.mark.parametrize("split", [True, False])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[1]}x{c[0]}-tt{c[3]:g}-R{c[4]}")
def test_early_stop_evaluation_bitwise(ora, case, split, ordered, monkeypatch):
    from rajzel_amd import native as N
    nx, ny, sigma, ttn, R, step = case
    sv = W.edge_case(nx, ny, nshots=3, seed=nx + ny, sigma=sigma, ttnorm=ttn, rec_step=step)
    oracle_picks(ora, sv)
    sv.tobs[1, 2] = -1.0                                    # a muted pick (carry)
    print(f"[case] stopping {ny}x{nx} ttnorm={ttn} R={R} split={split} ordered={ordered}", flush=True)
    monkeypatch.setenv("RAJZEL_AMD_ORDERED_NORM", str(int(ordered)))
    o = ora.grad_obj(sv.S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.tobs, sv.dh, sv.time, sv.ttnorm,
                     sv.eps_adj)
    with N.Context([model_dict(sv, max_iter=100)], rows_per_lane=R) as ctx:
        ctx.set_split(split)
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.gradient()
        itf, ita = ctx.iterations(0)
        assert np.array_equal(itf, o["it_f"]) and np.array_equal(ita, o["it_a"]), (itf, ita, o["it_f"], o["it_a"])
        g = ctx.grad(0)
        assert bits_equal(g, o["grad"]), f"grad: {ndiff(g, o['grad'])} cells differ"
        assert bits_equal(ctx.tmod(0), o["tmod"]) and bits_equal(ctx.tres(0), o["tres"])
        L2, l2s = ctx.misfit(0)
        assert L2 == o["L2"] and bits_equal(l2s, o["l2_shot"])
        assert bits_equal(ctx.field(0, 2, "T"), o["T"])
        assert bits_equal(ctx.field(0, 2, "lambda"), o["lam"])
        assert ctx.split_state()[1] == 0
        sp = ctx.stop_passes(0)
        if ordered == 1:   # every decision after a nonzero change took the ordered sum
            assert sp["ordered_f"].sum() > 0 and sp["ordered_a"].sum() > 0
        if ordered == 2:   # ... counted the changed cells
            assert sp["count_f"].sum() > 0 and sp["count_a"].sum() > 0


def _norm1_sum(a, b):
    """norm1.c:21-25: float sum of |a - b| (float differences), x outer."""
    d = np.abs(np.asarray(a, np.float32) - np.asarray(b, np.float32)).T.ravel()
    return np.add.accumulate(d, dtype=np.float32)[-1]


@pytest.mark.parametrize("split", [True, False])
@pytest.mark.parametrize("case", CASES[:4], ids=lambda c: f"{c[1]}x{c[0]}-tt{c[3]:g}-R{c[4]}")
def test_ordered_sum_bits(ora, case, split, monkeypatch):
    """With every decision taking the ordered sum, the sum of each shot's last
    iteration equals the sequential float32 sum of |T_k - T_(k-1)| (and of
    |lambda_k - lambda_(k-1)|) over the restatement's iterates, bit for bit."""
    from rajzel_amd import native as N
    nx, ny, sigma, ttn, R, step = case
    sv = W.edge_case(nx, ny, nshots=3, seed=nx + ny, sigma=sigma, ttnorm=ttn, rec_step=step)
    oracle_picks(ora, sv)
    print(f"[case] ordered sum bits {ny}x{nx} ttnorm={ttn} R={R} split={split}", flush=True)
    monkeypatch.setenv("RAJZEL_AMD_ORDERED_NORM", "1")
    o = ora.grad_obj(sv.S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.tobs, sv.dh, sv.time, sv.ttnorm,
                     sv.eps_adj)
    with N.Context([model_dict(sv, max_iter=100)], rows_per_lane=R) as ctx:
        ctx.set_split(split)
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.gradient()
        itf, ita = ctx.iterations(0)
        sp = ctx.stop_passes(0)
    checked = 0
    for s in range(sv.nshots):
        kf, ka = int(itf[s]), int(ita[s])
        T1, _ = ora.eikonal(sv.S, sv.src_ix[s], sv.src_iy[s], sv.dh, sv.time, sv.ttnorm, maxit=kf)
        if kf > 1:
            T0, _ = ora.eikonal(sv.S, sv.src_ix[s], sv.src_iy[s], sv.dh, sv.time, sv.ttnorm, maxit=kf - 1)
        else:
            T0 = np.full_like(T1, np.float32(2.0 * sv.time))
            T0[sv.src_iy[s], sv.src_ix[s]] = 0
        want = _norm1_sum(T1, T0)
        if want > 0:   # a zero sum is decided without the ordered sum
            assert np.float32(sp["sum_f"][s]).view(np.uint32) == want.view(np.uint32), (s, sp["sum_f"][s], want)
            checked += 1
        L1, _ = ora.eikonal_adj(T1, sv.rec_ix, sv.rec_iy, o["tres"][s], sv.dh, sv.time, sv.ttnorm, sv.eps_adj,
                                maxit=ka)
        if ka > 1:
            L0, _ = ora.eikonal_adj(T1, sv.rec_ix, sv.rec_iy, o["tres"][s], sv.dh, sv.time, sv.ttnorm,
                                    sv.eps_adj, maxit=ka - 1)
        else:
            L0 = ora.init_travel_adj(sv.nx, sv.ny, sv.time, sv.rec_ix, sv.rec_iy, o["tres"][s])
        want = _norm1_sum(L1, L0)
        if want > 0:
            assert np.float32(sp["sum_a"][s]).view(np.uint32) == want.view(np.uint32), (s, sp["sum_a"][s], want)
            checked += 1
    assert checked > 0


def test_early_stop_forward_only(ora, monkeypatch):
    """rz_forward alone (no adjoint) with solves ending in either A buffer."""
    from rajzel_amd import native as N
    sv = W.edge_case(600, 200, nshots=4, seed=5, sigma=2.0, ttnorm=1e-4, rec_step=9)
    print("[case] stopping forward only 200x600", flush=True)
    with N.Context([model_dict(sv, max_iter=100)], rows_per_lane=2) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.forward()
        itf, _ = ctx.iterations(0)
        for s in range(sv.nshots):
            T_o, it_o = ora.eikonal(sv.S, sv.src_ix[s], sv.src_iy[s], sv.dh, sv.time, sv.ttnorm)
            assert it_o == itf[s]
            T_g = ctx.field(0, s, "T")
            assert bits_equal(T_g, T_o), f"shot {s}: {ndiff(T_g, T_o)} cells differ"
    assert set(int(i) % 2 for i in itf) == {0, 1}   # converged T in TA and in LA
