"""GPU parity: the HIP path through the C ABI against the CPU restatement
(itself pinned to the compiled reference, tests/test_oracle_reference.py).

The bar is bitwise equality for every field, receiver time, residual,
gradient and iteration count (float32 path), and the double misfit.
"""
import numpy as np
import pytest

from rajzel_amd import workloads as W
from tests.helpers import bits_equal, model_dict, ndiff, oracle_picks

pytestmark = pytest.mark.gpu


def _native():
    from rajzel_amd import native as N
    return N


def _case(nx, ny, nshots, seed, rec_row=None, ntr=None):
    # receivers alternate between row 2 and a line near the bottom: the
    # adjoint field then spans the model
    rec_row = (2, ny - 1) if rec_row is None else rec_row
    return W.small_case(nx=nx, ny=ny, nshots=nshots, seed=seed, rec_row=rec_row, ntr=ntr)


@pytest.mark.parametrize("R", [1, 2, 4, 8])
def test_forward_small_bitwise(ora, R):
    sv = _case(53, 37, 3, seed=7)
    N = _native()
    print(f"[case] forward 37x53 R={R}", flush=True)
    with N.Context([model_dict(sv, max_iter=50)], rows_per_lane=R) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.forward()
        tm = ctx.tmod(0)
        tn = ctx.tnb(0)
        itf, _ = ctx.iterations(0)
        ry = np.asarray(sv.rec_iy)
        kn = np.where(ry <= sv.ny - 2, ry + 1, ry - 1)   # calc_FA_res.c:52-69 slope neighbour
        for s in range(sv.nshots):
            T_o, it_o = ora.eikonal(sv.S, sv.src_ix[s], sv.src_iy[s], sv.dh)
            T_g = ctx.field(0, s, "T")
            assert it_o == itf[s]
            assert bits_equal(T_g, T_o), f"shot {s}: {ndiff(T_g, T_o)} cells differ"
            assert bits_equal(tm[s], T_o[sv.rec_iy, sv.rec_ix])
            assert bits_equal(tn[s], T_o[kn, sv.rec_ix])


@pytest.mark.parametrize("R,shape", [(1, (150, 90)), (2, (300, 70)), (1, (64, 40)), (1, (65, 41)),
                                     (4, (520, 33)), (1, (200, 5)), (1, (512, 37)), (2, (1001, 29)),
                                     (8, (1100, 21))])
def test_forward_multiband_bitwise(ora, R, shape):
    ny, nx = shape
    sv = _case(nx, ny, 2, seed=11)
    N = _native()
    print(f"[case] forward multiband {ny}x{nx} R={R}", flush=True)
    with N.Context([model_dict(sv, max_iter=50)], rows_per_lane=R) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.forward()
        itf, _ = ctx.iterations(0)
        for s in range(sv.nshots):
            T_o, it_o = ora.eikonal(sv.S, sv.src_ix[s], sv.src_iy[s], sv.dh)
            T_g = ctx.field(0, s, "T")
            assert it_o == itf[s]
            assert bits_equal(T_g, T_o), f"shot {s}: {ndiff(T_g, T_o)} cells differ"


def _grad_case(ora, nx, ny, nshots, seed, mute=()):
    sv = _case(nx, ny, nshots, seed=seed)
    tobs = oracle_picks(ora, sv).copy()
    for (s, i) in mute:
        tobs[s, i] = -1.0
    sv.tobs = tobs
    return sv


def _check_grad(ora, sv, ctx, tres0=None):
    O = ora.grad_obj(sv.S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.tobs, sv.dh, sv.time,
                     sv.ttnorm, sv.eps_adj, tres0=tres0)
    g = ctx.grad(0)
    assert bits_equal(g, O["grad"]), f"grad: {ndiff(g, O['grad'])} cells differ"
    assert bits_equal(ctx.tmod(0), O["tmod"])
    assert bits_equal(ctx.tres(0), O["tres"])
    L2, l2s = ctx.misfit(0)
    assert L2 == O["L2"] and bits_equal(l2s, O["l2_shot"])
    itf, ita = ctx.iterations(0)
    assert np.array_equal(itf, O["it_f"]) and np.array_equal(ita, O["it_a"])
    assert bits_equal(ctx.tres_carry(0), O["tres_carry"])
    return O


@pytest.mark.parametrize("R", [1, 2, 4, 8])
def test_gradient_small_bitwise(ora, R):
    sv = _grad_case(ora, 53, 37, 4, seed=3)
    N = _native()
    print(f"[case] gradient 37x53 R={R}", flush=True)
    with N.Context([model_dict(sv, max_iter=50)], rows_per_lane=R) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.gradient()
        O = _check_grad(ora, sv, ctx)
        lam_last = ctx.field(0, sv.nshots - 1, "lambda")
        assert bits_equal(lam_last, O["lam"]), f"lambda: {ndiff(lam_last, O['lam'])} cells differ"
        T_last = ctx.field(0, sv.nshots - 1, "T")
        assert bits_equal(T_last, O["T"])


@pytest.mark.parametrize("R,ny", [(1, 150), (1, 500), (2, 700)])
def test_gradient_multiband_muted_carry(ora, R, ny):
    sv = _grad_case(ora, 70, ny, 5, seed=5, mute=[(0, 3), (1, 3), (2, 10), (4, 0)])
    N = _native()
    tres0 = np.linspace(-1e-3, 1e-3, sv.ntr).astype(np.float32)
    print(f"[case] gradient {ny}x70 R={R} muted", flush=True)
    with N.Context([model_dict(sv, max_iter=50)], rows_per_lane=R) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.set_tres_carry(0, tres0)
        ctx.gradient()
        _check_grad(ora, sv, ctx, tres0=tres0)


def test_per_shot_and_budget_equal_batched(ora):
    sv = _grad_case(ora, 60, 90, 6, seed=9)
    N = _native()
    print("[case] per-shot / budget vs batched", flush=True)
    with N.Context([model_dict(sv, max_iter=50)], rows_per_lane=1) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.gradient()
        g_b = ctx.grad(0)
        ctx.gradient(per_shot=True)
        g_p = ctx.grad(0)
        assert bits_equal(g_b, g_p)
        _check_grad(ora, sv, ctx)
    # a budget that fits two shots' fields: 3 batches
    R = 1
    with N.Context([model_dict(sv, max_iter=50)], rows_per_lane=R) as ctx:
        slot = ctx.slot_bytes
    with N.Context([model_dict(sv, max_iter=50)], rows_per_lane=R, hbm_budget=2 * slot) as ctx:
        assert ctx.batch_shots == 2
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.gradient()
        assert bits_equal(ctx.grad(0), g_b)
        _check_grad(ora, sv, ctx)
        ctx.gradient(per_shot=True)
        assert bits_equal(ctx.grad(0), g_b)


def test_c1_forward_bitwise(ora):
    sv = W.c1_forward()
    N = _native()
    print("[case] C1 forward 200x600", flush=True)
    with N.Context([model_dict(sv)]) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.forward()
        T_o, it_o = ora.eikonal(sv.S, sv.src_ix[0], sv.src_iy[0], sv.dh)
        T_g = ctx.field(0, 0, "T")
        assert ctx.iterations(0)[0][0] == it_o
        assert bits_equal(T_g, T_o), f"{ndiff(T_g, T_o)} cells differ"
        assert bits_equal(ctx.tmod(0)[0], T_o[sv.rec_iy, sv.rec_ix])


def test_single_shot_entry_points(ora):
    sv = _grad_case(ora, 48, 70, 2, seed=21)
    N = _native()
    print("[case] single-shot entry points", flush=True)
    with N.Context([model_dict(sv, max_iter=50)], rows_per_lane=1) as ctx:
        ctx.set_slowness(0, sv.S)
        T_g, tm_g, it_g = ctx.eikonal_shot(0, sv.src_ix[1], sv.src_iy[1])
        T_o, it_o = ora.eikonal(sv.S, sv.src_ix[1], sv.src_iy[1], sv.dh)
        assert it_g == it_o and bits_equal(T_g, T_o)
        tres = (np.arange(sv.ntr, dtype=np.float32) - sv.ntr / 2) * np.float32(1e-4)
        lam_g, ita = ctx.eikonal_adj_shot(0, T_o, tres)
        lam_o, ita_o = ora.eikonal_adj(T_o, sv.rec_ix, sv.rec_iy, tres, sv.dh)
        assert ita == ita_o and bits_equal(lam_g, lam_o), f"{ndiff(lam_g, lam_o)} cells differ"
        g0 = np.full(T_o.shape, 0.25, np.float32)
        gs_g, g_g = ctx.ass_grad_shot(0, lam_o, g0)
        gs_o, g_o = ora.ass_grad(sv.S, lam_o, g0)
        assert bits_equal(gs_g, gs_o) and bits_equal(g_g, g_o)


def test_single_shot_adjoint_negative_zero_traveltime(ora):
    """rz_eikonal_adj_shot with T = -0.0 at the source (an interior cell,
    not a receiver): the reference treats -0 as +0 and updates the cell, so
    lambda equals the +0 run's bit for bit."""
    sv = _grad_case(ora, 48, 70, 1, seed=22)
    N = _native()
    print("[case] single-shot adjoint with -0.0 in T", flush=True)
    sy, sx = int(sv.src_iy[0]) + 20, int(sv.src_ix[0])   # a source inside the grid
    T, _ = ora.eikonal(sv.S, sx, sy, sv.dh)
    assert T[sy, sx] == 0 and 0 < sy < sv.ny - 1 and 0 < sx < sv.nx - 1
    Tn = T.copy()
    Tn[sy, sx] = np.float32(-0.0)
    tres = (np.arange(sv.ntr, dtype=np.float32) - sv.ntr / 2) * np.float32(1e-4)
    with N.Context([model_dict(sv, max_iter=50)], rows_per_lane=1) as ctx:
        ctx.set_slowness(0, sv.S)
        lam_p, it_p = ctx.eikonal_adj_shot(0, T, tres)
        lam_n, it_n = ctx.eikonal_adj_shot(0, Tn, tres)
    lam_o, it_o = ora.eikonal_adj(T, sv.rec_ix, sv.rec_iy, tres, sv.dh)
    assert it_p == it_n == it_o
    assert bits_equal(lam_n, lam_p) and bits_equal(lam_p, lam_o)


def test_adjoint_inf_guard(ora):
    """sweep_adj.c:68: lambda = 0 next to traveltimes >= INF (9999999)."""
    sv = _grad_case(ora, 40, 30, 1, seed=23)
    N = _native()
    print("[case] adjoint INF guard", flush=True)
    T, _ = ora.eikonal(sv.S, sv.src_ix[0], sv.src_iy[0], sv.dh)
    T = T.copy()
    T[10:14, 20:25] = np.float32(1.5e7)
    T[25, 5] = np.float32(9999999.0)
    tres = np.linspace(-2e-3, 2e-3, sv.ntr).astype(np.float32)
    with N.Context([model_dict(sv, max_iter=50)], rows_per_lane=1) as ctx:
        ctx.set_slowness(0, sv.S)
        lam_g, it_g = ctx.eikonal_adj_shot(0, T, tres)
    lam_o, it_o = \
        ora.eikonal_adj(T, sv.rec_ix, sv.rec_iy, tres, sv.dh, maxit=50)
    assert it_g == it_o and bits_equal(lam_g, lam_o), f"{ndiff(lam_g, lam_o)} cells differ"
    assert np.any(lam_o[9:15, 19:26] == 0)
