"""Tall grids: 2049 to 4096 rows need 5 to 8 bands of 512 rows (8 rows per
lane), more than the two sweep groups of a workgroup hold at 4 bands each.
Such grids run unpaired: the workgroup's 8 waves form one group, one wave
per band, and the four sweeps of an iteration (eikonal.c:50-53,
eikonal_adj.c:43-46) run one after the other.  Bitwise against the
restatement, as every other grid."""
import numpy as np
import pytest

from rajzel_amd import workloads as W
from tests.helpers import bits_equal, model_dict, ndiff, oracle_picks
from tests.test_gpu_parity import _check_grad

pytestmark = pytest.mark.gpu


def _tall(ora, nx, ny, nshots, seed, mute=()):
    # receivers on row 2 and on the row above the bottom, so that the
    # adjoint field spans every band
    sv = W.small_case(nx=nx, ny=ny, nshots=nshots, seed=seed, rec_row=(2, ny - 1))
    tobs = oracle_picks(ora, sv).copy()
    for (s, i) in mute:
        tobs[s, i] = -1.0
    sv.tobs = tobs
    return sv


@pytest.mark.parametrize("ny,nx", [(2049, 23), (3000, 40), (4096, 17)])
def test_tall_gradient_bitwise(ora, ny, nx):
    from rajzel_amd import native as N
    sv = _tall(ora, nx, ny, 3, seed=41, mute=[(1, 2)])
    print(f"[case] tall gradient {ny}x{nx}", flush=True)
    with N.Context([model_dict(sv, max_iter=50)]) as ctx:
        assert ctx.rows_per_lane == 8
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.gradient()
        O = _check_grad(ora, sv, ctx)
        T_last = ctx.field(0, sv.nshots - 1, "T")
        assert bits_equal(T_last, O["T"]), f"T: {ndiff(T_last, O['T'])} cells differ"
        lam_last = ctx.field(0, sv.nshots - 1, "lambda")
        assert bits_equal(lam_last, O["lam"]), f"lambda: {ndiff(lam_last, O['lam'])} cells differ"
        for s in range(sv.nshots - 1):
            T_o, it_o = ora.eikonal(sv.S, sv.src_ix[s], sv.src_iy[s], sv.dh)
            assert bits_equal(ctx.field(0, s, "T"), T_o)


def test_tall_beside_a_paired_model(ora):
    """One launch with a tall model (6 bands) and a short one (1 band): the
    launch runs unpaired for both, and each equals its own restatement."""
    from rajzel_amd import native as N
    tall = _tall(ora, 31, 3000, 2, seed=43)
    short = _tall(ora, 53, 37, 3, seed=44)
    print("[case] tall beside short", flush=True)
    with N.Context([model_dict(tall, max_iter=50), model_dict(short, max_iter=50)]) as ctx:
        for m, sv in enumerate((tall, short)):
            ctx.set_slowness(m, sv.S)
            ctx.set_picks(m, sv.tobs)
        ctx.gradient()
        for m, sv in enumerate((tall, short)):
            O = ora.grad_obj(sv.S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.tobs, sv.dh, sv.time,
                             sv.ttnorm, sv.eps_adj)
            g = ctx.grad(m)
            assert bits_equal(g, O["grad"]), f"model {m}: {ndiff(g, O['grad'])} cells differ"
            assert bits_equal(ctx.tmod(m), O["tmod"])
            itf, ita = ctx.iterations(m)
            assert np.array_equal(itf, O["it_f"]) and np.array_equal(ita, O["it_a"])


def test_tall_single_shot_entry_points_and_autotune(ora):
    from rajzel_amd import native as N
    sv = _tall(ora, 19, 2600, 1, seed=45)
    print("[case] tall single shot", flush=True)
    with N.Context([model_dict(sv, max_iter=50)], rows_per_lane=N.ROWS_AUTOTUNE) as ctx:
        assert ctx.rows_per_lane == 8          # the only setting that holds 2600 rows
        ctx.set_slowness(0, sv.S)
        T_g, tm_g, it_g = ctx.eikonal_shot(0, sv.src_ix[0], sv.src_iy[0])
        T_o, it_o = ora.eikonal(sv.S, sv.src_ix[0], sv.src_iy[0], sv.dh)
        assert it_g == it_o and bits_equal(T_g, T_o), f"{ndiff(T_g, T_o)} cells differ"
        tres = (np.arange(sv.ntr, dtype=np.float32) - sv.ntr / 2) * np.float32(1e-4)
        lam_g, ita = ctx.eikonal_adj_shot(0, T_o, tres)
        lam_o, ita_o = ora.eikonal_adj(T_o, sv.rec_ix, sv.rec_iy, tres, sv.dh)
        assert ita == ita_o and bits_equal(lam_g, lam_o), f"{ndiff(lam_g, lam_o)} cells differ"


# Grids taller than a workgroup holds (NY > 4096) whose NX fits are stored
# and solved transposed: rows as stored are the grid's columns, the sweep
# directions are transposed, norm1's sum keeps the grid's order and every
# field, gradient and receiver time is returned in the grid's orientation.

@pytest.mark.parametrize("ny,nx", [(5000, 23), (10000, 12)])
def test_transposed_gradient_bitwise(ora, ny, nx):
    from rajzel_amd import native as N
    sv = _tall(ora, nx, ny, 3, seed=47, mute=[(0, 5), (2, 1)])
    print(f"[case] transposed gradient {ny}x{nx}", flush=True)
    with N.Context([model_dict(sv, max_iter=50)]) as ctx:
        assert ctx.rows_per_lane == 8
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.gradient()
        O = _check_grad(ora, sv, ctx)
        T_last = ctx.field(0, sv.nshots - 1, "T")
        assert bits_equal(T_last, O["T"]), f"T: {ndiff(T_last, O['T'])} cells differ"
        lam_last = ctx.field(0, sv.nshots - 1, "lambda")
        assert bits_equal(lam_last, O["lam"]), f"lambda: {ndiff(lam_last, O['lam'])} cells differ"


def test_transposed_ordered_norm_and_early_stop(ora, monkeypatch):
    """Every stopping decision that is not zero through norm1's float sum in
    the grid's order (RAJZEL_AMD_ORDERED_NORM=1), and a TTNORM that stops
    the solves before a fixed point."""
    from rajzel_amd import native as N
    monkeypatch.setenv("RAJZEL_AMD_ORDERED_NORM", "1")
    sv = _tall(ora, 17, 4500, 2, seed=48)
    sv.ttnorm = 1e-6
    sv.tobs = oracle_picks(ora, sv)
    print("[case] transposed, ordered norm, TTNORM 1e-6", flush=True)
    with N.Context([model_dict(sv, max_iter=50)]) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.gradient()
        _check_grad(ora, sv, ctx)
        assert sum(v for k, v in ctx.stop_passes(0).items() if k.startswith("ordered")).sum() > 0


def test_transposed_beside_a_stored_model_and_shot_taper(ora):
    """A transposed model and an ordinary one in one launch, the per-shot
    radial taper (SWS_TAPER_CIRCULAR_PER_SHOT = 1) on the transposed one."""
    from rajzel_amd import native as N
    tall = _tall(ora, 21, 4200, 3, seed=49)
    short = _tall(ora, 40, 30, 2, seed=50)
    print("[case] transposed beside stored, taper", flush=True)
    with N.Context([model_dict(tall, max_iter=50), model_dict(short, max_iter=50)]) as ctx:
        for m, sv in enumerate((tall, short)):
            ctx.set_slowness(m, sv.S)
            ctx.set_picks(m, sv.tobs)
        ctx.set_shot_taper(0, 1, radius_m=30.0, src_x_m=tall.src_x_m, src_z_m=tall.src_z_m)
        ctx.gradient()
        ora.set_shot_taper(1, radius=30.0, src_x_m=tall.src_x_m, src_z_m=tall.src_z_m)
        try:
            O = ora.grad_obj(tall.S, tall.src_ix, tall.src_iy, tall.rec_ix, tall.rec_iy, tall.tobs, tall.dh,
                             tall.time, tall.ttnorm, tall.eps_adj)
        finally:
            ora.set_shot_taper(0)
        g = ctx.grad(0)
        assert bits_equal(g, O["grad"]), f"tapered transposed: {ndiff(g, O['grad'])} cells differ"
        O2 = ora.grad_obj(short.S, short.src_ix, short.src_iy, short.rec_ix, short.rec_iy, short.tobs, short.dh,
                          short.time, short.ttnorm, short.eps_adj)
        assert bits_equal(ctx.grad(1), O2["grad"])


def test_transposed_single_shot_entry_points(ora):
    from rajzel_amd import native as N
    sv = _tall(ora, 14, 6000, 1, seed=51)
    print("[case] transposed single shot", flush=True)
    with N.Context([model_dict(sv, max_iter=50)]) as ctx:
        ctx.set_slowness(0, sv.S)
        T_g, tm_g, it_g = ctx.eikonal_shot(0, sv.src_ix[0], sv.src_iy[0])
        T_o, it_o = ora.eikonal(sv.S, sv.src_ix[0], sv.src_iy[0], sv.dh)
        assert it_g == it_o and bits_equal(T_g, T_o), f"{ndiff(T_g, T_o)} cells differ"
        assert bits_equal(tm_g, T_o[sv.rec_iy, sv.rec_ix])
        tres = (np.arange(sv.ntr, dtype=np.float32) - sv.ntr / 2) * np.float32(1e-4)
        lam_g, ita = ctx.eikonal_adj_shot(0, T_o, tres)
        lam_o, ita_o = ora.eikonal_adj(T_o, sv.rec_ix, sv.rec_iy, tres, sv.dh)
        assert ita == ita_o and bits_equal(lam_g, lam_o), f"{ndiff(lam_g, lam_o)} cells differ"
        g0 = np.full(T_o.shape, 0.25, np.float32)
        gs_g, g_g = ctx.ass_grad_shot(0, lam_o, g0)
        gs_o, g_o = ora.ass_grad(sv.S, lam_o, g0)
        assert bits_equal(gs_g, gs_o) and bits_equal(g_g, g_o)
