"""Forward solves that reach a fixed point, on the GPU.

A forward sweep that changes no cell leaves every later sweep nothing to
change (a cell's update is min(T, G) with G a function of its neighbours
alone), so once a sweep of an outer iteration changed no cell the kernels
skip the sweeps left and count the iteration that would follow it (norm1 =
0 ends the loop, eikonal.c:44) without running it.  Here every solve runs to
its fixed point (TTNORM = 1e-10) under every iteration limit from 1 to one
past the reference's count, so that the counted iteration also meets the
limit check at the loop's head (-1): iteration counts, receiver times and
fields against the CPU restatement, bit for bit, with one-workgroup and with
split shots.
"""
import numpy as np
import pytest

from rajzel_amd import workloads as W
from tests.helpers import bits_equal, model_dict, ndiff

pytestmark = pytest.mark.gpu

CASES = [  # (NX, NY, sigma, R): paired sweeps
    (300, 120, 2.0, 2),
    (400, 150, 6.0, 4),
    (64, 700, 1.0, 4),
]
# five bands of 512 rows (one wave group, sweeps one after the other), and a
# grid of more than 4096 rows, solved transposed
TALL = [(48, 2300, 3.0, 8), (20, 4500, 1.0, 8)]


def _check(ora, sv, R, split, limits):
    from rajzel_amd import native as N
    for maxit in limits:
        print(f"[case] fixed point {sv.ny}x{sv.nx} R={R} split={split} max_iter={maxit}", flush=True)
        tm_o, _, it_o = ora.forward(sv.S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.dh, sv.time,
                                    sv.ttnorm, maxit)
        with N.Context([model_dict(sv, max_iter=maxit)], rows_per_lane=R) as ctx:
            ctx.set_split(split)
            ctx.set_slowness(0, sv.S)
            if (it_o < 0).any():
                with pytest.raises(N.NotConverged):
                    ctx.forward()
            else:
                ctx.forward()
            itf, _ = ctx.iterations(0)
            assert np.array_equal(itf, it_o), (maxit, itf, it_o)
            assert bits_equal(ctx.tmod(0), tm_o), maxit
            for s in range(sv.nshots):
                T_o, _ = ora.eikonal(sv.S, sv.src_ix[s], sv.src_iy[s], sv.dh, sv.time, sv.ttnorm, maxit)
                T = ctx.field(0, s, "T")
                assert bits_equal(T, T_o), f"max_iter {maxit} shot {s}: {ndiff(T, T_o)} cells differ"
            assert ctx.split_state()[1] == 0


def _survey(nx, ny, sigma):
    return W.edge_case(nx, ny, nshots=3, seed=nx + 3 * ny, sigma=sigma, ttnorm=1e-10)


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[1]}x{c[0]}-s{c[2]:g}-R{c[3]}")
def test_forward_iteration_limits_around_the_fixed_point(ora, case, split):
    nx, ny, sigma, R = case
    sv = _survey(nx, ny, sigma)
    _, _, it_ref = ora.forward(sv.S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.dh, sv.time, sv.ttnorm)
    assert (it_ref > 1).all(), it_ref
    _check(ora, sv, R, split, range(1, int(it_ref.max()) + 2))


@pytest.mark.parametrize("case", TALL, ids=lambda c: f"{c[1]}x{c[0]}-s{c[2]:g}-R{c[3]}")
def test_forward_fixed_point_one_wave_group(ora, case):
    """The same with the sweeps run one after the other (more bands than a
    pair of wave groups holds), at each shot's own count and one either
    side of it."""
    nx, ny, sigma, R = case
    sv = _survey(nx, ny, sigma)
    _, _, it_ref = ora.forward(sv.S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.dh, sv.time, sv.ttnorm)
    limits = sorted({int(i) + d for i in it_ref for d in (-1, 0, 1)} | {0})
    _check(ora, sv, R, False, limits)
