"""Sources at the grid's edges and around epoch boundaries, on the GPU.

The first forward sweep only stores the initial value in the epochs of a
band whose keys all lie left of the source's column (those cells keep it),
so where that part ends depends on the source column and on the band's
offset in the pipeline.  Shots at the first and last columns, at columns
around multiples of the eight-step epoch and of the 64-lane skew, in the
first, a middle and the last row, on grids of two to four bands: receiver
times, iteration counts and the traveltime field of every shot against the
CPU restatement, bit for bit, one-workgroup and split.
"""
import numpy as np
import pytest

from rajzel_amd import workloads as W
from tests.helpers import bits_equal, model_dict, ndiff

pytestmark = pytest.mark.gpu

CASES = [  # (NX, NY, R)
    (300, 300, 4),   # two bands
    (260, 330, 2),   # three bands
    (200, 700, 4),   # three bands
]


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[1]}x{c[0]}-R{c[2]}")
def test_forward_with_sources_at_edge_columns(ora, case, split):
    from rajzel_amd import native as N
    nx, ny, R = case
    sv = W.edge_case(nx, ny, nshots=3, seed=nx + ny, sigma=3.0, ttnorm=1e-6)
    cols = [0, 1, 2, 7, 8, 9, 63, 64, 65, 71, 72, 73, nx // 2, nx - 3, nx - 2, nx - 1]
    rows = [0, ny // 2, ny - 1]
    src_ix = np.array([c for c in cols for _ in rows], np.int32)
    src_iy = np.array([r for _ in cols for r in rows], np.int32)
    print(f"[case] source columns {ny}x{nx} R={R} split={split} shots={src_ix.size}", flush=True)
    tm_o, _, it_o = ora.forward(sv.S, src_ix, src_iy, sv.rec_ix, sv.rec_iy, sv.dh, sv.time, sv.ttnorm)
    md = model_dict(sv, max_iter=200)
    md.update(src_ix=src_ix, src_iy=src_iy)
    with N.Context([md], rows_per_lane=R) as ctx:
        ctx.set_split(split)
        ctx.set_slowness(0, sv.S)
        ctx.forward()
        itf, _ = ctx.iterations(0)
        assert np.array_equal(itf, it_o), (itf, it_o)
        assert bits_equal(ctx.tmod(0), tm_o)
        for s in range(src_ix.size):
            T_o, _ = ora.eikonal(sv.S, src_ix[s], src_iy[s], sv.dh, sv.time, sv.ttnorm)
            T = ctx.field(0, s, "T")
            assert bits_equal(T, T_o), f"source ({src_iy[s]}, {src_ix[s]}): {ndiff(T, T_o)} cells differ"
        assert ctx.split_state()[1] == 0
