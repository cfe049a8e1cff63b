"""The rows-per-lane autotuner: it picks a feasible setting, and a context
created with it gives the same bits as the CPU restatement (the results of
the path do not depend on the tile height)."""
import numpy as np
import pytest

from rajzel_amd import workloads as W
from tests.helpers import bits_equal, model_dict, ndiff, oracle_picks

pytestmark = pytest.mark.gpu


def test_autotuned_context_bitwise(ora):
    from rajzel_amd import native as N
    sv = W.small_case(nx=70, ny=300, nshots=3, seed=21, rec_row=(2, 299))
    oracle_picks(ora, sv)
    models = [model_dict(sv, max_iter=200)]
    print("[case] autotune 300x70, 3 shots", flush=True)
    R, ms = N.autotune_rows(models)
    assert R in (1, 2, 4, 8) and ms > 0.0
    with N.Context(models, rows_per_lane=N.ROWS_AUTOTUNE) as ctx:
        assert ctx.rows_per_lane in (1, 2, 4, 8)
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.gradient()
        grad, L2 = ctx.grad(0), ctx.misfit(0)[0]
    o = ora.grad_obj(sv.S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.tobs, sv.dh)
    assert bits_equal(grad, o["grad"]), f"{ndiff(grad, o['grad'])} cells differ"
    assert L2 == o["L2"]
