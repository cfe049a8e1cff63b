"""Evaluations replayed as HIP graphs (rz_set_graph) against the same
evaluations launched one by one: every replay, with eager work and new
inputs (picks, slowness) between replays, must give the same bits."""
import numpy as np
import pytest

from rajzel_amd import workloads as W
from tests.helpers import bits_equal, model_dict, oracle_picks

pytestmark = pytest.mark.gpu


def test_graph_replays_equal_eager(ora):
    from rajzel_amd import native as N
    sv = W.small_case(nx=90, ny=60, nshots=4, seed=31, rec_row=(2, 59))
    oracle_picks(ora, sv)
    rng = np.random.default_rng(5)
    models = [sv.S, (sv.S * np.float32(1.01)).astype(np.float32), sv.S]
    picks = [sv.tobs, (sv.tobs + rng.normal(0, 1e-3, sv.tobs.shape)).astype(np.float32), sv.tobs]
    print("[case] graph replays vs eager launches", flush=True)
    with N.Context([model_dict(sv, max_iter=100)]) as eager, N.Context([model_dict(sv, max_iter=100)]) as g:
        g.set_graph(True)
        for rep in range(4):
            S, tobs = models[rep % 3], picks[rep % 3]
            for ctx in (eager, g):
                ctx.set_slowness(0, S)
                ctx.set_picks(0, tobs)
                ctx.gradient()
            assert bits_equal(g.grad(0), eager.grad(0)), rep
            assert bits_equal(g.tmod(0), eager.tmod(0)) and bits_equal(g.tres(0), eager.tres(0))
            assert g.misfit(0)[0] == eager.misfit(0)[0]
            eager.forward()
            g.forward()
            assert bits_equal(g.tmod(0), eager.tmod(0)), rep
        o = ora.grad_obj(sv.S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.tobs, sv.dh)
        g.set_slowness(0, sv.S)
        g.set_picks(0, sv.tobs)
        g.gradient()
        assert bits_equal(g.grad(0), o["grad"])
