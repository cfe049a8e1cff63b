"""Several independent models of uneven grid size in one context: their
shots share every launch.  Each model's results must be bit for bit those of
the same model alone (and of the CPU restatement on small cases)."""
import numpy as np
import pytest

from rajzel_amd import workloads as W
from tests.helpers import bits_equal, model_dict, ndiff, oracle_picks

pytestmark = pytest.mark.gpu


def _native():
    from rajzel_amd import native as N
    return N


def _run(N, models, surveys, R=0):
    with N.Context([model_dict(sv, max_iter=200) for sv in models], rows_per_lane=R) as ctx:
        for m, sv in enumerate(surveys):
            ctx.set_slowness(m, sv.S)
            ctx.set_picks(m, sv.tobs)
        ctx.gradient()
        out = []
        for m in range(len(surveys)):
            it_f, it_a = ctx.iterations(m)
            out.append(dict(grad=ctx.grad(m), tmod=ctx.tmod(m), tres=ctx.tres(m), L2=ctx.misfit(m)[0],
                            it_f=it_f, it_a=it_a))
        return out


def test_uneven_models_match_oracle(ora):
    shapes = [(37, 53, 3), (70, 40, 2), (45, 120, 4), (130, 66, 2)]
    surveys = []
    for j, (ny, nx, ns) in enumerate(shapes):
        sv = W.small_case(nx=nx, ny=ny, nshots=ns, seed=60 + j, rec_row=(2, ny - 1))
        oracle_picks(ora, sv)
        surveys.append(sv)
    N = _native()
    print("[case] 4 uneven models in one context", flush=True)
    res = _run(N, surveys, surveys)
    for sv, r in zip(surveys, res):
        o = ora.grad_obj(sv.S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.tobs, sv.dh)
        assert np.array_equal(r["it_f"], o["it_f"]) and np.array_equal(r["it_a"], o["it_a"])
        assert bits_equal(r["tmod"], o["tmod"]) and bits_equal(r["tres"], o["tres"])
        assert bits_equal(r["grad"], o["grad"]), f"{sv.name}: {ndiff(r['grad'], o['grad'])} cells differ"
        assert r["L2"] == o["L2"]


def test_c3_uneven_batch_equals_models_alone():
    """The uneven-batch configuration (32 models drawn from 200x200 ...
    600x1600, 8 shots each) in one context against each model alone."""
    surveys = W.c3_instances()
    rng = np.random.default_rng(77)
    for sv in surveys:
        sv.tobs = np.full((sv.nshots, sv.ntr), 1.0, np.float32) + rng.normal(0, 0.05, (sv.nshots, sv.ntr)).astype(
            np.float32)
    N = _native()
    print("[case] C3 batch: 32 models x 8 shots in one context", flush=True)
    together = _run(N, surveys, surveys)
    for m in (0, 5, 11, 17, 31):
        print(f"[case] C3 model {m} alone", flush=True)
        alone = _run(N, [surveys[m]], [surveys[m]], R=0)[0]
        t = together[m]
        for k in ("grad", "tmod", "tres"):
            assert bits_equal(t[k], alone[k]), f"model {m} {k}: {ndiff(t[k], alone[k])} differ"
        assert t["L2"] == alone["L2"]
        assert np.array_equal(t["it_f"], alone["it_f"]) and np.array_equal(t["it_a"], alone["it_a"])
