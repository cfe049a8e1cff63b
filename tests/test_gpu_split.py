"""Split shots (two workgroups per shot, rz_set_split): the same bits as one
workgroup per shot and as the CPU restatement, and the fallback when a
split launch reports failure."""
import numpy as np
import pytest

from rajzel_amd import workloads as W
from tests.helpers import bits_equal, model_dict, ndiff, oracle_picks

pytestmark = pytest.mark.gpu


def _run(N, sv, R, split):
    with N.Context([model_dict(sv, max_iter=200)], rows_per_lane=R) as ctx:
        ctx.set_split(split)
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.gradient()
        itf, ita = ctx.iterations(0)
        out = dict(grad=ctx.grad(0), tmod=ctx.tmod(0), tres=ctx.tres(0), L2=ctx.misfit(0)[0], itf=itf, ita=ita,
                   state=ctx.split_state())
        if split == 2:
            ctx.gradient()          # stays unsplit after the fallback
            out["again"] = ctx.grad(0)
            out["state2"] = ctx.split_state()
    return out


@pytest.mark.parametrize("R,shape,ns", [(1, (37, 53), 3), (2, (150, 70), 4), (4, (300, 90), 5), (8, (700, 40), 3),
                                        (1, (200, 41), 9)])
def test_split_equals_one_workgroup_and_oracle(ora, R, shape, ns):
    from rajzel_amd import native as N
    ny, nx = shape
    sv = W.small_case(nx=nx, ny=ny, nshots=ns, seed=90 + ny, rec_row=(2, ny - 1))
    oracle_picks(ora, sv)
    print(f"[case] split {ny}x{nx} R={R} shots={ns}", flush=True)
    a = _run(N, sv, R, True)
    assert a["state"] == (1, 0)
    b = _run(N, sv, R, False)
    for k in ("grad", "tmod", "tres"):
        assert bits_equal(a[k], b[k]), f"{k}: {ndiff(a[k], b[k])} differ"
    assert a["L2"] == b["L2"]
    assert np.array_equal(a["itf"], b["itf"]) and np.array_equal(a["ita"], b["ita"])
    if ny * nx <= 40000:
        o = ora.grad_obj(sv.S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.tobs, sv.dh)
        assert bits_equal(a["grad"], o["grad"]) and a["L2"] == o["L2"]


def test_failed_split_launch_falls_back():
    """The next split launch starts with its failure flag raised (forced from
    the host): the context repeats the work unsplit, reports failure 2 and
    stays unsplit; the results are those of one workgroup per shot."""
    from rajzel_amd import native as N
    sv = W.small_case(nx=64, ny=130, nshots=4, seed=5, rec_row=(2, 129))
    sv.tobs = np.full((sv.nshots, sv.ntr), 0.3, np.float32)
    print("[case] split fallback 130x64", flush=True)
    a = _run(N, sv, 2, 2)
    b = _run(N, sv, 2, False)
    assert a["state"] == (0, 2) and a["state2"] == (0, 2)
    assert bits_equal(a["grad"], b["grad"]) and bits_equal(a["again"], b["grad"])
    assert np.array_equal(a["itf"], b["itf"]) and np.array_equal(a["ita"], b["ita"])


def test_mixed_launch_after_history_equals_one_workgroup():
    """More shots than half the CUs: the second evaluation splits the shots
    that needed the most iterations in the first and runs the others
    beside them; every result equals the one-workgroup run's."""
    from rajzel_amd import native as N
    sv = W.small_case(nx=90, ny=150, nshots=160, seed=12, rec_row=(2, 149))
    rng = np.random.default_rng(3)
    sv.tobs = (0.2 + rng.uniform(0.0, 0.3, (sv.nshots, sv.ntr))).astype(np.float32)
    print("[case] mixed launch 150x90, 160 shots", flush=True)
    res = {}
    for split in (False, True):
        with N.Context([model_dict(sv, max_iter=200)], rows_per_lane=2) as ctx:
            ctx.set_split(split)
            ctx.set_slowness(0, sv.S)
            ctx.set_picks(0, sv.tobs)
            ctx.gradient()
            first = ctx.grad(0)
            ctx.gradient()
            itf, ita = ctx.iterations(0)
            res[split] = dict(first=first, grad=ctx.grad(0), tmod=ctx.tmod(0), tres=ctx.tres(0),
                              L2=ctx.misfit(0)[0], itf=itf, ita=ita, state=ctx.split_state())
    a, b = res[True], res[False]
    assert a["state"][1] == 0
    for k in ("first", "grad", "tmod", "tres"):
        assert bits_equal(a[k], b[k]), f"{k}: {ndiff(a[k], b[k])} differ"
    assert a["L2"] == b["L2"]
    assert np.array_equal(a["itf"], b["itf"]) and np.array_equal(a["ita"], b["ita"])


def test_mixed_launch_replayed_as_graph():
    """Mixed launches (side stream fork and join) recorded as HIP graphs and
    replayed: every replay equals the one-workgroup run."""
    from rajzel_amd import native as N
    sv = W.small_case(nx=70, ny=120, nshots=150, seed=14, rec_row=(2, 119))
    rng = np.random.default_rng(4)
    sv.tobs = (0.2 + rng.uniform(0.0, 0.3, (sv.nshots, sv.ntr))).astype(np.float32)
    print("[case] mixed launch as graph 120x70, 150 shots", flush=True)
    with N.Context([model_dict(sv, max_iter=200)], rows_per_lane=2) as ref:
        ref.set_split(False)
        ref.set_slowness(0, sv.S)
        ref.set_picks(0, sv.tobs)
        ref.gradient()
        want = ref.grad(0)
    with N.Context([model_dict(sv, max_iter=200)], rows_per_lane=2) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.gradient()                  # iteration history
        ctx.set_graph(True)
        for rep in range(3):
            ctx.gradient()
            ctx.sync()
            assert bits_equal(ctx.grad(0), want), f"replay {rep}: {ndiff(ctx.grad(0), want)} differ"
        assert ctx.split_state() == (1, 0)


def test_split_shot_counts_and_tuner_keep_results():
    """Shots per CU (rz_set_split_shots, rz_tune_split): with more shots than
    half the CUs, mixed launches running 0, some or all free CUs' worth of
    shots split give the same bits, and the tuner keeps one of its
    candidates with the results unchanged."""
    from rajzel_amd import native as N
    sv = W.small_case(nx=70, ny=60, nshots=200, seed=17, rec_row=2)
    sv.tobs = np.full((sv.nshots, sv.ntr), 0.4, np.float32)
    print("[case] split shot counts 60x70, 200 shots", flush=True)
    outs = {}
    with N.Context([model_dict(sv, max_iter=200)], rows_per_lane=1) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        F = ctx.compute_units() - sv.nshots
        ctx.gradient()                     # iteration counts that plan the mixed launches
        for n in (-1, 0, max(F // 2, 1)):
            ctx.set_split_shots(n)
            ctx.gradient()
            outs[n] = (ctx.grad(0), ctx.tmod(0), ctx.misfit(0)[0])
        n, ms = ctx.tune_split()
        assert n in (-1, 0, F // 4, F // 2, (3 * F) // 4) and ms > 0
        assert ctx.split_shots() == n
        ctx.gradient()
        outs["tuned"] = (ctx.grad(0), ctx.tmod(0), ctx.misfit(0)[0])
    ref = outs[-1]
    for k, v in outs.items():
        assert bits_equal(v[0], ref[0]) and bits_equal(v[1], ref[1]) and v[2] == ref[2], k


def test_failed_split_launch_retries_after_backoff():
    """After a failed split launch the context runs unsplit for 16 runs,
    then tries split launches again (the results are the same bits either
    way)."""
    from rajzel_amd import native as N
    sv = W.small_case(nx=64, ny=130, nshots=4, seed=6, rec_row=(2, 129))
    sv.tobs = np.full((sv.nshots, sv.ntr), 0.3, np.float32)
    print("[case] split retry 130x64", flush=True)
    ref = _run(N, sv, 2, False)
    with N.Context([model_dict(sv, max_iter=200)], rows_per_lane=2) as ctx:
        ctx.set_split(2)
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, sv.tobs)
        ctx.gradient()
        assert ctx.split_state() == (0, 2)
        states = []
        for _ in range(17):
            ctx.gradient()
            states.append(ctx.split_state()[0])
            assert bits_equal(ctx.grad(0), ref["grad"])
        # the failed run's own unsplit repetition counts as one of the 16
        assert 1 in states and states.index(1) >= 14 and states[-1] == 1, states
