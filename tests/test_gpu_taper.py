"""Per-shot gradient tapers (ass_grad.c:31-37, SWS_TAPER_CIRCULAR_PER_SHOT)
on the GPU against the reference's own grad_obj and ass_grad
(tests/golden/taper.npz, written by tests/golden/make_golden.py from
oracle/_ref):

  mode 1  taper_grad_shot.c:43-84, radial taper at the shot's source and
          every receiver (SRTRADIUS 35 and 80 m)
  mode 2  taper_grad_shot1.c:32-62, 1e-5 in the windows around all sources
          and receivers (FILTSIZE 1 and 0; windows that run off the grid
          sideways wrap onto the neighbouring row, as in the reference)

on a 34 x 60 survey with receivers on both edge columns, and on 3 shots of
the 200 x 600 C2 geometry (the radial taper at shots 57-59 of the 120-shot
source list, addressed through first_shot).  Gradient and L2 bit for bit,
in one batch, in batches of one shot (the sum's frame is copied after the
last), shot by shot, replayed as a HIP graph, and for single-shot ass_grad.
"""
import hashlib
import os

import numpy as np
import pytest

from rajzel_amd import workloads as W
from tests.helpers import bits_equal, model_dict, ndiff

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SMALL = ["r35", "r80", "w1", "w0"]


@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(GOLD, "taper.npz"))


def _set(fx, key):
    v = fx[key]
    return int(v[0]), float(v[1]), int(v[2])


@pytest.mark.parametrize("name", SMALL)
@pytest.mark.parametrize("how", ["batched", "one_shot_batches", "per_shot", "graph"])
def test_small_grad_obj(fx, name, how):
    from rajzel_amd import native as N
    sv = W.taper_case()
    mode, radius, filt = _set(fx, f"small_{name}_set")
    print(f"[case] taper {name} mode {mode} {how}", flush=True)
    budget = 0
    if how == "one_shot_batches":
        with N.Context([model_dict(sv)]) as probe:
            budget = probe.slot_bytes
    with N.Context([model_dict(sv)], hbm_budget=budget) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, fx["small_tobs"])
        ctx.set_shot_taper(0, mode, radius, filt, sv.src_x_m, sv.src_z_m)
        if how == "graph":
            ctx.set_graph(True)
            for _ in range(3):
                ctx.gradient()
                assert bits_equal(ctx.grad(0), fx[f"small_{name}_grad"])
        else:
            ctx.gradient(per_shot=how == "per_shot")
        if how == "one_shot_batches":
            assert ctx.batch_shots == 1
        g = ctx.grad(0)
        L2, _ = ctx.misfit(0)
    want = fx[f"small_{name}_grad"]
    assert bits_equal(g, want), f"{ndiff(g, want)} gradient values differ"
    assert L2 == float(fx[f"small_{name}_L2"])


@pytest.mark.parametrize("name", SMALL)
def test_small_ass_grad_shot(fx, name):
    from rajzel_amd import native as N
    sv = W.taper_case()
    mode, radius, filt = _set(fx, f"small_{name}_set")
    with N.Context([model_dict(sv)]) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.set_shot_taper(0, mode, radius, filt, sv.src_x_m, sv.src_z_m)
        for shot in (0, 3):
            gs, g = ctx.ass_grad_shot(0, fx[f"small_{name}_lam"], fx[f"small_{name}_g0"], src=shot)
            assert bits_equal(gs, fx[f"small_{name}_shot{shot}_gs"]), f"shot {shot}: grad_shot differs"
            assert bits_equal(g, fx[f"small_{name}_shot{shot}_g"]), f"shot {shot}: grad differs"


def test_taper_off_restores_plain_gradient(fx):
    """Setting mode 0 again gives the untapered gradient, and the tapered one
    differs from it."""
    from rajzel_amd import native as N
    sv = W.taper_case()
    with N.Context([model_dict(sv)]) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, fx["small_tobs"])
        ctx.gradient()
        plain = ctx.grad(0)
        ctx.set_shot_taper(0, 1, 80.0, 1, sv.src_x_m, sv.src_z_m)
        ctx.gradient()
        assert bits_equal(ctx.grad(0), fx["small_r80_grad"])
        ctx.set_shot_taper(0, 0)
        ctx.gradient()
        assert bits_equal(ctx.grad(0), plain)
    assert ndiff(plain, fx["small_r80_grad"]) > 0


@pytest.mark.parametrize("name", ["c2_r80", "c2_w0"])
def test_c2_geometry(fx, name):
    from rajzel_amd import native as N
    full = W.c2_gradient()
    mode, radius, filt = _set(fx, f"{name}_set")
    first = int(fx[f"{name}_set"][3])
    sv = W.c2_gradient()
    sv.src_x_m, sv.src_z_m = full.src_x_m[first:first + 3].copy(), full.src_z_m[first:first + 3].copy()
    print(f"[case] taper {name} on 200x600, shots {first}..{first + 2}", flush=True)
    with N.Context([model_dict(sv)]) as ctx:
        ctx.set_slowness(0, sv.S)
        ctx.set_picks(0, fx[f"{name}_tobs"])
        if mode == 1:   # the whole source list; this model's shots start at `first`
            ctx.set_shot_taper(0, mode, radius, filt, full.src_x_m, full.src_z_m, first_shot=first)
        else:           # the reference run held these 3 sources only
            ctx.set_shot_taper(0, mode, radius, filt, sv.src_x_m, sv.src_z_m)
        ctx.gradient()
        g = ctx.grad(0)
        L2, _ = ctx.misfit(0)
    assert np.array_equal(g[[1, 2, 100, 198]].view(np.uint32), fx[f"{name}_grad_rows"].view(np.uint32))
    assert hashlib.sha256(np.ascontiguousarray(g).tobytes()).hexdigest() == str(fx[f"{name}_d_grad"])
    assert L2 == float(fx[f"{name}_L2"])


def test_taper_arguments_rejected():
    from rajzel_amd import native as N
    sv = W.taper_case()
    with N.Context([model_dict(sv)]) as ctx:
        for kw in (dict(mode=3), dict(mode=1, radius_m=0.0), dict(mode=1, radius_m=float("nan")),
                   dict(mode=1, radius_m=80.0, first_shot=1)):   # shots 1..4 of 4 sources
            with pytest.raises(N.RzError) as e:
                ctx.set_shot_taper(0, src_x_m=sv.src_x_m, src_z_m=sv.src_z_m, **kw)
            assert e.value.code == N.RZ_ERR_INVALID
        ctx.set_shot_taper(0, 1, 80.0, 1, sv.src_x_m, sv.src_z_m)
        with pytest.raises(N.RzError):
            ctx.ass_grad_shot(0, np.zeros((sv.ny, sv.nx), np.float32), np.zeros((sv.ny, sv.nx), np.float32), src=4)
