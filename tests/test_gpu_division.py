"""Division modes of the adjoint.  The T-only quotients x/DH (slopes and
the terms of lhs) run as a float FMA with a split reciprocal where the
model's DH and slowness allow it (checked on the host); these tests run the
same cases with that mode turned off (RAJZEL_AMD_FLOAT_T=0) and require the
same bits, and the CPU restatement's bits where it is cheap."""
import os

import numpy as np
import pytest

from rajzel_amd import workloads as W
from tests.helpers import bits_equal, model_dict, ndiff, oracle_picks

pytestmark = pytest.mark.gpu


def _grad(N, sv, R, float_t, split=True, S=None):
    old = os.environ.get("RAJZEL_AMD_FLOAT_T")
    os.environ["RAJZEL_AMD_FLOAT_T"] = "1" if float_t else "0"
    try:
        with N.Context([model_dict(sv, max_iter=200)], rows_per_lane=R) as ctx:
            ctx.set_split(split)
            ctx.set_slowness(0, sv.S if S is None else S)
            ctx.set_picks(0, sv.tobs)
            ctx.gradient()
            itf, ita = ctx.iterations(0)
            return dict(grad=ctx.grad(0), L2=ctx.misfit(0)[0], tres=ctx.tres(0), itf=itf, ita=ita)
    finally:
        if old is None:
            del os.environ["RAJZEL_AMD_FLOAT_T"]
        else:
            os.environ["RAJZEL_AMD_FLOAT_T"] = old


def _same(a, b):
    for k in ("grad", "tres"):
        assert bits_equal(a[k], b[k]), f"{k}: {ndiff(a[k], b[k])} differ"
    assert a["L2"] == b["L2"]
    assert np.array_equal(a["itf"], b["itf"]) and np.array_equal(a["ita"], b["ita"])


# DH 10 and 5 have short significands (the float quotient applies); 4.7 and
# 3.3 have long ones (the host check decides; either way the bits must hold).
@pytest.mark.parametrize("dh,R,shape,ns,split", [
    (10.0, 1, (37, 53), 3, True), (5.0, 4, (300, 90), 4, True), (5.0, 2, (150, 70), 3, False),
    (4.7, 1, (37, 53), 3, False), (3.3, 8, (600, 40), 2, True), (2.5, 2, (190, 61), 5, True)])
def test_float_t_quotients_same_bits(ora, dh, R, shape, ns, split):
    from rajzel_amd import native as N
    ny, nx = shape
    sv = W.small_case(nx=nx, ny=ny, dh=dh, nshots=ns, seed=300 + ny, rec_row=(2, ny - 1))
    oracle_picks(ora, sv)
    print(f"[case] float T quotients {ny}x{nx} DH={dh} R={R} split={split}", flush=True)
    a = _grad(N, sv, R, True, split)
    b = _grad(N, sv, R, False, split)
    _same(a, b)
    if ny * nx <= 40000:
        o = ora.grad_obj(sv.S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.tobs, sv.dh)
        assert bits_equal(a["grad"], o["grad"]) and a["L2"] == o["L2"]


def test_float_t_with_tiny_slowness(ora):
    """Slowness far below physical values around a source (traveltimes of
    1e-34 s): the host turns the float quotients off for this model; the
    result is the restatement's either way."""
    from rajzel_amd import native as N
    sv = W.small_case(nx=53, ny=37, nshots=2, seed=41, rec_row=(2, 36))
    oracle_picks(ora, sv)
    S = sv.S.copy()
    for s in range(sv.nshots):
        y, x = int(sv.src_iy[s]), int(sv.src_ix[s])
        S[max(0, y - 2):y + 3, max(0, x - 2):x + 3] = np.float32(1e-36)
    print("[case] float T quotients, tiny slowness 37x53", flush=True)
    a = _grad(N, sv, 1, True, S=S)
    b = _grad(N, sv, 1, False, S=S)
    _same(a, b)
    o = ora.grad_obj(S, sv.src_ix, sv.src_iy, sv.rec_ix, sv.rec_iy, sv.tobs, sv.dh)
    assert bits_equal(a["grad"], o["grad"]) and a["L2"] == o["L2"]


def test_float_t_per_shot_entry_with_callers_T(ora):
    """rz_eikonal_adj_shot takes the caller's T: the host checks that T
    (not the slowness) before the float quotients run; the per-shot lambda
    equals the batched adjoint's and the restatement's."""
    from rajzel_amd import native as N
    sv = W.small_case(nx=53, ny=37, nshots=2, seed=43, rec_row=(2, 36))
    oracle_picks(ora, sv)
    print("[case] float T quotients, per-shot adjoint 37x53", flush=True)
    outs = []
    for float_t in (True, False):
        old = os.environ.get("RAJZEL_AMD_FLOAT_T")
        os.environ["RAJZEL_AMD_FLOAT_T"] = "1" if float_t else "0"
        try:
            with N.Context([model_dict(sv, max_iter=200)], rows_per_lane=1) as ctx:
                ctx.set_slowness(0, sv.S)
                T, tmod, _ = ctx.eikonal_shot(0, sv.src_ix[0], sv.src_iy[0])
                tres = ((tmod - sv.tobs[0]) * np.float32(10.0)).astype(np.float32)
                lam, it = ctx.eikonal_adj_shot(0, T, tres)
                outs.append((T, lam, it))
        finally:
            if old is None:
                del os.environ["RAJZEL_AMD_FLOAT_T"]
            else:
                os.environ["RAJZEL_AMD_FLOAT_T"] = old
    assert bits_equal(outs[0][1], outs[1][1]) and outs[0][2] == outs[1][2]
    lam_o, it_o = ora.eikonal_adj(outs[0][0], sv.rec_ix, sv.rec_iy, tres, sv.dh)
    assert bits_equal(outs[0][1], lam_o) and outs[0][2] == it_o
