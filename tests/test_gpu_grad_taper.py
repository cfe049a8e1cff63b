"""Gradient tapers of precond.c:43-53 (taper_grad.c) on the GPU through
rz_taper_grad, against the reference's own taper_grad (tests/golden/
grad_taper.npz, written by tests/golden/make_golden.py from oracle/_ref):

  sws 1  vertical taper, five GRADT sets (window ends inside and past the
         grid, an empty taper length, a negative GRADT1)
  sws 2  horizontal taper, five GRADT/EXP_TAPER_GRAD_HOR sets, and the
         stock setting (22, 25, EXP 0) on the 1001 x 3001 grid
  sws 3  radial taper at every source and receiver (SRTRADIUS 35, 80 m) on
         a 34 x 60 survey and on the 200 x 600 C2 geometry
  sws 4, 6  the coefficients of taper.bin

Fields and the taper_coeff files the reference writes, bit for bit, from
host arrays and from device tensors.
"""
import hashlib
import os

import numpy as np
import pytest

from rajzel_amd import workloads as W
from tests.helpers import bits_equal, ndiff

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(GOLD, "grad_taper.npz"))


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _check(got, want, what):
    assert bits_equal(got, want), f"{what}: {ndiff(got, want)} values differ"


@pytest.mark.parametrize("i", range(5))
def test_vertical(fx, i):
    from rajzel_amd import native as N
    out, coeff = N.taper_grad(fx["small_in"], 1, gradt=fx[f"v{i}_gradt"], want_coeff=True)
    _check(out, fx[f"v{i}_out"], "field")
    _check(coeff, fx[f"v{i}_coeff"], "taper_coeff")


@pytest.mark.parametrize("i", range(5))
def test_horizontal(fx, i):
    from rajzel_amd import native as N
    out, coeff = N.taper_grad(fx["small_in"], 2, gradt=fx[f"h{i}_gradt"], dh=10.0, exp_hor=float(fx[f"h{i}_exp"]),
                              want_coeff=True)
    _check(out, fx[f"h{i}_out"], "field")
    _check(coeff, fx[f"h{i}_coeff"], "taper_coeff")


@pytest.mark.parametrize("i", range(2))
def test_sources_and_receivers(fx, i):
    from rajzel_amd import native as N
    sv = W.taper_case()
    for _ in range(2):   # the second call reuses the factor tables
        out = N.taper_grad(fx["small_in"], 3, dh=sv.dh, srtradius=float(fx[f"r{i}_radius"]), src_x_m=sv.src_x_m,
                           src_z_m=sv.src_z_m, rec_ix=sv.rec_ix, rec_iy=sv.rec_iy)
        _check(out, fx[f"r{i}_out"], "field")


@pytest.mark.parametrize("sws", [4, 6])
def test_file(fx, sws):
    from rajzel_amd import native as N
    out, coeff = N.taper_grad(fx["small_in"], sws, taper_file=fx["taper_file"], want_coeff=True)
    _check(out, fx[f"f{sws}_out"], "field")
    _check(coeff, fx[f"f{sws}_coeff"], "taper_coeff")


def test_stock_horizontal_1001x3001(fx):
    from rajzel_amd import native as N
    big = np.random.default_rng(506).normal(0, 1e-3, (1001, 3001)).astype(np.float32)
    out, coeff = N.taper_grad(big, 2, gradt=(22, 25, 490, 500), dh=5.0, exp_hor=0.0, want_coeff=True)
    assert np.array_equal(out[[0, 21, 22, 24, 25, 1000]].view(np.uint32), fx["c4hor_rows"].view(np.uint32))
    assert _digest(out) == str(fx["c4hor_d_out"])
    assert _digest(coeff) == str(fx["c4hor_d_coeff"])


def test_c2_sources_and_receivers_on_device(fx):
    """The C2 geometry's 720 discs on a device tensor, in place; then a
    different geometry in between, so the tables are rebuilt."""
    torch = pytest.importorskip("torch")
    from rajzel_amd import native as N
    c2 = W.c2_gradient()
    mid = np.random.default_rng(507).normal(0, 1e-3, (c2.ny, c2.nx)).astype(np.float32)
    sv = W.taper_case()
    for k in range(2):
        t = torch.tensor(mid, device="cuda:0")
        N.taper_grad(t, 3, dh=c2.dh, srtradius=80.0, src_x_m=c2.src_x_m, src_z_m=c2.src_z_m, rec_ix=c2.rec_ix,
                     rec_iy=c2.rec_iy, stream=torch.cuda.current_stream().cuda_stream)
        out = t.cpu().numpy()
        assert np.array_equal(out[[0, 1, 2, 5, 100, 199]].view(np.uint32), fx["c2disc_rows"].view(np.uint32))
        assert _digest(out) == str(fx["c2disc_d_out"])
        _check(N.taper_grad(fx["small_in"], 3, dh=sv.dh, srtradius=35.0, src_x_m=sv.src_x_m, src_z_m=sv.src_z_m,
                            rec_ix=sv.rec_ix, rec_iy=sv.rec_iy), fx["r0_out"], "field")


def test_arguments_rejected(fx):
    from rajzel_amd import native as N
    A = fx["small_in"]
    for kw in (dict(sws=0), dict(sws=7), dict(sws=3, dh=10.0, srtradius=0.0), dict(sws=3, dh=10.0, srtradius=80.0,
                                                                                  src_x_m=[1.0], src_z_m=[]),
               dict(sws=4)):
        with pytest.raises(N.RzError) as e:
            N.taper_grad(A, **kw)
        assert e.value.code == N.RZ_ERR_INVALID
