"""precond.c's other gradient smoothings on the device, bit for bit against
the reference's own smooth_grad.c + median2d.c (SPATFILTER = 3) and
smooth2.c + tripd (SPATFILTER = 2), compiled from its sources
(oracle/_ref/librzref.so)."""
import os
import tempfile

import numpy as np
import pytest

from oracle.bindings import Reference
from rajzel_amd import native as N
from tests.helpers import bits_equal, ndiff

pytestmark = pytest.mark.gpu


def _field(ny, nx, seed, kind):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((ny, nx)).astype(np.float32)
    if kind == "ties":                       # few distinct values, signed zeros among them
        A = np.round(A * 2.0).astype(np.float32) / np.float32(2.0)
        A[rng.random((ny, nx)) < 0.2] = -0.0
    elif kind == "gradient":                 # a gradient-like field: spikes on a smooth background
        X, Z = np.meshgrid(np.arange(nx), np.arange(ny))
        A = (1e-6 * np.sin(X / 17.0) * np.cos(Z / 11.0)).astype(np.float32)
        A[rng.integers(0, ny, 40), rng.integers(0, nx, 40)] = np.float32(3e-4)
    return A


@pytest.mark.parametrize("ny,nx,fsize,kind", [(37, 53, 3, "random"), (37, 53, 4, "ties"), (40, 29, 7, "ties"),
                                              (64, 120, -5, "random"), (12, 9, 15, "random"),
                                              (1001, 3001, 5, "gradient")])
def test_median_filter_equals_reference(ny, nx, fsize, kind):
    print(f"[case] median {ny}x{nx} size {fsize} {kind}", flush=True)
    A = _field(ny, nx, 3 + ny, kind)
    ref = Reference()
    ref.setup(nx, ny, 5.0)
    want, fs_ref = ref.smooth_grad(A, fsize)
    got, fs = N.median_filt(A, fsize)
    assert fs == fs_ref
    assert bits_equal(got, want), f"{ndiff(got, want)} cells differ"


def test_median_filter_size_zero_leaves_field():
    A = _field(20, 30, 1, "random")
    got, fs = N.median_filt(A, 0)
    assert fs == 0 and bits_equal(got, A)


@pytest.mark.parametrize("ny,nx,wd,wd1,kind", [(37, 53, 3.0, 2.0, "random"), (40, 29, 0.5, 7.0, "ties"),
                                                (5, 4, 1.0, 1.0, "random"), (1001, 3001, 10.0, 10.0, "gradient")])
def test_smooth2_equals_reference(ny, nx, wd, wd1, kind):
    print(f"[case] smooth2 {ny}x{nx} WD_DAMP={wd} WD_DAMP1={wd1} {kind}", flush=True)
    A = _field(ny, nx, 7 + nx, kind)
    ref = Reference()
    ref.setup(nx, ny, 5.0)
    with tempfile.TemporaryDirectory(prefix="rzs", dir="/tmp") as d:
        want = ref.smooth2(A, wd, wd1, os.path.join(d, "j"))
    got = N.smooth2(A, wd, wd1)
    assert bits_equal(got, want), f"{ndiff(got, want)} cells differ"
