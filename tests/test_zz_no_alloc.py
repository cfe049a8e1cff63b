"""rz_gauss_filt_noalloc (the Gaussian filter on a static device scratch
array, no device allocation) against rz_gauss_filt, bit for bit, through the
C ABI: two calls in a row on one device, the second on a smaller grid and a
smaller window, so that a stale weight or input of the first call would
show."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import bits_equal, ndiff

pytestmark = pytest.mark.gpu


def _call_noalloc(N, t, fsx, fsy, stream):
    ny, nx = t.shape
    x, y = C.c_int(fsx), C.c_int(fsy)
    rc = N.lib().rz_gauss_filt_noalloc(0, stream, C.c_void_p(t.data_ptr()), nx, ny, C.byref(x), C.byref(y))
    if rc != N.RZ_OK:
        raise N.RzError(rc, N.lib().rz_last_error(None).decode())
    return x.value, y.value


def test_noalloc_equals_allocating_filter_twice():
    import torch
    from rajzel_amd import native as N
    stream = torch.cuda.current_stream().cuda_stream
    cases = [((1001, 3001), (81, 41), 31), ((150, 333), (30, 14), 32)]   # second: smaller grid, even sizes
    for (ny, nx), (fx, fy), seed in cases:
        print(f"[case] gauss_filt_noalloc {ny}x{nx} window {fx}x{fy}", flush=True)
        A = np.random.default_rng(seed).normal(0.0, 1e-3, (ny, nx)).astype(np.float32)
        ref, rx, ry = N.gauss_filt(A, fx, fy)                       # allocating entry point, host field
        t = torch.from_numpy(A).to("cuda")
        gx, gy = _call_noalloc(N, t, fx, fy, stream)
        torch.cuda.synchronize()
        assert (gx, gy) == (rx, ry)
        out = t.cpu().numpy()
        assert bits_equal(out, ref), f"{ndiff(out, ref)} cells differ"


def test_noalloc_rejects_inputs_beyond_the_scratch():
    import torch
    from rajzel_amd import native as N
    limit = (8 << 20) - (16 << 10)                                   # RZ_GAUSS_NOALLOC_* of the header
    t = torch.zeros(1, dtype=torch.float32, device="cuda")
    x, y = C.c_int(5), C.c_int(5)
    # more cells than the scratch holds: rejected before anything is read
    rc = N.lib().rz_gauss_filt_noalloc(0, None, C.c_void_p(t.data_ptr()), limit + 1, 1, C.byref(x), C.byref(y))
    assert rc == N.RZ_ERR_INVALID and b"scratch" in N.lib().rz_last_error(None)
    assert (x.value, y.value) == (5, 5)
    # a window with more weights than the scratch keeps
    x, y = C.c_int(201), C.c_int(201)
    rc = N.lib().rz_gauss_filt_noalloc(0, None, C.c_void_p(t.data_ptr()), 1, 1, C.byref(x), C.byref(y))
    assert rc == N.RZ_ERR_INVALID and b"weights" in N.lib().rz_last_error(None)
