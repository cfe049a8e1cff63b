"""Host side of the C ABI, without a GPU: the library loads, exports every
function include/rajzel_amd.h declares, and rejects invalid inputs before
any device call."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "rajzel_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rz_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from rajzel_amd import native as N
    lib = N.lib()
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(N.EXPORTS)


def test_grid_index_matches_reference_rounding():
    from rajzel_amd import native as N
    from rajzel_amd.workloads import iround_index
    xs = np.array([10.0, 3000.0, 37.5, 12.49, 12.5, 7.5, 2.5, 14962.5], np.float32)
    for dh in (10.0, 5.0, 2.5):
        assert [N.grid_index(x, dh) for x in xs] == list(iround_index(xs, dh))


def _model(**kw):
    m = dict(nx=20, ny=15, dh=10.0, src_ix=[3], src_iy=[0], rec_ix=[1, 2], rec_iy=[1, 1])
    m.update(kw)
    return m


@pytest.mark.parametrize("kw,msg", [
    (dict(nx=4), "NX, NY must be >= 5"),
    (dict(ny=3), "NX, NY must be >= 5"),
    (dict(dh=0.0), "bad DH"),
    (dict(dh=float("nan")), "bad DH"),
    (dict(src_ix=[20]), "outside the grid"),
    (dict(src_iy=[-1]), "outside the grid"),
    (dict(rec_ix=[1, 20]), "outside the grid"),
    (dict(rec_iy=[1, 15]), "outside the grid"),
    (dict(time=float("inf")), "must be finite"),
    (dict(ny=10000, nx=5000), "bands"),          # too tall stored either way
    (dict(ny=4097, nx=4097), "bands"),           # 9 bands of 512 rows: one more than a workgroup holds
    (dict(nx=9000000, ny=5), "exceeds 2 GiB"),   # 45M cells, but one band of 64 rows: 2.3 GB per field
])
def test_invalid_models_rejected_before_device(kw, msg):
    from rajzel_amd import native as N
    with pytest.raises(N.RzError) as e:
        N.Context([_model(**kw)])
    assert e.value.code == N.RZ_ERR_INVALID and msg in str(e.value)


def test_transposed_grids_need_eight_rows_per_lane():
    """NY = 10000 with NX = 20 is stored transposed, which only R = 8 runs."""
    from rajzel_amd import native as N
    with pytest.raises(N.RzError) as e:
        N.Context([_model(ny=10000)], rows_per_lane=4)
    assert e.value.code == N.RZ_ERR_INVALID and "transposed" in str(e.value)


def test_invalid_rows_per_lane():
    from rajzel_amd import native as N
    with pytest.raises(N.RzError) as e:
        N.Context([_model()], rows_per_lane=3)
    assert "rows_per_lane" in str(e.value)


def test_autotune_of_infeasible_model_fails_before_device():
    from rajzel_amd import native as N
    with pytest.raises(N.RzError) as e:
        N.autotune_rows([_model(ny=10000, nx=5000)])
    assert e.value.code == N.RZ_ERR_INVALID and "rows-per-lane" in str(e.value)
    with pytest.raises(N.RzError) as e:
        N.Context([_model(ny=10000, nx=5000)], rows_per_lane=N.ROWS_AUTOTUNE)
    assert "autotune" in str(e.value)


def test_gauss_filt_argument_checks_before_device():
    import ctypes as C
    from rajzel_amd import native as N
    L = N.lib()
    x, y = C.c_int(9), C.c_int(5)
    a = np.zeros((4, 4), np.float32)
    p = a.ctypes.data_as(C.c_void_p)
    assert L.rz_gauss_filt(0, None, None, 4, 4, C.byref(x), C.byref(y), 0) == N.RZ_ERR_INVALID
    assert "null" in L.rz_last_error(None).decode()
    assert L.rz_gauss_filt(0, None, p, 0, 4, C.byref(x), C.byref(y), 0) == N.RZ_ERR_INVALID
    # size 0 returns at once (gauss_filt.c:27), without touching a device;
    # an even size before it is made odd first
    x, y = C.c_int(8), C.c_int(0)
    assert L.rz_gauss_filt(0, None, p, 4, 4, C.byref(x), C.byref(y), 0) == N.RZ_OK
    assert (x.value, y.value) == (9, 0)


def test_null_context_calls_fail_cleanly():
    from rajzel_amd import native as N
    L = N.lib()
    assert L.rz_forward(None) == N.RZ_ERR_INVALID
    assert L.rz_sync(None) == N.RZ_ERR_INVALID
    assert L.rz_set_profiling(None, 1) == N.RZ_ERR_INVALID
    assert L.rz_rows_per_lane(None) == 0
    assert L.rz_get_stream(None) is None
    assert L.rz_device_bytes(None) == 0


def test_header_compiles_as_c():
    """The header is plain C (callers are C drivers)."""
    import subprocess
    import tempfile
    src = '#include "rajzel_amd.h"\nint main(void){ rz_ctx *c = 0; (void)c; return 0; }\n'
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "t.c")
        open(p, "w").write(src)
        r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", p,
                            "-o", os.path.join(d, "t.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
