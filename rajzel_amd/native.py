"""ctypes binding of librajzel_amd.so (include/rajzel_amd.h).

This is the only way Python reaches the GPU path: there is no fallback.  If
the library is missing, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librajzel_amd.so")

RZ_OK, RZ_ERR_INVALID, RZ_ERR_HIP, RZ_ERR_NOMEM, RZ_ERR_NOT_CONVERGED, RZ_ERR_STATE = range(6)

# Every symbol declared in include/rajzel_amd.h.
EXPORTS = (
    "rz_grid_index", "rz_create", "rz_destroy", "rz_last_error", "rz_set_stream", "rz_set_slowness",
    "rz_set_picks", "rz_set_tres_carry", "rz_forward", "rz_gradient", "rz_gradient_per_shot",
    "rz_get_tmod", "rz_get_tres", "rz_get_grad", "rz_get_misfit", "rz_get_iterations",
    "rz_get_tres_carry", "rz_get_field", "rz_eikonal_shot", "rz_eikonal_adj_shot", "rz_ass_grad_shot",
    "rz_sync", "rz_rows_per_lane", "rz_device_bytes", "rz_batch_shots", "rz_set_profiling",
    "rz_get_profile", "rz_set_residuals", "rz_adjoint", "rz_get_stream", "rz_set_graph",
    "rz_autotune_rows", "rz_gauss_filt", "rz_get_tnb", "rz_set_split", "rz_get_split",
    "rz_read_picks", "rz_picks_parses", "rz_picks_clear", "rz_gauss_filt_noalloc",
    "rz_slot_bytes", "rz_get_stop_passes", "rz_set_shot_taper", "rz_ass_grad_shot_at", "rz_taper_grad",
    "rz_dotp_ordered", "rz_lbfgs_create", "rz_lbfgs_destroy", "rz_lbfgs_error", "rz_lbfgs_zero",
    "rz_lbfgs_has_previous", "rz_lbfgs_set_previous", "rz_lbfgs_step", "rz_compute_units",
    "rz_set_split_shots", "rz_get_split_shots", "rz_tune_split", "rz_median_filt", "rz_smooth2",
)


class RzError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rajzel_amd error {code}: {msg}")
        self.code = code


class NotConverged(RzError):
    pass


class ModelDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("dh", C.c_float), ("time", C.c_float),
        ("ttnorm", C.c_float), ("eps_adj", C.c_float), ("max_iter", C.c_int), ("nshots", C.c_int),
        ("src_ix", C.POINTER(C.c_int)), ("src_iy", C.POINTER(C.c_int)), ("ntr", C.c_int),
        ("rec_ix", C.POINTER(C.c_int)), ("rec_iy", C.POINTER(C.c_int)),
    ]


_lib = None
_loaded_after_torch = False
_lock = threading.Lock()


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib, _loaded_after_torch
    with _lock:
        if _lib is None:
            # torch bundles its own libamdhip64.so.7; whichever is loaded first
            # serves the whole process (same soname), so record the order.
            import sys
            _loaded_after_torch = "torch" in sys.modules
            if not os.path.exists(LIB_PATH):
                raise RzError(-1, f"{LIB_PATH} not built: run __graft_entry__.build() or "
                                  "python rajzel_amd/build.py")
            L = C.CDLL(LIB_PATH)
            vp, i, f, d = C.c_void_p, C.c_int, C.c_float, C.c_double
            sig = {
                "rz_grid_index": (i, [f, f]),
                "rz_create": (i, [C.POINTER(vp), i, i, C.POINTER(ModelDesc), i, C.c_size_t]),
                "rz_destroy": (None, [vp]),
                "rz_last_error": (C.c_char_p, [vp]),
                "rz_set_stream": (i, [vp, vp]),
                "rz_set_slowness": (i, [vp, i, vp, i]),
                "rz_set_picks": (i, [vp, i, vp, i]),
                "rz_set_tres_carry": (i, [vp, i, vp]),
                "rz_forward": (i, [vp]),
                "rz_gradient": (i, [vp]),
                "rz_gradient_per_shot": (i, [vp]),
                "rz_get_tmod": (i, [vp, i, vp, i]),
                "rz_get_tres": (i, [vp, i, vp, i]),
                "rz_get_grad": (i, [vp, i, vp, i]),
                "rz_get_misfit": (i, [vp, i, C.POINTER(d), vp]),
                "rz_get_iterations": (i, [vp, i, vp, vp]),
                "rz_get_tres_carry": (i, [vp, i, vp]),
                "rz_get_field": (i, [vp, i, i, i, vp, i]),
                "rz_eikonal_shot": (i, [vp, i, i, i, vp, vp, C.POINTER(i)]),
                "rz_eikonal_adj_shot": (i, [vp, i, vp, vp, vp, C.POINTER(i)]),
                "rz_ass_grad_shot": (i, [vp, i, vp, vp, vp]),
                "rz_ass_grad_shot_at": (i, [vp, i, i, vp, vp, vp]),
                "rz_set_shot_taper": (i, [vp, i, i, f, i, i, vp, vp, i]),
                "rz_sync": (i, [vp]),
                "rz_rows_per_lane": (i, [vp]),
                "rz_compute_units": (i, [vp]),
                "rz_set_split_shots": (i, [vp, i]),
                "rz_get_split_shots": (i, [vp]),
                "rz_tune_split": (i, [vp, C.POINTER(i), C.POINTER(d)]),
                "rz_median_filt": (i, [i, vp, vp, i, i, C.POINTER(i), i]),
                "rz_smooth2": (i, [i, vp, vp, vp, i, i, f, f]),
                "rz_device_bytes": (C.c_size_t, [vp]),
                "rz_batch_shots": (i, [vp]),
                "rz_slot_bytes": (C.c_size_t, [vp]),
                "rz_get_stop_passes": (i, [vp, i, vp, vp, vp, vp, vp, vp]),
                "rz_set_profiling": (i, [vp, i]),
                "rz_get_profile": (i, [vp, vp, vp]),
                "rz_set_residuals": (i, [vp, i, vp, i]),
                "rz_adjoint": (i, [vp]),
                "rz_get_stream": (vp, [vp]),
                "rz_set_graph": (i, [vp, i]),
                "rz_autotune_rows": (i, [i, i, C.POINTER(ModelDesc), C.c_size_t, C.POINTER(i), C.POINTER(d)]),
                "rz_gauss_filt": (i, [i, vp, vp, i, i, C.POINTER(i), C.POINTER(i), i]),
                "rz_gauss_filt_noalloc": (i, [i, vp, vp, i, i, C.POINTER(i), C.POINTER(i)]),
                "rz_taper_grad": (i, [i, vp, vp, i, i, vp, vp, i]),
                "rz_get_tnb": (i, [vp, i, vp, i]),
                "rz_set_split": (i, [vp, i]),
                "rz_get_split": (i, [vp, C.POINTER(i), C.POINTER(i)]),
                "rz_read_picks": (i, [C.c_char_p, vp, i]),
                "rz_picks_parses": (C.c_long, []),
                "rz_picks_clear": (None, []),
                "rz_dotp_ordered": (i, [i, vp, vp, C.c_longlong, i, C.POINTER(f)]),
                "rz_lbfgs_create": (i, [C.POINTER(vp), i, i, i, i]),
                "rz_lbfgs_destroy": (None, [vp]),
                "rz_lbfgs_error": (C.c_char_p, [vp]),
                "rz_lbfgs_zero": (i, [vp]),
                "rz_lbfgs_has_previous": (i, [vp]),
                "rz_lbfgs_set_previous": (i, [vp, vp, vp]),
                "rz_lbfgs_step": (i, [vp, i, i, vp, vp, vp, vp, vp, vp, vp, C.POINTER(f), vp, vp]),
            }
            for name, (res, args) in sig.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
        return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def grid_index(pos_m, dh):
    return lib().rz_grid_index(float(pos_m), float(dh))


ROWS_AUTOTUNE = -1


def _descs(models):
    """rz_model_desc array for `models`, the arrays it points into (keep
    them alive for the call) and (ny, nx, nshots, ntr) per model."""
    keep, shapes = [], []
    descs = (ModelDesc * len(models))()
    for k, m in enumerate(models):
        g = (lambda n, dflt=None: getattr(m, n, m.get(n, dflt) if isinstance(m, dict) else dflt))
        sx = np.ascontiguousarray(g("src_ix"), np.int32)
        sy = np.ascontiguousarray(g("src_iy"), np.int32)
        rx = np.ascontiguousarray(g("rec_ix"), np.int32)
        ry = np.ascontiguousarray(g("rec_iy"), np.int32)
        keep += [sx, sy, rx, ry]
        d = descs[k]
        d.nx, d.ny, d.dh = int(g("nx")), int(g("ny")), float(g("dh"))
        d.time = float(g("time", 6.0))
        d.ttnorm = float(g("ttnorm", 1e-10))
        d.eps_adj = float(g("eps_adj", 1e-9))
        d.max_iter = int(g("max_iter", 0) or 0)
        d.nshots, d.ntr = len(sx), len(rx)
        ip = C.POINTER(C.c_int)
        d.src_ix, d.src_iy = sx.ctypes.data_as(ip), sy.ctypes.data_as(ip)
        d.rec_ix, d.rec_iy = rx.ctypes.data_as(ip), ry.ctypes.data_as(ip)
        shapes.append((d.ny, d.nx, d.nshots, d.ntr))
    return descs, keep, shapes


def autotune_rows(models, device=0, hbm_budget=0):
    """(rows per lane, ms of one outer forward + adjoint iteration) of the
    fastest feasible setting for these models (rz_autotune_rows)."""
    descs, keep, _ = _descs(models)
    R, ms = C.c_int(0), C.c_double(0.0)
    rc = lib().rz_autotune_rows(device, len(models), descs, hbm_budget, C.byref(R), C.byref(ms))
    del keep
    if rc != RZ_OK:
        raise RzError(rc, lib().rz_last_error(None).decode())
    return R.value, ms.value


def gauss_filt(field, fsx, fsy, device=0, stream=None):
    """gauss_filt.c on a numpy (NY, NX) float32 array (in place on a copy)
    or a torch CUDA tensor (in place).  Returns (field, fsx, fsy) with the
    sizes made odd as the reference does."""
    x, y = C.c_int(int(fsx)), C.c_int(int(fsy))
    if hasattr(field, "data_ptr"):
        ny, nx = field.shape
        rc = lib().rz_gauss_filt(device, stream, C.c_void_p(field.data_ptr()), nx, ny, C.byref(x), C.byref(y), 1)
        out = field
    else:
        out = np.ascontiguousarray(field, np.float32).copy()
        ny, nx = out.shape
        rc = lib().rz_gauss_filt(device, stream, out.ctypes.data_as(C.c_void_p), nx, ny, C.byref(x), C.byref(y), 0)
    if rc != RZ_OK:
        raise RzError(rc, lib().rz_last_error(None).decode())
    return out, x.value, y.value


class GradTaper(C.Structure):
    _fields_ = [
        ("sws", C.c_int), ("gradt1", C.c_int), ("gradt2", C.c_int), ("gradt3", C.c_int), ("gradt4", C.c_int),
        ("dh", C.c_float), ("exp_taper_grad_hor", C.c_float), ("srtradius", C.c_float),
        ("nsrc", C.c_int), ("src_x_m", C.POINTER(C.c_float)), ("src_z_m", C.POINTER(C.c_float)),
        ("ntr", C.c_int), ("rec_ix", C.POINTER(C.c_int)), ("rec_iy", C.POINTER(C.c_int)),
        ("file", C.POINTER(C.c_float)),
    ]


def taper_grad(field, sws, gradt=(0, 0, 0, 0), dh=0.0, exp_hor=0.0, srtradius=0.0, src_x_m=(), src_z_m=(),
               rec_ix=(), rec_iy=(), taper_file=None, device=0, stream=None, want_coeff=False):
    """taper_grad.c (precond.c:43-53) through rz_taper_grad on a numpy
    (NY, NX) float32 array (on a copy) or a torch CUDA tensor (in place):
    sws 1 vertical, 2 horizontal, 3 at all sources and receivers, 4..6
    from taper_file (NX*NY floats, x outer).  Returns the field, and with
    want_coeff also taper_coeff as the reference writes it (None for sws 3)."""
    keep = []

    def arr(a, dt, ct):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(ct)) if a.size else None
    sx, sz = arr(src_x_m, np.float32, C.c_float), arr(src_z_m, np.float32, C.c_float)
    rx, ry = arr(rec_ix, np.int32, C.c_int), arr(rec_iy, np.int32, C.c_int)
    t = GradTaper(int(sws), *[int(g) for g in gradt], float(dh), float(exp_hor), float(srtradius),
                  len(np.atleast_1d(src_x_m)) if np.size(src_x_m) else 0, sx, sz,
                  len(np.atleast_1d(rec_ix)) if np.size(rec_ix) else 0, rx, ry,
                  arr(taper_file, np.float32, C.c_float) if taper_file is not None else None)
    if hasattr(field, "data_ptr"):
        ny, nx = field.shape
        out, ptr, on_dev = field, C.c_void_p(field.data_ptr()), 1
    else:
        out = np.ascontiguousarray(field, np.float32).copy()
        ny, nx = out.shape
        ptr, on_dev = out.ctypes.data_as(C.c_void_p), 0
    coeff = np.zeros((ny, nx), np.float32) if want_coeff and sws != 3 else None
    rc = lib().rz_taper_grad(device, stream, ptr, nx, ny, C.byref(t), _ptr(coeff), on_dev)
    if rc != RZ_OK:
        raise RzError(rc, lib().rz_last_error(None).decode())
    return (out, coeff) if want_coeff else out


def median_filt(field, fsize, device=0, stream=None):
    """smooth_grad.c's median filter on a numpy (NY, NX) float32 array (a
    copy).  Returns (field, fsize made odd)."""
    out = np.ascontiguousarray(field, np.float32).copy()
    ny, nx = out.shape
    fs = C.c_int(int(fsize))
    rc = lib().rz_median_filt(device, stream, _ptr(out), nx, ny, C.byref(fs), 0)
    if rc != RZ_OK:
        raise RzError(rc, lib().rz_last_error(None).decode())
    return out, fs.value


def smooth2(field, wd_damp, wd_damp1, device=0, stream=None):
    """smooth2.c's tridiagonal smoothing of a numpy (NY, NX) float32 field
    (unstrided); returns the smoothed field (NY, NX) as smooth2 leaves it in
    grad."""
    A = np.ascontiguousarray(field, np.float32)
    ny, nx = A.shape
    xin = np.ascontiguousarray(A.T)                       # x slow, z fast
    out = np.empty_like(xin)
    rc = lib().rz_smooth2(device, stream, _ptr(xin), _ptr(out), nx, ny, float(wd_damp), float(wd_damp1))
    if rc != RZ_OK:
        raise RzError(rc, lib().rz_last_error(None).decode())
    return np.ascontiguousarray(out.T)


def dotp_ordered(a, b, device=0):
    """dotp.c's serial float dot product sum(a[i]*b[i]), i ascending, of two
    float32 vectors, on the device (rz_dotp_ordered)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    if a.shape != b.shape or a.ndim != 1:
        raise ValueError("a and b must be 1-D float32 vectors of one length")
    out = C.c_float()
    rc = lib().rz_dotp_ordered(device, _ptr(a), _ptr(b), a.size, 0, C.byref(out))
    if rc != RZ_OK:
        raise RzError(rc, lib().rz_last_error(None).decode())
    return np.float32(out.value)


class LBFGS:
    """LBFGS.c's two-loop recursion on the device (rz_lbfgs_*), for an
    NY x NX grid and nlbfgs history slots.  step() takes the row-major fields
    grad, gradm, vp and returns the direction hgrad and the reference's
    scalars and file contents."""

    def __init__(self, nx, ny, nlbfgs, device=0):
        self.nx, self.ny, self.K = nx, ny, nlbfgs
        h = C.c_void_p()
        rc = lib().rz_lbfgs_create(C.byref(h), device, nx, ny, nlbfgs)
        if rc != RZ_OK:
            raise RzError(rc, lib().rz_last_error(None).decode())
        self.h = h

    def _chk(self, rc):
        if rc != RZ_OK:
            raise RzError(rc, lib().rz_lbfgs_error(self.h).decode())

    def close(self):
        if self.h:
            lib().rz_lbfgs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def zero(self):
        self._chk(lib().rz_lbfgs_zero(self.h))

    def has_previous(self):
        return bool(lib().rz_lbfgs_has_previous(self.h))

    def set_previous(self, grad_h, vp_h):
        g = np.ascontiguousarray(grad_h, np.float32)
        v = np.ascontiguousarray(vp_h, np.float32)
        self._chk(lib().rz_lbfgs_set_previous(self.h, _ptr(g), _ptr(v)))

    def step(self, it, pointer, grad, gradm, vp):
        """One LBFGS() call; returns a dict with hgrad (NY, NX), grad_h,
        vp_h, hgrad_h (h order), gamma, rho, alpha (iter > 1)."""
        shp = (self.ny, self.nx)
        f = [np.ascontiguousarray(x, np.float32).reshape(shp) for x in (grad, gradm, vp)]
        out = {"hgrad": np.empty(shp, np.float32)}
        for k in ("grad_h", "vp_h", "hgrad_h"):
            out[k] = np.empty(self.nx * self.ny, np.float32)
        gamma = C.c_float(np.nan)
        rho = np.full(self.K, np.nan, np.float32)
        alpha = np.full(self.K, np.nan, np.float32)
        self._chk(lib().rz_lbfgs_step(self.h, it, pointer, _ptr(f[0]), _ptr(f[1]), _ptr(f[2]), _ptr(out["hgrad"]),
                                      _ptr(out["grad_h"]), _ptr(out["vp_h"]), _ptr(out["hgrad_h"]),
                                      C.byref(gamma), _ptr(rho), _ptr(alpha)))
        out.update(gamma=np.float32(gamma.value), rho=rho, alpha=alpha)
        return out


def read_picks(path, tobs):
    """calc_FA_res.c:28-37 through rz_read_picks: the pick file at `path`
    into tobs[1..ntr] (tobs: float32 numpy array of ntr + 1 values, index 0
    unused, the reference's 1-based vector), in place.  Values that do not
    parse leave their entries as they were.  Host only."""
    if tobs.dtype != np.float32 or not tobs.flags.c_contiguous or tobs.ndim != 1 or tobs.size < 1:
        raise ValueError("tobs must be a contiguous 1-D float32 array of ntr + 1 values")
    rc = lib().rz_read_picks(os.fsencode(path), tobs.ctypes.data_as(C.c_void_p), tobs.size - 1)
    if rc != 0:
        raise OSError(f"pick file {path} could not be opened")
    return tobs


def picks_parses():
    """Number of pick-file parses rz_read_picks has made in this process."""
    return lib().rz_picks_parses()


def picks_clear():
    lib().rz_picks_clear()


class Context:
    """One context over a list of models (dicts or Survey-like objects).

    Each model needs nx, ny, dh, src_ix, src_iy, rec_ix, rec_iy (0-based) and
    optionally time, ttnorm, eps_adj, max_iter.
    """

    def __init__(self, models, device=0, rows_per_lane=0, hbm_budget=0):
        """rows_per_lane: 1, 2, 4, 8; 0 = by band count; ROWS_AUTOTUNE = timed."""
        L = lib()
        descs, self._keep, self.shapes = _descs(models)
        h = C.c_void_p()
        rc = L.rz_create(C.byref(h), device, len(models), descs, rows_per_lane, hbm_budget)
        if rc != RZ_OK:
            raise RzError(rc, L.rz_last_error(None).decode())
        self.h = h
        self._keep = None

    # -- lifecycle -----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            lib().rz_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc):
        if rc != RZ_OK:
            msg = lib().rz_last_error(self.h).decode()
            raise (NotConverged if rc == RZ_ERR_NOT_CONVERGED else RzError)(rc, msg)

    # -- inputs --------------------------------------------------------------
    def set_stream(self, stream_handle):
        """Run on the hipStream_t `stream_handle` (int); 0 / None = HIP's null stream."""
        self._chk(lib().rz_set_stream(self.h, C.c_void_p(stream_handle) if stream_handle else None))

    def get_stream(self):
        return lib().rz_get_stream(self.h) or 0

    def set_slowness(self, model, S):
        """S: numpy (ny, nx) float32, or a device pointer (int)."""
        if isinstance(S, int):
            self._chk(lib().rz_set_slowness(self.h, model, C.c_void_p(S), 1))
            return
        ny, nx = self.shapes[model][:2]
        S = np.ascontiguousarray(S, np.float32)
        assert S.shape == (ny, nx), (S.shape, (ny, nx))
        self._chk(lib().rz_set_slowness(self.h, model, _ptr(S), 0))

    def set_picks(self, model, tobs):
        if isinstance(tobs, int):
            self._chk(lib().rz_set_picks(self.h, model, C.c_void_p(tobs), 1))
            return
        ns, ntr = self.shapes[model][2:]
        tobs = np.ascontiguousarray(tobs, np.float32).reshape(ns, ntr)
        self._chk(lib().rz_set_picks(self.h, model, _ptr(tobs), 0))

    def set_residuals(self, model, dT):
        """Adjoint source dL/dTmod, numpy (nshots, ntr) or a device pointer."""
        if isinstance(dT, int):
            self._chk(lib().rz_set_residuals(self.h, model, C.c_void_p(dT), 1))
            return
        ns, ntr = self.shapes[model][2:]
        dT = np.ascontiguousarray(dT, np.float32).reshape(ns, ntr)
        self._chk(lib().rz_set_residuals(self.h, model, _ptr(dT), 0))

    def adjoint(self):
        self._chk(lib().rz_adjoint(self.h))

    def set_tres_carry(self, model, tres0):
        t = None if tres0 is None else np.ascontiguousarray(tres0, np.float32)
        self._chk(lib().rz_set_tres_carry(self.h, model, _ptr(t)))

    # -- runs ----------------------------------------------------------------
    def forward(self):
        self._chk(lib().rz_forward(self.h))

    def gradient(self, per_shot=False):
        self._chk((lib().rz_gradient_per_shot if per_shot else lib().rz_gradient)(self.h))

    def sync(self):
        self._chk(lib().rz_sync(self.h))

    # -- outputs -------------------------------------------------------------
    def tnb(self, model=0):
        """T at each receiver's slope neighbour row, (nshots, ntr)."""
        ny, nx, ns, ntr = self.shapes[model]
        out = np.empty((ns, ntr), np.float32)
        self._chk(lib().rz_get_tnb(self.h, model, _ptr(out), 0))
        return out

    def tmod(self, model=0, device_ptr=None):
        ny, nx, ns, ntr = self.shapes[model]
        if device_ptr is not None:
            self._chk(lib().rz_get_tmod(self.h, model, C.c_void_p(device_ptr), 1))
            return None
        out = np.empty((ns, ntr), np.float32)
        self._chk(lib().rz_get_tmod(self.h, model, _ptr(out), 0))
        return out

    def tres(self, model=0):
        ny, nx, ns, ntr = self.shapes[model]
        out = np.empty((ns, ntr), np.float32)
        self._chk(lib().rz_get_tres(self.h, model, _ptr(out), 0))
        return out

    def grad(self, model=0, device_ptr=None):
        ny, nx = self.shapes[model][:2]
        if device_ptr is not None:
            self._chk(lib().rz_get_grad(self.h, model, C.c_void_p(device_ptr), 1))
            return None
        out = np.empty((ny, nx), np.float32)
        self._chk(lib().rz_get_grad(self.h, model, _ptr(out), 0))
        return out

    def misfit(self, model=0):
        ns = self.shapes[model][2]
        L2 = C.c_double()
        per = np.empty(max(ns, 1), np.float64)
        self._chk(lib().rz_get_misfit(self.h, model, C.byref(L2), _ptr(per)))
        return L2.value, per[:ns]

    def iterations(self, model=0):
        ns = self.shapes[model][2]
        f = np.empty(max(ns, 1), np.int32)
        a = np.empty(max(ns, 1), np.int32)
        self._chk(lib().rz_get_iterations(self.h, model, _ptr(f), _ptr(a)))
        return f[:ns], a[:ns]

    def stop_passes(self, model=0):
        """Per shot, the stopping decisions of the last run that read the
        fields: counts of changed cells ("count_f", "count_a"), norm1's
        ordered float sums ("ordered_f", "ordered_a"), and the ordered sum of
        the last iteration's decision where one was taken, else -1 ("sum_f",
        "sum_a"); forward and adjoint."""
        ns = self.shapes[model][2]
        keys = ("count_f", "count_a", "ordered_f", "ordered_a", "sum_f", "sum_a")
        out = {k: np.empty(max(ns, 1), np.float32 if k.startswith("sum") else np.int32) for k in keys}
        self._chk(lib().rz_get_stop_passes(self.h, model, *(_ptr(out[k]) for k in keys)))
        return {k: v[:ns] for k, v in out.items()}

    def tres_carry(self, model=0):
        ntr = self.shapes[model][3]
        out = np.empty(max(ntr, 1), np.float32)
        self._chk(lib().rz_get_tres_carry(self.h, model, _ptr(out)))
        return out[:ntr]

    def field(self, model, shot, which):
        ny, nx = self.shapes[model][:2]
        out = np.empty((ny, nx), np.float32)
        self._chk(lib().rz_get_field(self.h, model, shot, 0 if which in (0, "T") else 1, _ptr(out), 0))
        return out

    # -- single-shot reference-surface entry points --------------------------
    def eikonal_shot(self, model, src_ix, src_iy):
        ny, nx, ns, ntr = self.shapes[model]
        T = np.empty((ny, nx), np.float32)
        tm = np.empty(max(ntr, 1), np.float32)
        it = C.c_int()
        self._chk(lib().rz_eikonal_shot(self.h, model, int(src_ix), int(src_iy), _ptr(T), _ptr(tm), C.byref(it)))
        return T, tm[:ntr], it.value

    def eikonal_adj_shot(self, model, T, tres):
        ny, nx, ns, ntr = self.shapes[model]
        T = np.ascontiguousarray(T, np.float32)
        tr = np.ascontiguousarray(tres, np.float32)
        lam = np.empty((ny, nx), np.float32)
        it = C.c_int()
        self._chk(lib().rz_eikonal_adj_shot(self.h, model, _ptr(T), _ptr(tr), _ptr(lam), C.byref(it)))
        return lam, it.value

    def ass_grad_shot(self, model, lam, grad, src=None):
        """ass_grad for one shot: (grad_shot, grad + grad_shot); with `src`
        (the shot's index in the taper's sources) the per-shot taper set by
        set_shot_taper applies."""
        lam = np.ascontiguousarray(lam, np.float32)
        g = np.ascontiguousarray(grad, np.float32).copy()
        gs = np.empty_like(lam)
        if src is None:
            self._chk(lib().rz_ass_grad_shot(self.h, model, _ptr(lam), _ptr(gs), _ptr(g)))
        else:
            self._chk(lib().rz_ass_grad_shot_at(self.h, model, int(src), _ptr(lam), _ptr(gs), _ptr(g)))
        return gs, g

    def set_shot_taper(self, model, mode, radius_m=0.0, filtsize=0, src_x_m=(), src_z_m=(), first_shot=0):
        """Per-shot gradient taper (rz_set_shot_taper): mode 1 radial with
        radius_m, mode 2 windows of half-size filtsize; source positions of
        the whole survey in metres; shot s of the model is source
        first_shot + s."""
        sx = np.ascontiguousarray(src_x_m, np.float32)
        sz = np.ascontiguousarray(src_z_m, np.float32)
        if sx.shape != sz.shape:
            raise ValueError("src_x_m and src_z_m differ in length")
        self._chk(lib().rz_set_shot_taper(self.h, model, int(mode), float(radius_m), int(filtsize), sx.size,
                                          _ptr(sx) if sx.size else None, _ptr(sz) if sz.size else None,
                                          int(first_shot)))

    def set_graph(self, on=True):
        """Replay whole evaluations as HIP graphs (rz_set_graph)."""
        self._chk(lib().rz_set_graph(self.h, 1 if on else 0))

    def set_profiling(self, on=True):
        self._chk(lib().rz_set_profiling(self.h, 1 if on else 0))

    def profile(self):
        """{phase: (total ms, launches)} since the last set_profiling()."""
        ms = np.zeros(4, np.float64)
        n = np.zeros(4, np.int64)
        self._chk(lib().rz_get_profile(self.h, _ptr(ms), _ptr(n)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(("forward", "residual", "adjoint", "gradient"))}

    def set_split(self, on=True):
        """Two workgroups per shot (rz_set_split; on=2 forces the fallback)."""
        self._chk(lib().rz_set_split(self.h, int(on)))

    def set_split_shots(self, n):
        """Split shots per mixed launch (-1 = every CU the others leave, 0 = none)."""
        self._chk(lib().rz_set_split_shots(self.h, n))

    def split_shots(self):
        return lib().rz_get_split_shots(self.h)

    def tune_split(self):
        """Time this context's evaluation with each split-shot candidate and
        keep the fastest; returns (split_shots, ms)."""
        n, ms = C.c_int(), C.c_double()
        self._chk(lib().rz_tune_split(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def compute_units(self):
        return lib().rz_compute_units(self.h)

    def split_state(self):
        """(active, failure) of rz_get_split."""
        a, f = C.c_int(0), C.c_int(0)
        self._chk(lib().rz_get_split(self.h, C.byref(a), C.byref(f)))
        return a.value, f.value

    @property
    def rows_per_lane(self):
        return lib().rz_rows_per_lane(self.h)

    @property
    def device_bytes(self):
        return lib().rz_device_bytes(self.h)

    @property
    def batch_shots(self):
        return lib().rz_batch_shots(self.h)

    @property
    def slot_bytes(self):
        """Device bytes of one shot's fields: the unit of an hbm_budget."""
        return lib().rz_slot_bytes(self.h)
