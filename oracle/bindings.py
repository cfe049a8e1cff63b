"""ctypes bindings of the checker libraries.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

``Restatement(f64=False)`` wraps oracle/_build/librzoracle.so (float build)
or librzoracle64.so (double build); ``Reference()`` wraps
oracle/_ref/librzref.so, the reference's own sources.  All arrays are numpy,
row-major (NY, NX); restatement indices are 0-based, reference indices
1-based (converted here), positions in metres.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = np.ctypeslib.ndpointer
_f32 = _P(np.float32, flags="C_CONTIGUOUS")
_f64 = _P(np.float64, flags="C_CONTIGUOUS")
_i32 = _P(np.int32, flags="C_CONTIGUOUS")


def build():
    """Compile the checker libraries (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _opt(ptr_type):
    """ndpointer that also accepts None."""
    class _T(ptr_type):
        @classmethod
        def from_param(cls, obj):
            if obj is None:
                return None
            return ptr_type.from_param(obj)
    return _T


class Restatement:
    """The CPU restatement, float (default) or double build."""

    def __init__(self, f64=False):
        name = "librzoracle64.so" if f64 else "librzoracle.so"
        path = os.path.join(HERE, "_build", name)
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        self.f64 = f64
        self.dt = np.float64 if f64 else np.float32
        fr = _f64 if f64 else _f32
        sfx = "_f64" if f64 else ""
        L = self.lib

        def fn(n):
            return getattr(L, n + sfx)

        self._sweep = fn("rzo_sweep")
        self._sweep.argtypes = [C.c_int, C.c_int, C.c_float, fr, fr, C.c_int, C.c_int]
        self._sweep.restype = None
        self._eik = fn("rzo_eikonal")
        self._eik.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, fr,
                              C.c_int, C.c_int, fr]
        self._eik.restype = C.c_int
        self._res = fn("rzo_calc_fa_res")
        self._res.argtypes = [C.c_int, C.c_int, C.c_float, fr, C.c_int, _i32, _i32, fr, fr, fr]
        self._res.restype = C.c_double
        self._flag = fn("rzo_init_recflag")
        self._flag.argtypes = [C.c_int, C.c_int, C.c_int, _i32, _i32, _i32]
        self._flag.restype = None
        self._initadj = fn("rzo_init_travel_adj")
        self._initadj.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int, _i32, _i32, fr, fr]
        self._initadj.restype = None
        self._sadj = fn("rzo_sweep_adj")
        self._sadj.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, fr, _i32, fr, C.c_int, C.c_int]
        self._sadj.restype = None
        self._adj = fn("rzo_eikonal_adj")
        self._adj.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                              fr, C.c_int, _i32, _i32, fr, _i32, fr]
        self._adj.restype = C.c_int
        self._ass = fn("rzo_ass_grad")
        self._ass.argtypes = [C.c_int, C.c_int, fr, fr, fr, fr]
        self._ass.restype = None
        self._frame = fn("rzo_cp_grad_frame")
        self._frame.argtypes = [C.c_int, C.c_int, fr]
        self._frame.restype = None
        o_fr, o_d, o_i = _opt(fr), _opt(_f64), _opt(_i32)
        self._go = fn("rzo_grad_obj")
        self._go.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                             fr, C.c_int, _i32, _i32, C.c_int, _i32, _i32, fr, fr, fr, fr, fr,
                             o_fr, o_fr, o_d, o_i, o_i]
        self._go.restype = C.c_double
        self._fwd = fn("rzo_forward")
        self._fwd.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, fr,
                              C.c_int, _i32, _i32, C.c_int, _i32, _i32, fr, fr, o_i]
        self._fwd.restype = None
        self._taper = fn("rzo_set_shot_taper")
        self._taper.argtypes = [C.c_int, C.c_float, C.c_int, C.c_int, _f32, _f32, C.c_int]
        self._taper.restype = None
        self._taper_keep = None
        self._gauss = fn("rzo_gauss_filt")
        self._gauss.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), fr]
        self._gauss.restype = C.c_int

    def _a(self, x):
        return np.ascontiguousarray(x, dtype=self.dt)

    @staticmethod
    def _i(x):
        return np.ascontiguousarray(x, dtype=np.int32)

    def sweep(self, S, T, sx, sy, dh):
        T = self._a(T).copy()
        ny, nx = T.shape
        self._sweep(nx, ny, dh, self._a(S), T, sx, sy)
        return T

    def eikonal(self, S, isx, isy, dh, time=6.0, ttnorm=1e-10, maxit=0):
        S = self._a(S)
        ny, nx = S.shape
        T = np.empty_like(S)
        it = self._eik(nx, ny, dh, time, ttnorm, maxit, S, int(isx), int(isy), T)
        return T, it

    def calc_fa_res(self, T, rx, ry, tmod, tobs, tres, dh):
        T = self._a(T)
        ny, nx = T.shape
        tres = self._a(tres).copy()
        l2 = self._res(nx, ny, dh, T, len(rx), self._i(rx), self._i(ry), self._a(tmod), self._a(tobs), tres)
        return tres, l2

    def init_recflag(self, nx, ny, rx, ry):
        f = np.empty((ny, nx), np.int32)
        self._flag(nx, ny, len(rx), self._i(rx), self._i(ry), f)
        return f

    def init_travel_adj(self, nx, ny, time, rx, ry, tres):
        lam = np.empty((ny, nx), self.dt)
        self._initadj(nx, ny, time, len(rx), self._i(rx), self._i(ry), self._a(tres), lam)
        return lam

    def sweep_adj(self, lam, T, flag, sx, sy, dh, eps_adj=1e-9):
        lam = self._a(lam).copy()
        ny, nx = lam.shape
        self._sadj(nx, ny, dh, eps_adj, self._a(T), self._i(flag), lam, sx, sy)
        return lam

    def eikonal_adj(self, T, rx, ry, tres, dh, time=6.0, ttnorm=1e-10, eps_adj=1e-9, maxit=0):
        T = self._a(T)
        ny, nx = T.shape
        flag = self.init_recflag(nx, ny, rx, ry)
        lam = np.empty_like(T)
        it = self._adj(nx, ny, dh, time, ttnorm, eps_adj, maxit, T, len(rx), self._i(rx), self._i(ry),
                       self._a(tres), flag, lam)
        return lam, it

    def ass_grad(self, S, lam, grad):
        S = self._a(S)
        ny, nx = S.shape
        g = self._a(grad).copy()
        gs = np.empty_like(S)
        self._ass(nx, ny, S, self._a(lam), gs, g)
        return gs, g

    def cp_grad_frame(self, A):
        A = self._a(A).copy()
        ny, nx = A.shape
        self._frame(nx, ny, A)
        return A

    def set_shot_taper(self, mode, radius=80.0, filtsize=1, src_x_m=(0.0,), src_z_m=(0.0,), first_shot=0):
        """Per-shot gradient taper of grad_obj (ass_grad.c:31-37): mode =
        SWS_TAPER_CIRCULAR_PER_SHOT, radius = SRTRADIUS (m), filtsize =
        FILTSIZE; src_x_m/src_z_m the sources of every shot of the survey
        (metres), of which grad_obj's shots start at first_shot."""
        sx = np.ascontiguousarray(src_x_m, np.float32)
        sz = np.ascontiguousarray(src_z_m, np.float32)
        self._taper_keep = (sx, sz)
        self._taper(int(mode), float(radius), int(filtsize), sx.size, sx, sz, int(first_shot))

    def gauss_filt(self, A, fsx, fsy):
        """gauss_filt.c on a copy of A; returns (filtered, fsx, fsy as made odd)."""
        A = self._a(A).copy()
        ny, nx = A.shape
        x, y = C.c_int(fsx), C.c_int(fsy)
        self._gauss(nx, ny, C.byref(x), C.byref(y), A)
        return A, x.value, y.value

    def grad_obj(self, S, src_ix, src_iy, rec_ix, rec_iy, tobs, dh, time=6.0, ttnorm=1e-10,
                 eps_adj=1e-9, tres0=None, maxit=0):
        """One evaluation; returns dict of L2, grad, last T/lam, per-shot data."""
        S = self._a(S)
        ny, nx = S.shape
        ns, ntr = len(src_ix), len(rec_ix)
        tres = self._a(np.zeros(ntr) if tres0 is None else tres0).copy()
        grad = np.empty_like(S)
        T = np.empty_like(S)
        lam = np.empty_like(S)
        tmod = np.empty((ns, ntr), self.dt)
        tres_all = np.empty((ns, ntr), self.dt)
        l2s = np.empty(ns, np.float64)
        itf = np.empty(ns, np.int32)
        ita = np.empty(ns, np.int32)
        L2 = self._go(nx, ny, dh, time, ttnorm, eps_adj, maxit, S, ns, self._i(src_ix), self._i(src_iy),
                      ntr, self._i(rec_ix), self._i(rec_iy), self._a(tobs).reshape(ns, ntr), tres, grad,
                      T, lam, tmod, tres_all, l2s, itf, ita)
        return dict(L2=L2, grad=grad, T=T, lam=lam, tmod=tmod, tres=tres_all, tres_carry=tres,
                    l2_shot=l2s, it_f=itf, it_a=ita)

    def forward(self, S, src_ix, src_iy, rec_ix, rec_iy, dh, time=6.0, ttnorm=1e-10, maxit=0):
        S = self._a(S)
        ny, nx = S.shape
        ns, ntr = len(src_ix), len(rec_ix)
        tmod = np.empty((ns, ntr), self.dt)
        T = np.empty_like(S)
        itf = np.empty(ns, np.int32)
        self._fwd(nx, ny, dh, time, ttnorm, maxit, S, ns, self._i(src_ix), self._i(src_iy), ntr,
                  self._i(rec_ix), self._i(rec_iy), tmod, T, itf)
        return tmod, T, itf


class Reference:
    """The reference's own hot-path sources (oracle/_ref/librzref.so)."""

    PATH = os.path.join(HERE, "_ref", "librzref.so")

    @classmethod
    def available(cls):
        return os.path.exists(cls.PATH)

    def __init__(self):
        if not self.available():
            build()
        self.lib = C.CDLL(self.PATH)
        L = self.lib
        L.rzref_setup.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]
        L.rzref_setup.restype = None
        L.rzref_sweep.argtypes = [_f32, _f32, C.c_int, C.c_int]
        L.rzref_sweep.restype = None
        L.rzref_sweep_adj_dir.argtypes = [_f32, _f32, _i32, C.c_int, C.c_int]
        L.rzref_sweep_adj_dir.restype = None
        L.rzref_eikonal.argtypes = [_f32, C.c_float, C.c_float, C.c_int, _i32, _i32, _f32, _f32]
        L.rzref_eikonal.restype = C.c_int
        L.rzref_eikonal_adj.argtypes = [_f32, C.c_int, _i32, _i32, _f32, _f32]
        L.rzref_eikonal_adj.restype = C.c_int
        L.rzref_init_recflag.argtypes = [C.c_int, _i32, _i32, _i32]
        L.rzref_init_recflag.restype = None
        L.rzref_ass_grad.argtypes = [_f32, _f32, _f32, _f32]
        L.rzref_ass_grad.restype = None
        o_f, o_d, o_i = _opt(_f32), _opt(_f64), _opt(_i32)
        L.rzref_grad_obj.argtypes = [_f32, C.c_int, _f32, _f32, C.c_int, _i32, _i32, _f32, C.c_char_p,
                                     _f32, _f32, _f32, _f32, o_f, o_f, o_f, o_d, o_i, o_i, C.c_int]
        L.rzref_grad_obj.restype = C.c_double
        L.rzref_gauss_filt.argtypes = [_f32, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.rzref_gauss_filt.restype = None
        L.rzref_set_taper.argtypes = [C.c_int, C.c_float, C.c_int]
        L.rzref_set_taper.restype = None
        L.rzref_ass_grad_at.argtypes = [_f32, _f32, C.c_int, _f32, _f32, C.c_int, _i32, _i32, C.c_int, _f32, _f32]
        L.rzref_ass_grad_at.restype = None
        L.rzref_taper_grad.argtypes = [_f32, C.c_int, _f32, _f32, C.c_int, _i32, _i32, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.c_float, C.c_float]
        L.rzref_taper_grad.restype = None
        L.rzref_lbfgs.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _f32, _f32, _f32,
                                  _f32, _f32, _f32, _f32, _f32, _f32, _f32, _f32]
        L.rzref_lbfgs.restype = None
        L.rzref_smooth_grad.argtypes = [_f32, C.POINTER(C.c_int)]
        L.rzref_smooth_grad.restype = None
        L.rzref_smooth2.argtypes = [_f32, C.c_float, C.c_float, C.c_char_p]
        L.rzref_smooth2.restype = None

    def smooth_grad(self, A, fsize):
        """The reference's median filter (SPATFILTER = 3); returns (field,
        FILT_SIZE_GRAD as it leaves it)."""
        a = self._f(A).copy()
        fs = C.c_int(int(fsize))
        self.lib.rzref_smooth_grad(a, C.byref(fs))
        return a, fs.value

    def smooth2(self, A, wd_damp, wd_damp1, jac):
        """The reference's smooth2 (SPATFILTER = 2) with its files under jac."""
        a = self._f(A).copy()
        self.lib.rzref_smooth2(a, wd_damp, wd_damp1, os.fsencode(jac))
        return a

    def lbfgs_state(self, nlbfgs):
        """Zeroed history and work vectors for lbfgs(), 1-based as fatt.c
        allocates them (fatt.c:74-82)."""
        n = self.nx * self.ny
        return {"y": np.zeros(nlbfgs * n + 1, np.float32), "s": np.zeros(nlbfgs * n + 1, np.float32),
                "q": np.zeros(n + 1, np.float32), "r": np.zeros(n + 1, np.float32),
                "rho": np.zeros(nlbfgs + 1, np.float32), "alpha": np.zeros(nlbfgs + 1, np.float32),
                "beta": np.zeros(nlbfgs + 1, np.float32), "nlbfgs": nlbfgs}

    def lbfgs(self, state, jac, it, pointer, grad, gradm, vp, hgrad=None, idx=1, idy=1):
        """The reference's LBFGS() on row-major (NY, NX) fields; state from
        lbfgs_state() is updated in place.  Returns Hgrad (cells it does not
        write keep hgrad's values, zeros by default)."""
        H = np.zeros((self.ny, self.nx), np.float32) if hgrad is None else self._f(hgrad).copy()
        st = state
        self.lib.rzref_lbfgs(os.fsencode(jac), idx, idy, it, pointer, st["nlbfgs"], st["y"], st["s"], st["rho"],
                             st["alpha"], st["beta"], st["q"], st["r"], self._f(grad), self._f(gradm),
                             self._f(vp), H)
        return H

    def setup(self, nx, ny, dh, time=6.0, ttnorm=1e-10, eps_adj=1e-9):
        self.nx, self.ny = nx, ny
        self.lib.rzref_setup(nx, ny, dh, time, ttnorm, eps_adj)

    def set_taper(self, sws, srtradius=80.0, filtsize=1):
        """SWS_TAPER_CIRCULAR_PER_SHOT, SRTRADIUS (m), FILTSIZE for grad_obj
        (reset to 0 by setup)."""
        self.lib.rzref_set_taper(sws, srtradius, filtsize)

    @staticmethod
    def _f(x):
        return np.ascontiguousarray(x, dtype=np.float32)

    @staticmethod
    def _i1(x):
        return np.ascontiguousarray(np.asarray(x) + 1, dtype=np.int32)  # 0-based -> 1-based

    def sweep(self, S, T, sx, sy):
        T = self._f(T).copy()
        self.lib.rzref_sweep(self._f(S), T, sx, sy)
        return T

    def sweep_adj(self, lam, T, flag, sx, sy):
        lam = self._f(lam).copy()
        self.lib.rzref_sweep_adj_dir(lam, self._f(T), np.ascontiguousarray(flag, np.int32), sx, sy)
        return lam

    def eikonal(self, S, srcx_m, srcy_m, rec_ix=(), rec_iy=()):
        T = np.empty((self.ny, self.nx), np.float32)
        ntr = len(rec_ix)
        tmod = np.empty(max(ntr, 1), np.float32)
        it = self.lib.rzref_eikonal(self._f(S), srcx_m, srcy_m, ntr, self._i1(rec_ix), self._i1(rec_iy), T, tmod)
        return T, tmod[:ntr], it

    def eikonal_adj(self, T, rec_ix, rec_iy, tres):
        lam = np.empty((self.ny, self.nx), np.float32)
        it = self.lib.rzref_eikonal_adj(self._f(T), len(rec_ix), self._i1(rec_ix), self._i1(rec_iy),
                                        self._f(tres), lam)
        return lam, it

    def init_recflag(self, rec_ix, rec_iy):
        f = np.empty((self.ny, self.nx), np.int32)
        self.lib.rzref_init_recflag(len(rec_ix), self._i1(rec_ix), self._i1(rec_iy), f)
        return f

    def ass_grad(self, S, lam, grad):
        g = self._f(grad).copy()
        gs = np.empty((self.ny, self.nx), np.float32)
        self.lib.rzref_ass_grad(self._f(S), self._f(lam), gs, g)
        return gs, g

    def ass_grad_at(self, S, lam, grad, src_x_m, src_z_m, rec_ix, rec_iy, ishot):
        """The reference's ass_grad() for shot ishot (0-based) of the given
        sources (metres) and receivers (0-based cells), with the per-shot
        taper of set_taper."""
        g = self._f(grad).copy()
        gs = np.empty((self.ny, self.nx), np.float32)
        sx, sz = self._f(src_x_m), self._f(src_z_m)
        self.lib.rzref_ass_grad_at(self._f(S), self._f(lam), len(sx), sx, sz, len(rec_ix), self._i1(rec_ix),
                                   self._i1(rec_iy), int(ishot) + 1, gs, g)
        return gs, g

    def taper_grad(self, A, sws, gradt=(0, 0, 0, 0), exp_hor=0.0, srtradius=80.0, src_x_m=(), src_z_m=(),
                   rec_ix=(), rec_iy=(), taper_file=None):
        """The reference's taper_grad() (NX, NY, DH from setup) on a copy of A:
        returns (tapered field, taper_coeff as written to taper_coeff_*.bin
        or None).  It runs in a scratch directory: taper_file (NX*NY floats,
        x outer) is written there as taper.bin for sws 4..6."""
        import tempfile
        A = self._f(A).copy()
        sx, sz = self._f(np.asarray(src_x_m, np.float32)), self._f(np.asarray(src_z_m, np.float32))
        name = {1: "taper_coeff_vert.bin", 2: "taper_coeff_hor.bin"}.get(sws, "taper_coeff_file.bin")
        cwd = os.getcwd()
        with tempfile.TemporaryDirectory(dir="/tmp") as d:
            os.chdir(d)
            try:
                if taper_file is not None:
                    np.ascontiguousarray(taper_file, np.float32).tofile("taper.bin")
                self.lib.rzref_taper_grad(A, len(sx), sx, sz, len(rec_ix), self._i1(rec_ix), self._i1(rec_iy), sws,
                                          *[int(g) for g in gradt], exp_hor, srtradius)
                coeff = None
                if sws != 3 and os.path.exists(name):   # x outer, y inner
                    coeff = np.fromfile(name, np.float32).reshape(self.nx, self.ny).T.copy()
            finally:
                os.chdir(cwd)
        return A, coeff

    def gauss_filt(self, A, fsx, fsy):
        """The reference's gauss_filt() (NX, NY from setup) on a copy of A."""
        A = self._f(A).copy()
        x, y = C.c_int(fsx), C.c_int(fsy)
        self.lib.rzref_gauss_filt(A, C.byref(x), C.byref(y))
        return A, x.value, y.value

    def grad_obj(self, S, src_x_m, src_z_m, rec_ix, rec_iy, tobs, tres0=None, workdir=None):
        ns, ntr = len(src_x_m), len(rec_ix)
        tres = self._f(np.zeros(ntr) if tres0 is None else tres0).copy()
        out = dict(grad=np.empty((self.ny, self.nx), np.float32), T=np.empty((self.ny, self.nx), np.float32),
                   lam=np.empty((self.ny, self.nx), np.float32), tobs=np.empty((ns, ntr), np.float32),
                   tmod=np.empty((ns, ntr), np.float32), tres=np.empty((ns, ntr), np.float32),
                   l2_shot=np.empty(ns, np.float64), it_f=np.empty(ns, np.int32), it_a=np.empty(ns, np.int32))
        with tempfile.TemporaryDirectory(dir=workdir) as d:
            stem = os.path.join(d, "p")
            if len(stem) >= 73:
                raise ValueError("pick path too long for the reference's 74-char buffers")
            L2 = self.lib.rzref_grad_obj(self._f(S), ns, self._f(src_x_m), self._f(src_z_m), ntr,
                                         self._i1(rec_ix), self._i1(rec_iy), self._f(tobs).reshape(ns, ntr),
                                         stem.encode(), tres, out["grad"], out["T"], out["lam"], out["tobs"],
                                         out["tmod"], out["tres"], out["l2_shot"], out["it_f"], out["it_a"], 1)
        out["L2"] = L2
        out["tres_carry"] = tres
        return out
