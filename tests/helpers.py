"""Shared helpers for the parity tests (no GPU needed to import)."""
import numpy as np


def bits_equal_nan(a, b):
    """bits_equal, except that a NaN matches any NaN: NaN payloads are not
    carried across machines (x86's default NaN has the sign bit set, the
    GPU's does not)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return bits_equal(np.where(na, 0, a).astype(a.dtype), np.where(nb, 0, b).astype(b.dtype))


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    if a.dtype == np.float32:
        return np.array_equal(a.view(np.uint32), b.view(np.uint32))
    if a.dtype == np.float64:
        return np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)


def ndiff(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    v = np.uint32 if a.dtype == np.float32 else np.uint64
    return int(np.count_nonzero(a.view(v) != b.view(v)))


def oracle_picks(ora, sv, seed=None):
    """Picks from the restatement's forward on the true model plus noise."""
    from rajzel_amd import workloads as W

    def fwd(S, s):
        tm, _, _ = ora.forward(S, s.src_ix, s.src_iy, s.rec_ix, s.rec_iy, s.dh, s.time, s.ttnorm)
        return tm
    return W.synthesize_picks(sv, fwd, seed)


def model_dict(sv, **kw):
    d = dict(nx=sv.nx, ny=sv.ny, dh=sv.dh, time=sv.time, ttnorm=sv.ttnorm, eps_adj=sv.eps_adj,
             src_ix=sv.src_ix, src_iy=sv.src_iy, rec_ix=sv.rec_ix, rec_iy=sv.rec_iy)
    d.update(kw)
    return d
