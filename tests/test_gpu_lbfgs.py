"""The l-BFGS two-loop recursion on the device (rz_lbfgs_*, LBFGS.c:12-226)
and dotp.c's serial float dot product (rz_dotp_ordered, dotp.c:22-36).

The dot product is checked against the float chain itself (numpy's
add.accumulate in float32 adds left to right, one rounding per term) on
vectors built to stress the exact parallel evaluation: sign changes,
cancellation, terms spanning every binade, subnormal sums, exact ties,
infinities and NaN.  The recursion is checked against the reference's own
LBFGS() (oracle/_ref/librzref.so) over a run of iterations whose history
slots wrap, with a history reset in between: direction, rho, alpha and the
three history files, bit for bit."""
import os
import tempfile

import numpy as np
import pytest

from oracle.bindings import Reference
from rajzel_amd import native as N

pytestmark = pytest.mark.gpu


def serial_dot(a, b):
    p = np.asarray(a, np.float32) * np.asarray(b, np.float32)
    return np.add.accumulate(np.concatenate([np.zeros(1, np.float32), p]), dtype=np.float32)[-1]


def same(x, y):
    x, y = np.float32(x), np.float32(y)
    if np.isnan(x) or np.isnan(y):
        return bool(np.isnan(x) and np.isnan(y))
    return x.view(np.uint32) == y.view(np.uint32)


def _cases():
    rng = np.random.default_rng(11)
    out = []
    for n in (0, 1, 15, 16, 17, 1000, 16383, 16384, 16385, 100003):
        out.append((f"normal{n}", rng.standard_normal(n), rng.standard_normal(n)))
    n = 3_004_001
    out.append(("walk3M", rng.standard_normal(n), rng.standard_normal(n)))
    out.append(("positive3M", np.abs(rng.standard_normal(n)), np.abs(rng.standard_normal(n))))
    e = rng.integers(-60, 60, 200_000)
    out.append(("binades", rng.standard_normal(e.size) * 2.0 ** e, rng.standard_normal(e.size) * 2.0 ** -e))
    # products near and below the smallest normal: a subnormal running sum
    out.append(("subnormal", rng.standard_normal(50_000) * 3e-20, rng.standard_normal(50_000) * 3e-20))
    # exact ties on a sum near 1: multiples of a quarter unit
    k = rng.integers(-8, 9, 60_000).astype(np.float64)
    a = np.concatenate([[1.0], k * 2.0 ** -25, [-1.0], k[:1000] * 2.0 ** -30])
    out.append(("ties", a, np.ones_like(a)))
    # cancellation to zero and sign changes around it
    big = rng.standard_normal(5000) * 1e6
    out.append(("cancel", np.concatenate([big, -big[::-1], rng.standard_normal(100)]), np.ones(10100)))
    out.append(("zeros", np.zeros(40_000), -np.ones(40_000)))
    # runs of zeros before, between and after nonzero terms (whole 64-term runs and parts)
    z = lambda k: np.zeros(k)
    a = np.concatenate([z(192), rng.standard_normal(3000), z(1024), rng.standard_normal(70), z(8192 + 64),
                        rng.standard_normal(20_000), z(333)])
    out.append(("zero_runs", a, rng.standard_normal(a.size)))
    out.append(("negzero_first", np.concatenate([[-0.0], rng.standard_normal(99)]), np.ones(100)))
    # overflow to inf, then inf - inf
    out.append(("overflow", np.full(100, 3e38), np.ones(100)))
    out.append(("inf_nan", np.concatenate([rng.standard_normal(5000), [np.inf], rng.standard_normal(10),
                                           [-np.inf], rng.standard_normal(10)]), np.ones(5022)))
    out.append(("nan", np.concatenate([rng.standard_normal(70_000), [np.nan], rng.standard_normal(10)]),
                np.ones(70_011)))
    return out


@pytest.mark.parametrize("name,a,b", _cases(), ids=[c[0] for c in _cases()])
def test_dotp_ordered_is_the_serial_float_sum(name, a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    print(f"[case] dotp {name} n={a.size}", flush=True)
    with np.errstate(all="ignore"):
        want = serial_dot(a, b)
    got = N.dotp_ordered(a, b)
    assert same(got, want), f"{name}: {got!r} != {want!r}"


def _iterations(nx, ny, nits, rng):
    vp = (2000.0 + 500.0 * rng.standard_normal((ny, nx))).astype(np.float32)
    for it in range(1, nits + 1):
        g = (rng.standard_normal((ny, nx)) * 10.0 ** rng.uniform(-8, -6)).astype(np.float32)
        g[rng.random((ny, nx)) < 0.05] = 0.0           # zero gradient cells: gradm = -0 there
        g[:, :3] = 0.0                                 # whole zero columns: zero runs in h order
        g[:, nx // 2:nx // 2 + 2] = 0.0
        yield it, g, -g, vp.copy()
        vp = (vp + rng.standard_normal((ny, nx)).astype(np.float32) * np.float32(5.0)).astype(np.float32)


def _pointer(it, K):
    # fatt.c:108-119: the pointer moves on from the third iteration, wrapping at NLBFGS
    return 1 if it <= 1 else (it - 2) % K + 1


def _run(nx, ny, K, nits, reset_at=None, seed=3):
    ref = Reference()
    ref.setup(nx, ny, 5.0)
    st = ref.lbfgs_state(K)
    dev = N.LBFGS(nx, ny, K)
    rng = np.random.default_rng(seed)
    with tempfile.TemporaryDirectory(prefix="rzlb", dir="/tmp") as d:
        jac = os.path.join(d, "j")
        for it, g, gm, vp in _iterations(nx, ny, nits, rng):
            p = _pointer(it, K)
            print(f"[case] lbfgs {ny}x{nx} K={K} iter={it} pointer={p}", flush=True)
            H = ref.lbfgs(st, jac, it, p, g, gm, vp)
            out = dev.step(it, p, g, gm, vp)
            assert np.array_equal(out["hgrad"].view(np.uint32), H.view(np.uint32)), f"direction, iter {it}"
            for suffix, key in (("_p.old", "grad_h"), ("_p_vp.old", "vp_h"), ("_c.old", "hgrad_h")):
                ref_file = np.fromfile(jac + suffix, np.float32)
                assert np.array_equal(ref_file.view(np.uint32), out[key].view(np.uint32)), f"{suffix}, iter {it}"
            if it > 1:
                assert np.array_equal(out["rho"].view(np.uint32), st["rho"][1:].view(np.uint32)), f"rho, iter {it}"
                assert np.array_equal(out["alpha"].view(np.uint32), st["alpha"][1:].view(np.uint32)), \
                    f"alpha, iter {it}"
            if reset_at is not None and it == reset_at:
                st["y"][:] = 0.0                                   # zero_LBFGS1 (check_descent.c:39)
                st["s"][:] = 0.0
                dev.zero()
    dev.close()


def test_lbfgs_small_grid_wrapping_history():
    _run(53, 37, K=3, nits=7)


def test_lbfgs_history_reset():
    _run(61, 29, K=4, nits=6, reset_at=3, seed=5)


def test_lbfgs_1001x3001_twenty_slots():
    _run(3001, 1001, K=20, nits=4, seed=7)


def test_lbfgs_rejects_out_of_order_calls():
    dev = N.LBFGS(10, 8, 2)
    z = np.zeros((8, 10), np.float32)
    with pytest.raises(N.RzError):
        dev.step(2, 1, z, z, z)           # iter 2 without a previous model and gradient
    with pytest.raises(N.RzError):
        dev.step(1, 3, z, z, z)           # pointer beyond NLBFGS
    dev.close()
