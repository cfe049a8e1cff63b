"""Arithmetic identities the kernels rely on for bitwise parity, checked on
the CPU with IEEE float32/float64 (numpy rounds to nearest-even like the
GPU's default mode).

* A float quotient x/y equals (float)((double)x * RN64(1/(double)y)): the
  exact quotient of two floats is never closer than about 2^-49 (relative)
  to a float rounding tie unless it is a tie, which the double product then
  also reproduces exactly, while the double product errs by < 2^-52.  The
  adjoint uses it for x/DH (sweep_adj.c:33-54) and for rhs/lhs
  (sweep_adj.c:60-64) with RN64(1/lhs) computed once per cell.
* (x + |x|)/2 and (x - |x|)/2 in float, and max(x, +0) / min(x, +0) with
  -0 ordered below +0 (v_max_f32 / v_min_f32 on gfx950), equal the
  reference's double evaluation rounded to float (sweep_adj.c:39-49),
  signed zeros included.
"""
import numpy as np

f32, f64 = np.float32, np.float64


def _spread_floats(rng, n, lo=-126, hi=127):
    """Random finite floats of both signs with unbiased exponents in
    [lo, hi) (below -126: subnormals)."""
    m = rng.integers(0, 1 << 23, n, dtype=np.uint64)
    e = rng.integers(max(lo + 127, 0), min(hi + 127, 255), n, dtype=np.uint64)   # 0: subnormal
    s = rng.integers(0, 2, n, dtype=np.uint64)
    bits = (s << 31) | (e << 23) | m
    return bits.astype(np.uint32).view(f32)


def _div_equal(x, y):
    with np.errstate(all="ignore"):
        ref = (x / y).astype(f32)
        alt = (x.astype(f64) * (1.0 / y.astype(f64))).astype(f32)
    ok = np.isfinite(ref)
    return np.array_equal(ref[ok].view(np.uint32), alt[ok].view(np.uint32)), int(ok.sum())


def test_quotient_by_double_reciprocal_random():
    rng = np.random.default_rng(1)
    for _ in range(8):
        x = _spread_floats(rng, 1_000_000, -40, 40)
        y = _spread_floats(rng, 1_000_000, -40, 40)
        eq, n = _div_equal(x, y)
        assert eq and n > 900_000


def test_quotient_by_double_reciprocal_wide_and_subnormal_results():
    rng = np.random.default_rng(2)
    x = _spread_floats(rng, 2_000_000, -149, 127)
    y = _spread_floats(rng, 2_000_000, -126, 127)
    eq, _ = _div_equal(x, y)
    assert eq
    # quotients in the subnormal range, including exact ties (y a power of two)
    k = rng.integers(1, 1 << 20, 500_000).astype(f64)
    y = np.ldexp(1.0, rng.integers(1, 30, 500_000)).astype(f32)
    x = (np.ldexp(2 * k + 1, -150) * y.astype(f64)).astype(f32)
    eq, n = _div_equal(x, y)
    assert eq and n > 0


def test_quotient_by_dh_values():
    rng = np.random.default_rng(3)
    x = _spread_floats(rng, 2_000_000, -30, 30)
    for dh in (0.5, 1.0, 2.5, 5.0, 7.5, 10.0, 12.5, 20.0, 3.0, 0.1, 6.25):
        eq, _ = _div_equal(x, np.full_like(x, f32(dh)))
        assert eq, dh


def _max0(x):
    """v_max_f32(x, +0) on gfx950: -0 orders below +0, so max(-0, +0) = +0."""
    return np.where(x > 0, x, f32(0.0)).astype(f32)


def _min0(x):
    """v_min_f32(x, +0) on gfx950: min(-0, +0) = -0 (and +0 for x = +0)."""
    neg_zero = (x == 0) & np.signbit(x)
    return np.where(x < 0, x, np.where(neg_zero, f32(-0.0), f32(0.0))).astype(f32)


def test_positive_and_negative_parts_keep_signed_zeros():
    # |x| < 1.7e38 (slopes of traveltimes are far below it)
    rng = np.random.default_rng(4)
    vals = np.concatenate([np.array([0.0, -0.0, 1.5, -1.5, 1e-45, -1e-45, 1.7e38, -1.7e38], f32),
                           _spread_floats(rng, 100_000, -149, 126)])
    ref_p = ((vals.astype(f64) + np.abs(vals.astype(f64))) / 2.0).astype(f32)
    ref_m = ((vals.astype(f64) - np.abs(vals.astype(f64))) / 2.0).astype(f32)
    # the float form of the formulas
    p = (vals + np.abs(vals)) * f32(0.5)
    m = (vals - np.abs(vals)) * f32(0.5)
    assert np.array_equal(p.view(np.uint32), ref_p.view(np.uint32))
    assert np.array_equal(m.view(np.uint32), ref_m.view(np.uint32))
    # the kernels' form: max(x, +0) and min(x, +0) with -0 below +0
    assert np.array_equal(_max0(vals).view(np.uint32), ref_p.view(np.uint32))
    assert np.array_equal(_min0(vals).view(np.uint32), ref_m.view(np.uint32))
