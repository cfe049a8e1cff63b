"""The method by which the kernels take norm1's float sum in the reference's
order (norm1.c:21-25: sum += |TT - TTold| with a float accumulator), checked
on the CPU against a sequential float32 sum.

A running float sum s in the binade [2^e, 2^(e+1)) (e >= -126) moves on a
grid of spacing u = 2^(e-23); s + d rounds to s plus d rounded to that grid,
where a tie (d/u exactly half-way between two integers) goes to the even
neighbour and so depends on the parity of s/u alone, until the sum leaves
the binade.  The kernels (ordered_norm1 in rz_kernels.hip) treat a run of
terms as a function of the incoming parity, scan those functions, and add
only the terms that leave a binade as floats.  This restates that method in
exact Python integers, chunk by chunk and run by run as the kernel does, and
requires the same bits as numpy's sequential float32 accumulate.  A planted
mistake (ties rounded up instead of to even) must fail.
"""
import numpy as np
import pytest


def _split(v):
    b = int(np.float32(v).view(np.uint32))
    E = (b >> 23) & 0xFF
    M = (b & 0x7FFFFF) | (0x800000 if E else 0)
    return M, (E if E else 1) - 150


def _inc(d, e):
    """d in units of 2^(e-23): (floor, tie, up, big)."""
    M, q = _split(d)
    if M == 0:
        return 0, False, False, False
    sh = q - e + 23
    if sh >= 24:
        return 0, False, False, True
    if sh >= 0:
        return M << sh, False, False, False
    r = -sh
    if r > 25:
        return 0, False, False, False
    f, rem, half = M >> r, M & ((1 << r) - 1), 1 << (r - 1)
    return f, rem == half, rem > half, False


def _inc_at(v, p, ties_to_even=True):
    f, tie, up, _ = v
    if tie:
        return f + (((p + f) & 1) if ties_to_even else 1)
    return f + (1 if up else 0)


def _state(s):
    b = int(np.float32(s).view(np.uint32))
    E = (b >> 23) & 0xFF
    return (E if E else 1) - 127, (b & 0x7FFFFF) | (0x800000 if E else 0)


def _fn(vals, ties_to_even):
    """A run of terms as a function of the incoming parity: [(inc, out)]."""
    out = []
    for p in (0, 1):
        t, q = 0, p
        for v in vals:
            a = _inc_at(v, q, ties_to_even)
            t, q = t + a, (q + a) & 1
        out.append((t, q))
    return out


def _compose(F, G):
    return [(F[p][0] + G[F[p][1]][0], G[F[p][1]][1]) for p in (0, 1)]


def ordered_sum(terms, chunk=64, per=4, ties_to_even=True):
    """The kernel's method: per chunk, runs of `per` terms (one thread each)
    as parity functions, an exclusive scan over them, a walk from each run's
    exact incoming sum, and a float add of the first term that leaves the
    binade."""
    d = np.asarray(terms, np.float32)
    s = np.float32(0.0)
    for q0 in range(0, d.size, chunk):
        part = d[q0:q0 + chunk]
        start = 0
        while start < part.size:
            if s == 0:
                nz = np.nonzero(part[start:])[0]
                if nz.size == 0:
                    break
                f = start + int(nz[0])
                s = part[f]
                start = f + 1
                continue
            e, N = _state(s)
            runs = [[(i, _inc(part[i], e)) for i in range(r0, min(r0 + per, part.size)) if i >= start]
                    for r0 in range(0, part.size, per)]
            pre = [(0, 0), (0, 1)]
            event = None
            for run in runs:
                base = pre[N & 1][0]
                acc = N + base
                for i, v in run:
                    a = _inc_at(v, acc & 1, ties_to_even)
                    if v[3] or acc + a >= 1 << 24:
                        event = (i, acc)
                        break
                    acc += a
                if event:
                    break
                pre = _compose(pre, _fn([v for _, v in run], ties_to_even))
            if event is None:
                s = np.float32(np.ldexp(float(N + pre[N & 1][0]), e - 23))
                break
            j, acc = event
            s = np.float32(np.float32(np.ldexp(float(acc), e - 23)) + part[j])
            start = j + 1
    return s


def sequential(terms):
    d = np.asarray(terms, np.float32)
    return np.add.accumulate(d, dtype=np.float32)[-1] if d.size else np.float32(0)


def _cases():
    rng = np.random.default_rng(5)
    yield "uniform", rng.uniform(0, 1, 5000)
    yield "lognormal", np.exp(rng.normal(-10, 6, 5000))
    yield "sparse", np.where(rng.uniform(size=5000) < 0.02, rng.uniform(0, 1e-3, 5000), 0.0)
    yield "subnormal", rng.integers(0, 2000, 3000) * float(np.float32(2.0 ** -149))
    # ties: halves of the grid once the sum sits in [1, 2) and [2, 4)
    yield "ties", np.concatenate([[1.0], np.full(3000, 2.0 ** -24), [1.5], np.full(2000, 2.0 ** -23),
                                  rng.integers(1, 8, 2000) * 2.0 ** -24])
    yield "growing", np.arange(1, 4000) * 1e-7
    yield "huge_then_tiny", np.concatenate([[3e5], rng.uniform(0, 1e-3, 4000), [7e6], np.full(500, 0.25)])
    yield "traveltime_like", np.abs(rng.normal(0, 3e-9, 20000)) * (rng.uniform(size=20000) < 0.3)


@pytest.mark.parametrize("name,terms", list(_cases()), ids=[c[0] for c in _cases()])
@pytest.mark.parametrize("chunk", [7, 64, 1000])
def test_binade_runs_equal_sequential_float32(name, terms, chunk):
    got = ordered_sum(terms, chunk, per=max(1, chunk // 8))
    want = sequential(terms)
    assert np.float32(got).view(np.uint32) == np.float32(want).view(np.uint32), (got, want)


def test_planted_tie_rounding_is_detected():
    """Rounding ties up instead of to the even neighbour gives other bits."""
    terms = np.concatenate([[1.0], np.full(3000, 2.0 ** -24), [1.5], np.full(100, 3 * 2.0 ** -25)])
    assert ordered_sum(terms, 64, 8) == sequential(terms)
    assert ordered_sum(terms, 64, 8, ties_to_even=False) != sequential(terms)
