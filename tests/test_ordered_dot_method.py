"""The method by which rz_dotp_ordered and the l-BFGS recursion take
dotp.c's serial float dot product (dotp.c:22-36: tmp += a[i]*b[i] with a
float accumulator), checked on the CPU against a sequential float32 sum.

While the running sum s keeps its sign and stays in one binade (or, below
2^-125, in the grid of step 2^-149), s + d is s plus d rounded to that grid:
floor(d/u) plus one above one half, and a tie goes to the even neighbour, so
it depends on the parity of s/u alone.  A run of terms is then, at a given
binade, a function of the incoming parity: an increment, an outgoing parity,
and the least and largest partial sums, which decide whether the run stays
in the binade for a given incoming sum.  The kernels (rz_lbfgs.hip) compute
these functions for runs of 64 terms at the binade the double prefix sums
predict, compose them into runs of 1024 and 8192 terms, and walk: a run
whose function applies at the current sum is one step, otherwise its parts
are tried, and a 64-term run that leaves the binade is added as floats.
This restates that walk in Python integers and requires numpy's sequential
float32 bits.  Planted mistakes (ties rounded up; a negative remainder
rounded as a positive one; a run of zeros treated as a function of its
first part only) must fail.
"""
import numpy as np
import pytest

M32 = 0xFFFFFFFF
TIE, UP, BIG = 1, 2, 4
SAT = 1 << 30
ZERO_RUN = -2
SUB, SEG_SUBS, SUP_SEGS = 64, 16, 8


def _bits(x):
    return int(np.float32(x).view(np.uint32))


def _key(s):
    b = _bits(s)
    E = (b >> 23) & 0xFF
    if (b & 0x7FFFFFFF) == 0 or E == 0xFF:
        return -1
    return ((b >> 31) << 8) | max(E, 1)


def _classify(d, key, plant):
    sgn, e = (key >> 8) << 31, (key & 0xFF) - 127
    b = _bits(d) ^ sgn
    E = (b >> 23) & 0xFF
    M = (b & 0x7FFFFF) | (0x800000 if E else 0)
    if M == 0:
        return 0, 0
    if E == 0xFF:
        return 0, BIG
    sh = (E if E else 1) - 127 - e
    neg = (b >> 31) != 0
    if sh >= 0:
        if sh >= 24 or (M << sh) >= (1 << 24):
            return 0, BIG
        return (-(M << sh) if neg else M << sh), 0
    k = min(-sh, 26)
    fa, rem, half = M >> k, M & ((1 << k) - 1), 1 << (k - 1)
    if not neg or plant == "negative_as_positive":
        return (-fa if neg else fa), (TIE if rem == half else 0) | (UP if rem > half else 0)
    return -fa - (1 if rem else 0), 0 if rem == 0 else TIE if rem == half else UP if rem < half else 0


def _inc(f, fl, p, plant):
    tie_up = (fl & TIE) and (plant == "ties_up" or ((p + f) & 1))
    return f + (1 if (fl & UP) or tie_up else 0)


def _identity(key):
    return {"key": key, "t": [0, 0], "q": [0, 1], "lo": [SAT, SAT], "hi": [-SAT, -SAT]}


def _clamp(x):
    return max(min(x, SAT), -SAT)


def _compose(F, G):
    k = G["key"] if F["key"] == ZERO_RUN else F["key"] if G["key"] == ZERO_RUN else \
        F["key"] if F["key"] == G["key"] else -1
    H = {"key": k, "t": [0, 0], "q": [0, 0], "lo": [0, 0], "hi": [0, 0]}
    for p in (0, 1):
        pm = F["q"][p]
        H["t"][p] = _clamp(F["t"][p] + G["t"][pm])
        H["lo"][p] = min(F["lo"][p], _clamp(F["t"][p] + G["lo"][pm]))
        H["hi"][p] = max(F["hi"][p], _clamp(F["t"][p] + G["hi"][pm]))
        H["q"][p] = G["q"][pm]
    return H


def _leaf(key, f, fl, plant):
    if fl & BIG:
        return _identity(-1)
    t = [_inc(f, fl, p, plant) for p in (0, 1)]
    return {"key": key, "t": t, "q": [(p + t[p]) & 1 for p in (0, 1)], "lo": [f, f], "hi": t}


def _functions(t, plant):
    """Run functions at the predicted binades: 64-term runs, then 1024, 8192."""
    dpre = np.concatenate([[0.0], np.cumsum(t.astype(np.float64).reshape(-1, SUB).sum(1))[:-1]])
    subs = []
    for g in range(t.size // SUB):
        run = t[g * SUB:(g + 1) * SUB]
        key = ZERO_RUN if not np.any(run) else _key(np.float32(dpre[g]))
        F = _identity(key)
        if key >= 0:
            for d in run:
                F = _compose(F, _leaf(key, *_classify(d, key, plant), plant))
                if F["key"] == -1:
                    break
        subs.append(F)

    def up(children, fan):
        out = []
        for g in range(len(children) // fan):
            F = children[g * fan]
            for c in children[g * fan + 1:(g + 1) * fan]:
                if F["key"] == -1 or (plant == "zero_run_first" and F["key"] == ZERO_RUN):
                    break
                F = _compose(F, c)
            out.append(F)
        return out
    segs = up(subs, SEG_SUBS)
    return subs, segs, up(segs, SUP_SEGS)


def ordered_sum(terms, plant=None):
    t = np.asarray(terms, np.float32)
    npad = max(-(-t.size // (SUB * SEG_SUBS * SUP_SEGS)), 1) * SUB * SEG_SUBS * SUP_SEGS
    t = np.concatenate([t, np.zeros(npad - t.size, np.float32)])
    subs, segs, sups = _functions(t, plant)
    state = {"s": np.float32(0)}

    def jump(F):
        if F["key"] == ZERO_RUN:
            return True
        s = state["s"]
        if F["key"] < 0 or _key(s) != F["key"]:
            return False
        b = _bits(s)
        E = (b >> 23) & 0xFF
        N = (b & 0x7FFFFF) | (0x800000 if E else 0)
        LO = 1 if E <= 1 else 1 << 23
        p = N & 1
        if N + F["lo"][p] < LO or N + F["hi"][p] >= 1 << 24:
            return False
        state["s"] = np.float32(np.copysign(np.ldexp(float(N + F["t"][p]), max(E, 1) - 150), s))
        return True

    with np.errstate(all="ignore"):
        for sp, F in enumerate(sups):
            if jump(F):
                continue
            for g in range(sp * SUP_SEGS, (sp + 1) * SUP_SEGS):
                if jump(segs[g]):
                    continue
                for u in range(g * SEG_SUBS, (g + 1) * SEG_SUBS):
                    if jump(subs[u]):
                        continue
                    s = state["s"]
                    for d in t[u * SUB:(u + 1) * SUB]:
                        s = np.float32(s + d)
                    state["s"] = s
    return state["s"]


def serial(terms):
    t = np.asarray(terms, np.float32)
    return np.add.accumulate(np.concatenate([np.zeros(1, np.float32), t]), dtype=np.float32)[-1]


def _cases():
    rng = np.random.default_rng(5)
    yield "signed", (rng.standard_normal(9000) * rng.standard_normal(9000)).astype(np.float32)
    yield "positive", np.abs(rng.standard_normal(20000)).astype(np.float32)
    e = rng.integers(-40, 40, 3000)
    yield "binades", (rng.standard_normal(3000) * 2.0 ** e).astype(np.float32)
    yield "subnormal", (rng.standard_normal(2000) * 1e-39).astype(np.float32)
    k = rng.integers(-8, 9, 9000).astype(np.float64)
    yield "ties", np.concatenate([[1.0], k * 2.0 ** -25, [-1.0], k[:300] * 2.0 ** -30]).astype(np.float32)
    big = rng.standard_normal(2000) * 1e6
    yield "cancel", np.concatenate([big, -big[::-1], rng.standard_normal(50)]).astype(np.float32)
    yield "zero_runs", np.concatenate([np.zeros(128), np.abs(rng.standard_normal(12000)) + 1.0,
                                       np.zeros(64), rng.standard_normal(500)]).astype(np.float32)


@pytest.mark.parametrize("name,terms", list(_cases()), ids=[c[0] for c in _cases()])
def test_method_gives_the_serial_float_sum(name, terms):
    assert _bits(ordered_sum(terms)) == _bits(serial(terms)), name


def test_runs_are_jumped():
    """On a sum that grows steadily most runs have a function at their
    predicted binade (the method is not the serial fallback in disguise)."""
    t = np.abs(np.random.default_rng(2).standard_normal(40000)).astype(np.float32) + 1.0
    subs, segs, sups = _functions(np.concatenate([t, np.zeros(-t.size % 8192, np.float32)]), None)
    assert sum(F["key"] >= 0 for F in subs) > 0.9 * (t.size // 64)
    assert sum(F["key"] >= 0 for F in segs) >= 20


@pytest.mark.parametrize("plant", ["ties_up", "negative_as_positive", "zero_run_first"])
def test_planted_mistakes_fail(plant):
    cases = dict(_cases())
    wrong = [n for n, t in cases.items() if _bits(ordered_sum(t, plant=plant)) != _bits(serial(t))]
    assert wrong, f"planted mistake {plant} went unnoticed"
