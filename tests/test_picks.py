"""rz_read_picks (calc_FA_res.c:28-37 with each file parsed once per change)
against the reference's own loop: one libc fscanf("%e") per receiver into
a 1-based Tobs, where a failed conversion leaves the entry as it was.  The
loop below is that loop, run through ctypes on the same files."""
import ctypes as C
import os
import subprocess
import sys
import time

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_libc = C.CDLL(None)
_libc.fopen.restype = C.c_void_p
_libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
_libc.fclose.argtypes = [C.c_void_p]


def _fscanf_loop(path, tobs):
    """for (i = 1; i <= ntr; i++) fscanf(fp, "%e", &Tobs[i]);"""
    fp = _libc.fopen(os.fsencode(path), b"r")
    assert fp
    v = C.c_float()
    for i in range(1, tobs.size):
        if _libc.fscanf(C.c_void_p(fp), b"%e", C.byref(v)) == 1:
            tobs[i] = v.value
    _libc.fclose(fp)
    return tobs


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def _bump_mtime(path, seconds):
    st = os.stat(path)
    os.utime(path, ns=(st.st_atime_ns, st.st_mtime_ns + int(seconds * 1e9)))


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture()
def N():
    from rajzel_amd import native
    native.picks_clear()
    yield native
    native.picks_clear()


def test_values_as_the_reference_loop_reads_them(N, tmp_path):
    rng = np.random.default_rng(5)
    vals = np.concatenate([rng.random(300).astype(np.float32) * 3, [-1.0, 0.0, 1e-30, 7.5e3]]).astype(np.float32)
    p = tmp_path / "picks_1.dat"
    _write(p, "".join(f"{v:e}\n" for v in vals))            # write_picks.c:26-31 format
    a = N.read_picks(p, np.full(vals.size + 1, 9.0, np.float32))
    b = _fscanf_loop(p, np.full(vals.size + 1, 9.0, np.float32))
    assert _same(a, b)
    assert _same(a[1:], np.array([np.float32(float(f"{v:e}")) for v in vals], np.float32))


def test_unchanged_file_is_not_parsed_again(N, tmp_path):
    p = tmp_path / "picks_2.dat"
    _write(p, "1.0e-01\n2.0e-01\n3.0e-01\n")
    n0 = N.picks_parses()
    a = N.read_picks(p, np.zeros(4, np.float32))
    assert N.picks_parses() == n0 + 1
    b = N.read_picks(p, np.zeros(4, np.float32))
    assert N.picks_parses() == n0 + 1
    assert _same(a, b) and _same(a, _fscanf_loop(p, np.zeros(4, np.float32)))


def test_changed_or_replaced_file_is_parsed_again(N, tmp_path):
    p = tmp_path / "picks_3.dat"
    _write(p, "1.0e-01\n2.0e-01\n")
    N.read_picks(p, np.zeros(3, np.float32))
    n0 = N.picks_parses()
    _write(p, "4.0e-01\n5.0e-01\n")                           # same size, new contents
    _bump_mtime(p, 1.0)
    a = N.read_picks(p, np.zeros(3, np.float32))
    assert N.picks_parses() == n0 + 1 and list(a[1:]) == [np.float32(0.4), np.float32(0.5)]
    q = tmp_path / "new.dat"                                   # replaced: another inode
    _write(q, "6.0e-01\n7.0e-01\n")
    os.replace(q, p)
    a = N.read_picks(p, np.zeros(3, np.float32))
    assert N.picks_parses() == n0 + 2 and list(a[1:]) == [np.float32(0.6), np.float32(0.7)]
    a = N.read_picks(p, np.zeros(4, np.float32))               # another receiver count
    assert N.picks_parses() == n0 + 3


@pytest.mark.parametrize("text", [
    "1.5\n2.5\n",                       # short file: the rest keep their values
    "1.5\nabc\n2.5\n",                  # a word stops every later conversion
    "1.5\n1e+ 2.5\n-0.5\n",             # glibc consumes "1e+" and fails; 2.5 then parses
    "",                                 # empty file
    "  \n 3.25e+00 \t -1.000000e+00\n",  # whitespace of any kind between values
])
def test_failed_conversions_keep_entries_cached_or_not(N, tmp_path, text):
    p = tmp_path / "picks_4.dat"
    _write(p, text)
    for sentinel in (9.0, -7.0):        # the second read comes from the kept parse
        a = N.read_picks(p, np.full(6, sentinel, np.float32))
        b = _fscanf_loop(p, np.full(6, sentinel, np.float32))
        assert _same(a, b), (text, a, b)


def test_missing_file(N, tmp_path):
    with pytest.raises(OSError):
        N.read_picks(tmp_path / "none.dat", np.zeros(3, np.float32))


def test_switch_off_parses_every_time(tmp_path):
    p = tmp_path / "picks_5.dat"
    _write(p, "1.0\n2.0\n")
    code = ("import numpy as np\nfrom rajzel_amd import native as N\n"
            f"for _ in range(3): N.read_picks({str(p)!r}, np.zeros(3, np.float32))\n"
            "print(N.picks_parses())\n")
    env = dict(os.environ, RAJZEL_AMD_PICK_CACHE="0", PYTHONPATH=ROOT)
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().splitlines()[-1] == "3"


def test_parse_cost_is_paid_once(N, tmp_path):
    """200 files of 1500 picks: the first pass parses, the second only
    stats (the reference parses on every evaluation)."""
    rng = np.random.default_rng(7)
    names = []
    for s in range(1, 201):
        p = tmp_path / f"p_{s}.dat"
        np.savetxt(p, rng.random(1500).astype(np.float32), fmt="%e")
        names.append(p)
    tobs = np.zeros(1501, np.float32)
    t0 = time.perf_counter()
    first = [N.read_picks(p, tobs).copy() for p in names]
    t1 = time.perf_counter()
    second = [N.read_picks(p, tobs).copy() for p in names]
    t2 = time.perf_counter()
    assert all(_same(a, b) for a, b in zip(first, second))
    assert (t2 - t1) < (t1 - t0)
