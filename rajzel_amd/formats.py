"""The reference's on-disk formats, read and written bit-compatibly.

* parameter file   positional slots (read_par.c:72-268)
* workflow file    header + one 6-column row per stage (read_par_inv.c:24-27)
* source file      nsrc, then x z y tshift fc amp [az type] per line (sources.c:37-71)
* receiver file    "x y" per line, count = number of newlines (receiver.c:34-58)
* model files      raw native float32, x slow / y fast (readmod.c:34-39, writemod.c:26-28)
* pick files       one "%e \\n" per receiver, <stem>_<shot>.dat, shot 1-based
                   (write_picks.c:26-31, read back with fscanf("%e") in calc_FA_res.c:28-37)

Numbers are converted with the C library's own strtol/strtof/printf so that
a value written by the reference and read here (or the reverse) is the same
float bit for bit.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import os

import numpy as np

_libc = ctypes.CDLL(ctypes.util.find_library("c") or None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]
_libc.strtol.restype = ctypes.c_long
_libc.strtol.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int]
_libc.snprintf.restype = ctypes.c_int

# ---------------------------------------------------------------------------
# scalar conversions with C semantics
# ---------------------------------------------------------------------------


def _c_scan(fn, text: bytes, pos: int, *extra):
    """Run strtof/strtol on text[pos:]; return (value, new_pos) or (None, pos)."""
    buf = ctypes.create_string_buffer(text[pos:])
    end = ctypes.c_char_p()
    v = fn(buf, ctypes.byref(end), *extra)
    used = ctypes.cast(end, ctypes.c_void_p).value - ctypes.addressof(buf)
    if used == 0:
        return None, pos
    return v, pos + used


def c_float(s: str) -> np.float32:
    """strtof (what fscanf %f / %e store)."""
    v, p = _c_scan(_libc.strtof, s.strip().encode(), 0)
    if v is None:
        raise ValueError(f"not a float: {s!r}")
    return np.float32(v)


def format_e(x) -> str:
    """printf("%e", (double)x) -- the reference writes floats promoted to double."""
    buf = ctypes.create_string_buffer(64)
    _libc.snprintf(buf, 64, b"%e", ctypes.c_double(float(x)))
    return buf.value.decode()


# ---------------------------------------------------------------------------
# picks
# ---------------------------------------------------------------------------


def pick_path(stem: str, shot1: int) -> str:
    """<stem>_<shot>.dat with the 1-based shot number (write_picks.c:26)."""
    return f"{stem}_{shot1}.dat"


def write_picks(path: str, times) -> None:
    """One "%e \\n" line per receiver (write_picks.c:30-32)."""
    with open(path, "w") as f:
        for t in np.asarray(times, np.float32).ravel():
            f.write(format_e(t) + " \n")


def read_picks(path: str, ntr: int) -> np.ndarray:
    """ntr values as fscanf("%e") stores them (calc_FA_res.c:36-37)."""
    with open(path, "rb") as f:
        text = f.read()
    out = np.empty(ntr, np.float32)
    pos = 0
    for i in range(ntr):
        v, pos2 = _c_scan(_libc.strtof, text, pos)
        if v is None:
            raise ValueError(f"{path}: only {i} of {ntr} picks")
        out[i] = v
        pos = pos2
    return out


def roundtrip_picks(times) -> np.ndarray:
    """Values as they come back from a pick file: printf %e then strtof."""
    t = np.asarray(times, np.float32)
    flat = np.array([c_float(format_e(x)) for x in t.ravel()], np.float32)
    return flat.reshape(t.shape)


# ---------------------------------------------------------------------------
# models
# ---------------------------------------------------------------------------


def read_model(path: str, nx: int, ny: int) -> np.ndarray:
    """Raw float32, x slow / y fast (readmod.c:34-39) -> row-major (ny, nx)."""
    a = np.fromfile(path, dtype=np.float32, count=nx * ny)
    if a.size != nx * ny:
        raise ValueError(f"{path}: {a.size} floats, expected {nx * ny}")
    return np.ascontiguousarray(a.reshape(nx, ny).T)


def write_model(path: str, field) -> None:
    """Binary format 3 of writemod.c:26-28 (IDX = IDY = 1)."""
    a = np.asarray(field, np.float32)
    np.ascontiguousarray(a.T).tofile(path)


# ---------------------------------------------------------------------------
# acquisition
# ---------------------------------------------------------------------------


def read_sources(path: str):
    """sources.c:37-71: first token = nsrc, first line skipped, then per line
    ``x z y tshift fc amp [az type]``.  Returns dict of float32 arrays; the
    reference keeps column 3 (y) as the depth srcpos[2]."""
    with open(path) as f:
        lines = f.read().split("\n")
    n = int(lines[0].split()[0])
    cols = []
    for l in range(1, n + 1):
        toks = lines[l].split()
        if len(toks) < 6:
            raise ValueError(f"{path}: source {l} has {len(toks)} values, need at least 6")
        cols.append([c_float(t) for t in toks[:8]] + [np.float32(0)] * (8 - min(len(toks), 8)))
    a = np.array(cols, np.float32).reshape(n, 8)
    return {"x": a[:, 0], "depth": a[:, 2], "tshift": a[:, 3], "fc": a[:, 4], "amp": a[:, 5],
            "az": a[:, 6], "type": a[:, 7]}


def read_receivers(path: str, refrec=(0.0, 0.0)):
    """receiver.c:34-58: count = newlines, ``x y`` per line; positions are
    shifted by REFREC.  Returns (x, y) float32 metres."""
    with open(path) as f:
        text = f.read()
    n = text.count("\n")
    vals = text.split()
    if len(vals) < 2 * n:
        raise ValueError(f"{path}: {len(vals)} values for {n} receivers")
    xy = np.array([c_float(v) for v in vals[:2 * n]], np.float32).reshape(n, 2)
    x = (xy[:, 0] + np.float32(refrec[0])).astype(np.float32)
    y = (xy[:, 1] + np.float32(refrec[1])).astype(np.float32)
    return x, y


def receiver_indices(x, y, dh):
    """0-based cells as receiver.c:55-58 (iround(pos/DH)), dropping receivers
    whose x index is <= 0 as receiver.c:77 does."""
    from .workloads import iround_index
    ix = iround_index(x, dh)
    iy = iround_index(y, dh)
    keep = ix + 1 > 0
    return ix[keep], iy[keep]


# ---------------------------------------------------------------------------
# parameter and workflow files
# ---------------------------------------------------------------------------

# (slot number, name(s), kind) of read_par.c:77-265; kind: i int, f float,
# s string, "f,f" etc. = comma-separated list after '='.
PAR_SLOTS = [
    (1, "INVMAT", "i"), (2, "NX", "i"), (3, "NY", "i"), (4, "DH", "f"), (5, "TIME", "f"),
    (6, "SOURCE_FILE", "s"), (7, "READMOD", "i"), (8, "MFILE", "s"), (9, "SNAP", "i"),
    (10, "IDX", "i"), (11, "IDY", "i"), (12, "SNAP_FORMAT", "i"), (13, "SNAP_FILE", "s"),
    (14, "SEISMO", "i"), (15, "REC_FILE", "s"), (16, ("REFREC1", "REFREC2"), "f,f"),
    (17, "N_STREAMER", "i"), (18, "REC_INCR_X", "f"), (19, "REC_INCR_Y", "f"),
    (20, "PICKS_FILE", "s"), (21, "TTNORM", "f"), (22, "LOG_FILE", "s"), (23, "LOG", "i"),
    (24, "ITERMAX", "i"), (25, "EPS_ADJ", "f"), (26, "JACOBIAN", "s"), (27, "DATA_DIR", "s"),
    (28, "TAPER", "i"), (29, "TAPERLENGTH", "i"), (30, "INVTYPE", "i"),
    (31, ("GRADT1", "GRADT2", "GRADT3", "GRADT4"), "i,i,i,i"), (32, "SWS_TAPER_GRAD_VERT", "i"),
    (33, "SWS_TAPER_GRAD_HOR", "i"), (34, "EXP_TAPER_GRAD_HOR", "f"), (35, "SWS_TAPER_GRAD_SOURCES", "i"),
    (36, "SWS_TAPER_CIRCULAR_PER_SHOT", "i"), (37, "SRTSHAPE", "i"), (38, "SRTRADIUS", "f"),
    (39, "FILTSIZE", "i"), (40, "SWS_TAPER_FILE", "i"), (41, "INV_MODELFILE", "s"), (42, "nfstart", "i"),
    (43, "nf", "i"), (44, "nfstart_jac", "i"), (45, "nf_jac", "i"), (46, "VPUPPERLIM", "f"),
    (47, "VPLOWERLIM", "f"), (48, "GRAD_METHOD", "i"), (49, "NLBFGS", "i"), (50, "MODEL_FILTER", "i"),
    (51, "FILT_SIZE", "i"), (52, "LINESEARCH", "i"), (53, "EPS_SCALE", "f"), (54, ("C1", "C2"), "f,f"),
    (55, "STEPMAX", "i"), (56, "SCALEFAC", "f"), (57, "MISFIT_LOG_FILE", "s"), (58, "MIN_ITER", "i"),
    (59, "GRAD_FILTER", "i"), (60, "FILT_SIZE_GRAD", "i"), (61, ("VP0_1", "VP0_2", "DVP0"), "f,f,f"),
    (62, ("GRAD0_1", "GRAD0_2", "DGRAD0"), "f,f,f"), (63, "GRIDSEARCH_FILE", "s"),
]


class _Scan:
    """Minimal fscanf/getc emulation over a byte string."""

    def __init__(self, data: bytes):
        self.d, self.p = data, 0

    def getc(self):
        if self.p >= len(self.d):
            return None
        c = self.d[self.p:self.p + 1]
        self.p += 1
        return c

    def _ws(self):
        while self.p < len(self.d) and self.d[self.p:self.p + 1].isspace():
            self.p += 1

    def token(self):              # %s
        self._ws()
        s = self.p
        while self.p < len(self.d) and not self.d[self.p:self.p + 1].isspace():
            self.p += 1
        return self.d[s:self.p].decode() if self.p > s else None

    def literal(self, ch: bytes, skip_ws=True):  # " =" / " ," ; "," alone: no whitespace before
        if skip_ws:
            self._ws()
        if self.d[self.p:self.p + 1] == ch:
            self.p += 1
            return True
        return False

    def integer(self):            # %i: strtol base 0
        self._ws()
        v, p = _c_scan(_libc.strtol, self.d, self.p, 0)
        if v is None:
            return None
        self.p = p
        return int(v)

    def real(self):               # %f
        self._ws()
        v, p = _c_scan(_libc.strtof, self.d, self.p)
        if v is None:
            return None
        self.p = p
        return np.float32(v)


def read_par(path: str) -> dict:
    """Parse a RAJZEL parameter file the way read_par.c:72-268 does: after
    every newline one character is consumed; unless it is '#', the next slot
    number is read as ``<key> = <value>`` (the key text itself is ignored,
    only its position counts)."""
    with open(path, "rb") as f:
        sc = _Scan(f.read())
    slots = {n: (name, kind) for n, name, kind in PAR_SLOTS}
    out, lineno = {}, 0
    while True:
        c = sc.getc()
        if c is None:
            break
        if c == b"\n":
            nxt = sc.getc()
            if nxt is None:
                break
            if nxt == b"#":
                continue
            lineno += 1
            if lineno not in slots:
                continue
            name, kind = slots[lineno]
            if sc.token() is None or not sc.literal(b"="):
                continue
            names = name if isinstance(name, tuple) else (name,)
            for j, (nm, k) in enumerate(zip(names, kind.split(","))):
                # "%f ,%f" for REFREC (read_par.c:123); "%i, %i" / "%f, %f"
                # elsewhere, where the comma must follow the number directly
                if j > 0 and not sc.literal(b",", skip_ws=(lineno == 16)):
                    break
                v = sc.integer() if k == "i" else sc.real() if k == "f" else sc.token()
                if v is None:
                    break
                out[nm] = v
    return out


def _par_value(v, kind):
    if kind == "i":
        return str(int(v))
    if kind == "f":
        return repr(float(np.float32(v)))
    return str(v)


def write_par(path: str, values: dict) -> None:
    """Write a parameter file that read_par.c:72-268 (and read_par above)
    reads back as ``values``: a comment line, then one ``KEY = value`` line
    per slot in slot order.  Every slot name of PAR_SLOTS must be given;
    lists take one value per name."""
    lines = ["# RAJZEL parameter file"]
    for _, name, kind in PAR_SLOTS:
        names = name if isinstance(name, tuple) else (name,)
        vals = [_par_value(values[nm], k) for nm, k in zip(names, kind.split(","))]
        key = "_".join(names)
        lines.append(f"{key} = {', '.join(vals)}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_workflow(path: str, rows) -> None:
    """Workflow file of read_par_inv.c:24-27: a header line, then per stage
    PRO SPATFILTER WD_DAMP WD_DAMP1 FILT_SIZE_GRAD FILT_SIZE_GRAD1."""
    out = ["PRO\tSPATFILTER\tWD_DAMP\tWD_DAMP1\tFILT_SIZE_GRAD\tFILT_SIZE_GRAD1"]
    for r in rows:
        out.append(f"{r['PRO']}\t{int(r['SPATFILTER'])}\t{r['WD_DAMP']}\t{r['WD_DAMP1']}\t"
                   f"{int(r['FILT_SIZE_GRAD'])}\t{int(r['FILT_SIZE_GRAD1'])}")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_sources(path: str, x, depth) -> None:
    """Source file of sources.c:37-71: the count, then ``x z y tshift fc amp``
    per line with the depth in the third column (the one kept as srcpos[2])."""
    x = np.asarray(x, np.float32)
    depth = np.asarray(depth, np.float32)
    with open(path, "w") as f:
        f.write(f"{len(x)}\n")
        for a, d in zip(x, depth):
            f.write(f"{repr(float(a))} 0.0 {repr(float(d))} 0.0 10.0 1.0\n")


def write_receivers(path: str, x, y) -> None:
    """Receiver file of receiver.c:34-58: ``x y`` per line, one line each."""
    with open(path, "w") as f:
        for a, b in zip(np.asarray(x, np.float32), np.asarray(y, np.float32)):
            f.write(f"{repr(float(a))} {repr(float(b))}\n")


def read_workflow(path: str):
    """Rows of PRO SPATFILTER WD_DAMP WD_DAMP1 FILT_SIZE_GRAD FILT_SIZE_GRAD1
    (read_par_inv.c:24-27).  The stage count of fatt.c:54-59 is the number of
    newlines minus one."""
    with open(path) as f:
        text = f.read()
    nstage = text.count("\n") - 1
    toks = text.split()[6:]
    rows = []
    for i in range(nstage):
        t = toks[6 * i:6 * i + 6]
        rows.append({"PRO": c_float(t[0]), "SPATFILTER": int(t[1]), "WD_DAMP": c_float(t[2]),
                     "WD_DAMP1": c_float(t[3]), "FILT_SIZE_GRAD": int(t[4]), "FILT_SIZE_GRAD1": int(t[5])})
    return rows


def survey_from_par(par_path: str, model_path: str | None = None):
    """Build a workloads.Survey from a parameter file and the files it names
    (paths relative to the parameter file's directory, as the reference runs
    from there)."""
    from .workloads import Survey
    base = os.path.dirname(os.path.abspath(par_path))
    P = read_par(par_path)
    nx, ny, dh = P["NX"], P["NY"], float(P["DH"])
    src = read_sources(os.path.join(base, P["SOURCE_FILE"]))
    rx, ry = read_receivers(os.path.join(base, P["REC_FILE"]), (P.get("REFREC1", 0.0), P.get("REFREC2", 0.0)))
    if model_path is None:
        model_path = os.path.join(base, P["MFILE"] + ".vp")
    vp = read_model(model_path, nx, ny) if os.path.exists(model_path) else None
    return Survey(name=os.path.basename(par_path), nx=nx, ny=ny, dh=dh, vp=vp, src_x_m=src["x"],
                  src_z_m=src["depth"], rec_x_m=rx, rec_z_m=ry, time=float(P["TIME"]),
                  ttnorm=float(P["TTNORM"]), eps_adj=float(P["EPS_ADJ"]), meta={"par": P})
