"""Seeded synthetic workloads for the FATT gradient path.

Every generator returns plain numpy data in the reference's conventions:

* velocity / slowness grids are float32, row-major ``(NY, NX)`` (x fastest),
  the block ``&Vp[1][1]`` of the reference's matrices (util.c:177-204);
* a cell with 0-based indices ``(k, h)`` sits at depth ``(k+1)*DH`` and offset
  ``(h+1)*DH`` (eikonal.c:31-32, receiver.c:55-56: index = iround(pos/DH));
* source and receiver positions are returned in metres *and* as 0-based grid
  indices computed with the reference's rounding ``(int)floor(pos/DH + 0.5)``
  in float, minus one.

Slowness is ``S = 1/Vp`` in float32 as calc_S.c:18-24.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def iround_index(pos_m, dh):
    """0-based grid index of a position, as the reference's iround(pos/DH) - 1.

    fd.h:18 defines iround(x) = (int)floor(x + 0.5); the division pos/DH is a
    float division (eikonal.c:31, receiver.c:55).
    """
    q = np.asarray(pos_m, dtype=np.float32) / np.float32(dh)
    return (np.floor(q.astype(np.float64) + 0.5).astype(np.int64) - 1).astype(np.int32)


def slowness(vp):
    """calc_S.c:18-24: S = 1.0/Vp, evaluated in double and stored as float."""
    return (1.0 / np.asarray(vp, dtype=np.float32).astype(np.float64)).astype(np.float32)


@dataclass
class Survey:
    """One model with its acquisition geometry and (optionally) picks."""

    name: str
    nx: int
    ny: int
    dh: float
    vp: np.ndarray                    # (ny, nx) float32, current model
    src_x_m: np.ndarray               # (nshots,) float32 metres
    src_z_m: np.ndarray
    rec_x_m: np.ndarray               # (ntr,) float32 metres
    rec_z_m: np.ndarray
    time: float = 6.0                 # TIME (T init = 2*TIME, lambda init = 10*TIME)
    ttnorm: float = 1e-10             # par/RAJZEL_marm_OBC.inp:48
    eps_adj: float = 1e-9             # par/RAJZEL_marm_OBC.inp:57
    vp_true: np.ndarray | None = None
    tobs: np.ndarray | None = None    # (nshots, ntr) float32, negative = muted
    noise_s: float = 0.0
    meta: dict = field(default_factory=dict)

    @property
    def S(self):
        return slowness(self.vp)

    @property
    def src_ix(self):
        return iround_index(self.src_x_m, self.dh)

    @property
    def src_iy(self):
        return iround_index(self.src_z_m, self.dh)

    @property
    def rec_ix(self):
        return iround_index(self.rec_x_m, self.dh)

    @property
    def rec_iy(self):
        return iround_index(self.rec_z_m, self.dh)

    @property
    def nshots(self):
        return int(self.src_x_m.shape[0])

    @property
    def ntr(self):
        return int(self.rec_x_m.shape[0])


def _grid_coords(nx, ny, dh):
    z = (np.arange(ny, dtype=np.float64) + 1.0) * dh
    x = (np.arange(nx, dtype=np.float64) + 1.0) * dh
    return np.meshgrid(x, z)  # X, Z each (ny, nx)


def _gaussians(rng, X, Z, n, amp_lo, amp_hi, r_lo, r_hi, dh, nx, ny):
    v = np.zeros_like(X)
    for _ in range(n):
        a = rng.uniform(amp_lo, amp_hi)
        r = rng.uniform(r_lo, r_hi)
        cx = rng.uniform(dh, nx * dh)
        cz = rng.uniform(dh, ny * dh)
        v += a * np.exp(-((X - cx) ** 2 + (Z - cz) ** 2) / (r * r))
    return v


def gradient_model(nx, ny, dh, v0=1500.0, grad=0.5):
    """1-D linear gradient v = v0 + grad*z (z = (k+1)*DH)."""
    X, Z = _grid_coords(nx, ny, dh)
    return (v0 + grad * Z).astype(np.float32)


def c1_model(seed=0, nx=600, ny=200, dh=10.0):
    """200 x 600, DH=10: v = 1500 + 0.5 z plus 5 seeded Gaussians
    (amplitude U(-300, 300) m/s, radius U(150, 250) m, centres uniform)."""
    rng = np.random.default_rng(seed)
    X, Z = _grid_coords(nx, ny, dh)
    v = 1500.0 + 0.5 * Z + _gaussians(rng, X, Z, 5, -300.0, 300.0, 150.0, 250.0, dh, nx, ny)
    return v.astype(np.float32)


def c1_forward(seed=0):
    """Plain forward, one shot at (3000 m, 10 m), 600 receivers at z = 10 m."""
    nx, ny, dh = 600, 200, 10.0
    vp = c1_model(seed, nx, ny, dh)
    return Survey(
        name="c1_forward_200x600", nx=nx, ny=ny, dh=dh, vp=vp,
        src_x_m=np.array([3000.0], np.float32), src_z_m=np.array([10.0], np.float32),
        rec_x_m=(10.0 * np.arange(1, 601)).astype(np.float32),
        rec_z_m=np.full(600, 10.0, np.float32),
    )


def c2_gradient(nshots=120, seed_model=0, seed_noise=1):
    """Full gradient, 200 x 600: truth = C1 model, start = 1500 + 0.5 z,
    shots every 50 m at z = 10 m, 600 receivers at z = 20 m (row 2, so that
    the adjoint source injects: init_travel_adj.c:44-55 seeds rows 2..NY-1
    only), picks = truth times + N(0, 2 ms).  Picks are filled in by
    :func:`synthesize_picks` (needs a forward solver)."""
    nx, ny, dh = 600, 200, 10.0
    return Survey(
        name="c2_gradient_200x600", nx=nx, ny=ny, dh=dh,
        vp=gradient_model(nx, ny, dh, 1500.0, 0.5),
        vp_true=c1_model(seed_model, nx, ny, dh),
        src_x_m=(50.0 * np.arange(1, nshots + 1)).astype(np.float32),
        src_z_m=np.full(nshots, 10.0, np.float32),
        rec_x_m=(10.0 * np.arange(1, 601)).astype(np.float32),
        rec_z_m=np.full(600, 20.0, np.float32),
        noise_s=0.002, meta={"seed_noise": seed_noise},
    )


def layered_model(nx, ny, dh, rng, nlayers=12, vmin=1500.0, vmax=5000.0, undulation=0.10):
    """nlayers constant-velocity layers, velocities linspace(vmin, vmax);
    interface i sits at mean depth i*Z/nlayers and undulates sinusoidally
    by +-undulation of that depth (seeded wavelength and phase)."""
    X, Z = _grid_coords(nx, ny, dh)
    zmax = ny * dh
    vel = np.linspace(vmin, vmax, nlayers)
    layer = np.zeros(X.shape, dtype=np.int64)
    for i in range(1, nlayers):
        lam = rng.uniform(0.25, 1.0) * nx * dh
        phi = rng.uniform(0.0, 2.0 * np.pi)
        zi = (i * zmax / nlayers) * (1.0 + undulation * np.sin(2.0 * np.pi * X / lam + phi))
        layer += (Z > zi).astype(np.int64)
    return vel[layer]


def c4_gradient(nshots=200, ntr=1500, seed=4, nx=3001, ny=1001, dh=5.0):
    """1001 x 3001, DH=5 m.  Truth = 12-layer model (1500-5000 m/s, +-10%
    sinusoidal interface undulation) plus 20 Gaussians (U(-300,300) m/s,
    radius U(50, 250) m); the current model is the layered model without the
    anomalies.  200 shots at x = 75 k - 37.5 m, z = 5 m; 1500 receivers at
    x = 10 k, z = 10 m (row 2).  Picks = truth times + N(0, 1 ms)."""
    rng = np.random.default_rng(seed)
    X, Z = _grid_coords(nx, ny, dh)
    base = layered_model(nx, ny, dh, rng)
    anom = _gaussians(rng, X, Z, 20, -300.0, 300.0, 50.0, 250.0, dh, nx, ny)
    spacing = 75.0 * (3001 / nx) if nx != 3001 else 75.0
    return Survey(
        name=f"c4_gradient_{ny}x{nx}", nx=nx, ny=ny, dh=dh,
        vp=base.astype(np.float32),
        vp_true=np.maximum(base + anom, 1000.0).astype(np.float32),
        src_x_m=(spacing * np.arange(1, nshots + 1) - spacing / 2).astype(np.float32),
        src_z_m=np.full(nshots, dh, np.float32),
        rec_x_m=(10.0 * np.arange(1, ntr + 1) * (nx - 1) / 3000.0).astype(np.float32),
        rec_z_m=np.full(ntr, 2 * dh, np.float32),
        noise_s=0.001, meta={"seed": seed},
    )


def c3_instances(nmodels=32, shots_per_model=8):
    """Uneven batch: per model m, default_rng(seed=m) draws a size from
    {200x200, 300x800, 400x1200, 600x1600} (NY x NX), DH=10, velocity
    1500 + 3000*minmax(gaussian_filter(U(0,1), sigma=8)); 8 shots at random
    surface columns (row 1); receivers at every interior column, row 2; picks
    are times in a 2%-faster copy plus N(0, 1 ms)."""
    from scipy.ndimage import gaussian_filter

    sizes = [(200, 200), (300, 800), (400, 1200), (600, 1600)]
    out = []
    for m in range(nmodels):
        rng = np.random.default_rng(m)
        ny, nx = sizes[int(rng.integers(0, len(sizes)))]
        dh = 10.0
        f = gaussian_filter(rng.uniform(0.0, 1.0, (ny, nx)), sigma=8.0)
        f = (f - f.min()) / (f.max() - f.min())
        vp = (1500.0 + 3000.0 * f).astype(np.float32)
        cols = rng.integers(2, nx, size=shots_per_model)  # 1-based 2..nx-1
        rec_cols = np.arange(2, nx)                           # interior, 1-based
        out.append(Survey(
            name=f"c3_m{m}_{ny}x{nx}", nx=nx, ny=ny, dh=dh, vp=vp,
            vp_true=(vp * np.float32(1.02)).astype(np.float32),
            src_x_m=(cols * dh).astype(np.float32), src_z_m=np.full(shots_per_model, dh, np.float32),
            rec_x_m=(rec_cols * dh).astype(np.float32), rec_z_m=np.full(rec_cols.size, 2 * dh, np.float32),
            noise_s=0.001, meta={"seed": m},
        ))
    return out


def small_case(nx=53, ny=37, dh=10.0, nshots=3, seed=7, ntr=None, rec_row=2):
    """Tiny seeded case for golden vectors and fast parity tests.

    ``rec_row`` is the 1-based receiver row (depth rec_row*DH), or a sequence
    of rows assigned to the receivers in turn (e.g. (2, ny-1) alternates a
    surface line with a line at depth, a transmission geometry in which the
    adjoint field fills the model)."""
    rng = np.random.default_rng(seed)
    X, Z = _grid_coords(nx, ny, dh)
    vtrue = 1500.0 + 0.5 * Z + _gaussians(rng, X, Z, 3, -300.0, 300.0, 3 * dh, 8 * dh, dh, nx, ny)
    cols = rng.integers(1, nx + 1, size=nshots)
    ntr = nx if ntr is None else ntr
    rc = np.linspace(1, nx, ntr).round().astype(np.int64)
    rows = np.resize(np.atleast_1d(np.asarray(rec_row, np.int64)), ntr)
    return Survey(
        name=f"small_{ny}x{nx}", nx=nx, ny=ny, dh=dh,
        vp=gradient_model(nx, ny, dh, 1500.0, 0.5), vp_true=vtrue.astype(np.float32),
        src_x_m=(cols * dh).astype(np.float32), src_z_m=np.full(nshots, dh, np.float32),
        rec_x_m=(rc * dh).astype(np.float32), rec_z_m=(rows * dh).astype(np.float32),
        noise_s=0.002, meta={"seed": seed},
    )


def edge_case(nx, ny, dh=10.0, nshots=3, seed=61, sigma=3.0, ttnorm=1e-4, rec_step=7):
    """Receiver edge cases on a rough model, for fixtures where the stopping
    test ends the solves before a fixed point.

    Velocity 1500 + 3000*minmax(gaussian_filter(U(0,1), sigma)) (rough for a
    small sigma, so that several outer iterations change cells), picks from a
    2%-faster copy.  Receivers, 0-based (row, column):
      * row 1 (the reference's row 2) every ``rec_step`` columns;
      * two receivers on one cell of row 1 and two on one cell at depth
        (init_recflag.c:24-30, init_travel_adj.c:49-55: the last one wins);
      * the top row 0 and the bottom row NY-1 (frame rows: no adjoint
        source, init_travel_adj.c:44; the bottom row's residual takes the
        backward slope, calc_FA_res.c:62-68);
      * the frame columns 0 and NX-1, and a line at depth NY/2.
    Sources at random surface columns (row 0) and one at depth."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.uniform(0.0, 1.0, (ny, nx)), sigma=sigma)
    f = (f - f.min()) / (f.max() - f.min())
    vp = (1500.0 + 3000.0 * f).astype(np.float32)
    cols = list(range(1, nx - 1, rec_step))
    rows = [1] * len(cols)
    dup_c, deep = cols[len(cols) // 2], ny // 2
    cells = [(1, dup_c), (deep, 2), (deep, 2), (1, dup_c),        # duplicates, interleaved
             (0, 3), (0, nx // 2), (ny - 1, 4), (ny - 1, nx - 3), # frame rows
             (ny // 3, 0), (2 * ny // 3, nx - 1)]                  # frame columns
    cells += [(deep, c) for c in range(3, nx - 1, max(1, 3 * rec_step))]
    rows += [k for k, _ in cells]
    cols += [c for _, c in cells]
    rk = np.asarray(rows, np.float64) + 1.0                     # 1-based
    rc = np.asarray(cols, np.float64) + 1.0
    scol = rng.integers(2, nx, size=nshots - 1)
    sx = np.concatenate([scol, [nx // 2]]).astype(np.float64)
    sz = np.concatenate([np.ones(nshots - 1), [ny // 2 + 1]])
    return Survey(
        name=f"edge_{ny}x{nx}", nx=nx, ny=ny, dh=dh, vp=vp,
        vp_true=(vp * np.float32(1.02)).astype(np.float32),
        src_x_m=(sx * dh).astype(np.float32), src_z_m=(sz * dh).astype(np.float32),
        rec_x_m=(rc * dh).astype(np.float32), rec_z_m=(rk * dh).astype(np.float32),
        ttnorm=ttnorm, noise_s=0.001, meta={"seed": seed},
    )


def taper_case(nx=60, ny=34, dh=10.0, nshots=4, seed=81):
    """A small survey for the per-shot gradient tapers (ass_grad.c:31-37):
    sources at depth 2*DH between the receivers, receivers on row 2 at both
    edge columns (1 and NX: a taper_grad_shot1.c window there runs off the
    grid sideways) and every 3 columns between, and a line at depth NY/2.
    Every (2*FILTSIZE+1)-row window of FILTSIZE <= 1 stays on the grid
    (the reference's behaviour for rows off the grid is undefined)."""
    rng = np.random.default_rng(seed)
    X, Z = _grid_coords(nx, ny, dh)
    vtrue = 1500.0 + 0.5 * Z + _gaussians(rng, X, Z, 3, -300.0, 300.0, 3 * dh, 8 * dh, dh, nx, ny)
    rc = np.concatenate([[1], np.arange(4, nx, 3), [nx], np.arange(3, nx, 7)]).astype(np.float64)
    rr = np.concatenate([np.full(len(rc) - len(np.arange(3, nx, 7)), 2), np.full(len(np.arange(3, nx, 7)), ny // 2)])
    sx = rng.uniform(3 * dh, (nx - 3) * dh, nshots)
    return Survey(
        name=f"taper_{ny}x{nx}", nx=nx, ny=ny, dh=dh,
        vp=gradient_model(nx, ny, dh, 1500.0, 0.5), vp_true=vtrue.astype(np.float32),
        src_x_m=sx.astype(np.float32), src_z_m=np.full(nshots, 2 * dh, np.float32),
        rec_x_m=(rc * dh).astype(np.float32), rec_z_m=(rr * dh).astype(np.float32),
        noise_s=0.002, meta={"seed": seed},
    )


def c5_gradient(nshots=400, ntr=3000, seed=5, nx=6000, ny=2000, dh=2.5):
    """2000 x 6000, DH = 2.5 m.  Truth = 1-D gradient 1500 + 0.6 z plus 40
    seeded Gaussians (U(-400, 400) m/s, radius U(50, 300) m); the current
    model is the 1-D gradient.  400 shots every 37.5 m (x = 37.5 k - 18.75
    m) at z = 2.5 m (row 1); 3000 receivers every 5 m at z = 5 m (row 2).
    Picks = truth times + N(0, 1 ms)."""
    rng = np.random.default_rng(seed)
    X, Z = _grid_coords(nx, ny, dh)
    base = 1500.0 + 0.6 * Z
    anom = _gaussians(rng, X, Z, 40, -400.0, 400.0, 50.0, 300.0, dh, nx, ny)
    return Survey(
        name=f"c5_gradient_{ny}x{nx}", nx=nx, ny=ny, dh=dh,
        vp=base.astype(np.float32),
        vp_true=np.maximum(base + anom, 1000.0).astype(np.float32),
        src_x_m=(37.5 * np.arange(1, nshots + 1) - 18.75).astype(np.float32),
        src_z_m=np.full(nshots, dh, np.float32),
        rec_x_m=(5.0 * np.arange(1, ntr + 1)).astype(np.float32),
        rec_z_m=np.full(ntr, 2 * dh, np.float32),
        noise_s=0.001, meta={"seed": seed},
    )


def synthesize_picks(survey: Survey, forward_fn, seed=None):
    """Fill ``survey.tobs`` with forward times in ``vp_true`` plus seeded
    Gaussian noise.  ``forward_fn(S, survey) -> (nshots, ntr) float32``."""
    vt = survey.vp_true if survey.vp_true is not None else survey.vp
    t = np.asarray(forward_fn(slowness(vt), survey), dtype=np.float32)
    if seed is None:
        seed = survey.meta.get("seed_noise", survey.meta.get("seed", 0) + 1000)
    rng = np.random.default_rng(seed)
    noise = rng.normal(0.0, survey.noise_s, size=t.shape) if survey.noise_s > 0 else 0.0
    survey.tobs = (t + noise).astype(np.float32)
    return survey.tobs
