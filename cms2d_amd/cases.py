"""Seeded synthetic inputs for the CMS-Flow implicit step workloads.

Every floating-point input is rounded to float32 before use: the reference reads
grid depths, Manning fields and cell sizes through single-precision buffers
(source/input/in_lib.f90:86, source/geospatial/geo_def.f90:118), so an input
that is not float32-representable would differ between the reference and any
FP64 build of it.

Workloads (all on Cartesian grids, cells numbered row-major from the south-west
corner, as source/geospatial/geospatial.F90:2709-2828 numbers them):

* ``basin``   -- 200 x 100 cells of 10 m, linear offshore slope 1 m -> 10 m
  towards the east plus seeded Gaussian bumps (sigma 0.2 m), M2 harmonic water
  level on the east side (0.5 m, 12.42 h), walls elsewhere, Manning 0.025,
  dt = 60 s, flow only.
* ``tidal``   -- 2000 x 1500 cells of 20 m (3.0 M cells), beach at the west
  (bed from +1 m to -15 m), wetting and drying with hdry = 0.05 m, lognormal
  Manning (median 0.025, sigma_log 0.2), smooth seeded radiation-stress field
  up to 0.05 N/m^2, M2 tide on the east side, dt = 30 s.
* ``square``  -- 1000 x 1000 cells of 20 m, the same physics as ``tidal``
  without the radiation stress.
* ``tidal_tel`` -- the ``tidal`` physics on a telescoping grid: base cells of
  20 m with a two-level band (10 m around a 5 m core) along the shore, cells
  numbered base cell by base cell (read_tel, geospatial.F90:403-795).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

RHOW = 1025.0
M2_SPEED_DEG_HR = 360.0 / 12.42        # harmonic speed of a 12.42 h period [deg/hr]


def f32(a):
    """Round to float32 and return float64."""
    return np.asarray(a, dtype=np.float32).astype(np.float64)


@dataclasses.dataclass
class Case:
    name: str
    nx: int
    ny: int
    dx: float
    dy: float
    zb: np.ndarray            # bed elevation [ny*nx] (row-major, j*nx+i), m (negative = below datum)
    mann: np.ndarray          # Manning n [ny*nx]
    wavestrx: np.ndarray | None
    wavestry: np.ndarray | None
    th_side: int              # 0=N,1=E,2=S,3=W, -1 none
    amp: list
    speed_deg_hr: list
    phase_deg: list
    dt: float
    nsteps: int
    rampdur: float = 0.0      # hours
    hdry: float = 0.05
    noptset: int = 2          # >=3: wave branches (radiation-stress forcing) active
    wallfric: int = 1
    ndsch: int = 4            # exponential scheme
    latitude: float = 0.0     # AVERAGE_LATITUDE, degrees (Coriolis)
    coriolis_plane: int = 0   # CORIOLIS_APPROXIMATION: 0 f-plane, 1 beta-plane
    mturbul: int = 5          # TURBULENCE_MODEL: 0 constant, 1 Falconer, 2 parabolic, 3 subgrid-Wu, 5 subgrid
    mbedfric: int = 2         # bottom friction law of `mann`: 2 Manning n, 1 drag coefficient,
                              # 3 roughness length z0 [m] (height / 30), 4 linear coefficient [1/s]
    wse_times: object = None  # water-level curve on the open side instead of the harmonics:
    wse_values: object = None # times [hr] and levels [m] (WSE_CURVE), Lagrange order wse_order
    wse_order: int = 0        # (0: the reference's default 2)
    wse_adjust: bool = True   # WSE_ADJUSTMENT of the curve string (wind and wave set-up)
    wse_grad: tuple = (0.0, 0.0)   # WSE_GRADIENTS of the curve string [m/m]
    wind_speed: float = 0.0   # constant 10-m wind speed [m/s] (0: no wind)
    wind_dir: float = 0.0     # the direction it comes from, degrees clockwise from north

    tel: object = None        # TelGrid: the cells of a telescoping grid (nx, ny, dx, dy then
                              # describe its base grid; zb, mann, wavestr are per cell)
    noskew: int = 0           # SKEWNESS_CORRECTION OFF (telescoping grids)
    dtvar: int = 0            # VARIABLE_TIMESTEP (the reference's card; the flow objects take it
                              # as an argument)
    q_side: int = -1          # side of a flux (river) string, 0=N,1=E,2=S,3=W, -1 none:
    q_const: float = 0.0      # its total inflow [m^3/s] (FLUX_VALUE ... 'm^3/s'), or
    q_times: object = None    # a curve: times [hr] and inflows [m^3/s] (FLUX_CURVE),
    q_values: object = None   # Lagrange order q_order (0: the reference's default 1)
    q_order: int = 0
    q_cmvel: float = 0.0      # CONVEYANCE coefficient (0: the reference's default 0.667)
    q_angle: object = None    # INFLOW_ANGLE [deg clockwise from north] (None: the inward normal)
    q_rating: bool = False    # q_times/q_values are a RATING_CURVE: mean depth [m] -> inflow

    def wind_vector(self):
        """(wndx, wndy) in grid coordinates as windcurve_init makes them
        (meteorological.F90:700-710, grid angle 0, 10-m height, factor 1):
        direction + 180 degrees, x = W sin, y = W cos."""
        d = np.deg2rad(self.wind_dir) + np.pi
        return self.wind_speed * np.sin(d), self.wind_speed * np.cos(d)

    @property
    def ncells(self) -> int:
        return self.nx * self.ny if self.tel is None else len(self.tel.geom)

    @property
    def speed_rad_hr(self):
        return [s * math.pi / 180.0 for s in self.speed_deg_hr]

    @property
    def phase_rad(self):
        return [p * math.pi / 180.0 for p in self.phase_deg]


def _bumps(rng, nx, ny, nbumps, sigma_amp, rmin, rmax):
    jj, ii = np.mgrid[0:ny, 0:nx]
    z = np.zeros((ny, nx))
    for _ in range(nbumps):
        ci, cj = rng.uniform(0, nx), rng.uniform(0, ny)
        r = rng.uniform(rmin, rmax)
        a = rng.normal(0.0, sigma_amp)
        z += a * np.exp(-((ii - ci) ** 2 + (jj - cj) ** 2) / (2 * r * r))
    return z


def _smooth_field(rng, nx, ny, nmodes, peak):
    jj, ii = np.mgrid[0:ny, 0:nx]
    x, y = ii / nx, jj / ny
    f = np.zeros((ny, nx))
    for _ in range(nmodes):
        kx, ky = rng.integers(1, 4), rng.integers(1, 4)
        ph1, ph2 = rng.uniform(0, 2 * np.pi, 2)
        f += rng.normal() * np.sin(np.pi * kx * x + ph1) * np.sin(np.pi * ky * y + ph2)
    return f * (peak / max(np.abs(f).max(), 1e-30))


def basin(nx=200, ny=100, dx=10.0, seed=0, nsteps=1440, dt=60.0) -> Case:
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    depth = 1.0 + 9.0 * x[None, :] * np.ones((ny, 1))
    depth = depth + _bumps(rng, nx, ny, 12, 0.2, max(2, nx / 40), max(3, nx / 12))
    zb = f32(-depth).ravel()
    mann = f32(np.full(nx * ny, 0.025))
    return Case("basin", nx, ny, dx, dx, zb, mann, None, None, 1, [0.5], [M2_SPEED_DEG_HR], [90.0],
                dt, nsteps)


def tidal(nx=2000, ny=1500, dx=20.0, seed=1, nsteps=2880, dt=30.0, radstress=True) -> Case:
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    depth = -1.0 + 16.0 * x[None, :] * np.ones((ny, 1))
    depth = depth + _bumps(rng, nx, ny, 40, 0.2, max(2, nx / 80), max(3, nx / 20))
    zb = f32(-depth).ravel()
    mann = f32(0.025 * np.exp(0.2 * rng.standard_normal(nx * ny)))
    wx = wy = None
    noptset = 2
    if radstress:
        tx = _smooth_field(rng, nx, ny, 6, 0.05)
        ty = _smooth_field(rng, nx, ny, 6, 0.05)
        mag = np.sqrt(tx * tx + ty * ty)
        scale = np.where(mag > 0.05, 0.05 / np.maximum(mag, 1e-30), 1.0)
        wx = f32((tx * scale / RHOW)).ravel()
        wy = f32((ty * scale / RHOW)).ravel()
        noptset = 3
    return Case("tidal", nx, ny, dx, dx, zb, mann, wx, wy, 1, [0.5], [M2_SPEED_DEG_HR], [90.0],
                dt, nsteps, noptset=noptset)


def river(nx=48, ny=32, dx=20.0, seed=4, nsteps=40, dt=30.0, q_side=3, q=40.0, q_curve=None, q_order=0,
          q_cmvel=0.0, shoal=False, q_angle=None, q_rating=None) -> Case:
    """An estuary: the M2 tide on the east side and a river inflow (flux
    string) on side q_side (west by default), depth 2 m at the west edge to
    12 m at the east with bumps; q_curve = (hours, m^3/s) replaces the
    constant q.  shoal=True lifts the west edge so that its southern half is
    dry land: the river then enters through the wet part of its string only."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    depth = 2.0 + 10.0 * x[None, :] * np.ones((ny, 1))
    if shoal:
        depth = depth - 2.4 * np.exp(-x[None, :] * 8.0) * (1.0 - y[:, None])
    depth = depth + _bumps(rng, nx, ny, 20, 0.15, max(2, nx / 40), max(3, nx / 12))
    zb = f32(-depth).ravel()
    mann = f32(0.025 * np.exp(0.2 * rng.standard_normal(nx * ny)))
    c = Case("river", nx, ny, dx, dx, zb, mann, None, None, 1, [0.5], [M2_SPEED_DEG_HR], [90.0], dt, nsteps)
    c.q_side, c.q_const, c.q_order, c.q_cmvel, c.q_angle = q_side, q, q_order, q_cmvel, q_angle
    if q_curve is not None:
        c.q_times, c.q_values = list(q_curve[0]), list(q_curve[1])
    if q_rating is not None:                 # (stages [m], inflows [m^3/s])
        c.q_times, c.q_values, c.q_rating = list(q_rating[0]), list(q_rating[1]), True
    return c


def river_tel(nx=40, ny=24, dx=20.0, seed=4, q=40.0, **kw) -> Case:
    """The estuary of ``river`` on a telescoping grid: a band of 10 m cells with
    a 5 m core near the river mouth (base columns 2-10, 4-8), the tide's
    string on the east edge, the flux string on the west edge (its cells
    south to north); kw are river's options of the inflow (q_curve, q_rating,
    q_angle, q_cmvel)."""
    levels = band_levels(nx, ny, (2, 10), (4, 8))
    tel = telescope(levels, dx, side=1)
    rng = np.random.default_rng(seed)
    x, y = tel.geom[:, 0], tel.geom[:, 1]
    xn = x / (nx * dx)
    depth = 2.0 + 10.0 * xn
    xi, yj = x / dx, y / dx
    for _ in range(20):
        ci, cj = rng.uniform(0, nx), rng.uniform(0, ny)
        r = rng.uniform(max(2, nx / 40), max(3, nx / 12))
        depth = depth + rng.normal(0.0, 0.15) * np.exp(-((xi - ci) ** 2 + (yj - cj) ** 2) / (2 * r * r))
    zb = f32(-depth)
    mann = f32(0.025 * np.exp(0.2 * rng.standard_normal(len(x))))
    c = Case("river_tel", nx, ny, dx, dx, zb, mann, None, None, 1, [0.5], [M2_SPEED_DEG_HR], [90.0], 30.0, 40)
    c.tel = tel
    west = np.nonzero(np.isclose(x - tel.geom[:, 2] / 2, 0.0))[0]
    tel.flux_cells = west[np.argsort(y[west])]
    c.q_side, c.q_const = 3, q
    for k, v in kw.items():
        if k == "q_curve":
            c.q_times, c.q_values = list(v[0]), list(v[1])
        elif k == "q_rating":
            c.q_times, c.q_values, c.q_rating = list(v[0]), list(v[1]), True
        else:
            setattr(c, k, v)
    return c


def square(nx=1000, ny=1000, dx=20.0, seed=3, nsteps=720, dt=60.0) -> Case:
    c = tidal(nx, ny, dx, seed=seed, nsteps=nsteps, dt=dt, radstress=False)
    c.name = "square"
    return c


def ensemble_members(n=256, seed=11, side_min=64, side_max=512, dx=20.0, dt=60.0, nsteps=720):
    """Independent basins of the batched ensemble: nx and ny uniform in
    [side_min, side_max], offshore depth slope, tide amplitude uniform in
    0.2-1.0 m and lognormal Manning with median uniform in 0.015-0.035, each
    member from its own seed."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n):
        nx, ny = (int(v) for v in rng.integers(side_min, side_max + 1, size=2))
        amp = float(rng.uniform(0.2, 1.0))
        med = float(rng.uniform(0.015, 0.035))
        deep = float(rng.uniform(5.0, 15.0))
        r = np.random.default_rng([seed, m])
        x = (np.arange(nx) + 0.5) / nx
        depth = 1.0 + (deep - 1.0) * x[None, :] * np.ones((ny, 1))
        depth = depth + _bumps(r, nx, ny, 8, 0.2, max(2, nx / 40), max(3, nx / 12))
        zb = f32(-depth).ravel()
        mann = f32(med * np.exp(0.2 * r.standard_normal(nx * ny)))
        out.append(Case(f"member{m}", nx, ny, dx, dx, zb, mann, None, None, 1, [amp], [M2_SPEED_DEG_HR],
                        [90.0], dt, nsteps))
    return out


def lake_at_rest(nx=40, ny=30, dx=10.0, seed=5) -> Case:
    """Closed basin with a still, flat water surface over a bumpy bed."""
    rng = np.random.default_rng(seed)
    depth = 3.0 + _bumps(rng, nx, ny, 6, 0.5, 2, 6)
    zb = f32(-depth).ravel()
    return Case("lake", nx, ny, dx, dx, zb, f32(np.full(nx * ny, 0.03)), None, None, 1,
                [0.0], [M2_SPEED_DEG_HR], [0.0], 60.0, 10)


# ---------------------------------------------------------------------------
# Telescoping grids
@dataclasses.dataclass
class TelGrid:
    """Cells of a telescoping (quadtree) grid in the layout of a .tel file
    (read_tel, source/geospatial/geospatial.F90:403-795).

    geom [n, 4]: centre x, y and size dx, dy [m]; nb [n, 8]: neighbour slots
    N1 N2 E1 E2 S1 S2 W1 W2 as 1-based cell numbers, 0 for none.  A side with
    one neighbour uses the first slot; a side with two (finer) neighbours lists
    them clockwise around the cell: north west then east, east north then
    south, south east then west, west south then north (the order the dummy
    cells of read_tel assume, :552-690).  base [n]: the base cell j*nx+i that
    holds the cell; level [n]: its refinement level.  bnd_cells: 0-based cells
    of the water-level string in string order."""
    geom: np.ndarray
    nb: np.ndarray
    base: np.ndarray
    level: np.ndarray
    bnd_cells: np.ndarray

    def edge_faces(self):
        """[n, 6] mask of the domain-edge faces in each cell's face order: the
        slots N1 N2 E1 E2 S1 S2 W1 W2 in turn give a face per listed neighbour
        and one to a dummy cell for a side that lists none (read_tel,
        geospatial.F90:552-690); the latter are the edge faces."""
        nb = np.asarray(self.nb)
        even = np.zeros(8, dtype=bool)
        even[0::2] = True
        pair_empty = np.repeat((nb[:, 0::2] == 0) & (nb[:, 1::2] == 0), 2, axis=1)
        edge = even & pair_empty
        face = edge | (nb > 0)
        k = np.cumsum(face, axis=1) - 1
        out = np.zeros((len(nb), 6), dtype=bool)
        rows, cols = np.nonzero(edge)
        out[rows, k[rows, cols]] = True
        return out

    def areas(self):
        g = np.asarray(self.geom)
        return g[:, 2] * g[:, 3]


def telescope(levels, dx, side=1) -> TelGrid:
    """Quadtree over an [ny, nx] base grid of dx x dx cells: base cell (i, j)
    is split into 2^L x 2^L cells with L = levels[j, i] (0..2).  Base cells that
    share an edge differ by one level at most.  Cells are numbered base cell
    by base cell (row-major), the cells of a split base cell row-major from its
    south-west corner.  side: the domain edge of the water-level string."""
    levels = np.asarray(levels, dtype=np.int64)
    ny, nx = levels.shape
    assert levels.min() >= 0 and levels.max() <= 2
    assert np.abs(np.diff(levels, axis=0)).max(initial=0) <= 1
    assert np.abs(np.diff(levels, axis=1)).max(initial=0) <= 1
    m = 1 << levels.ravel()                        # cells per side of each base cell
    cnt = m * m
    first = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    n = int(cnt.sum())
    base = np.repeat(np.arange(nx * ny), cnt)
    q = np.arange(n) - np.repeat(first, cnt)       # position inside the base cell
    mm = np.repeat(m, cnt)
    a, b = q % mm, q // mm
    lev = np.repeat(levels.ravel(), cnt)
    I = (base % nx) * mm + a                       # lattice position at the cell's own level
    J = (base // nx) * mm + b
    h = dx / mm
    geom = f32(np.stack([I * h + 0.5 * h, J * h + 0.5 * h, h, h], axis=1))
    def ids(L, Ia, Ja):
        # the number of the level-L cell at lattice position (Ia, Ja) of a base
        # cell of level L: the base cell's first cell plus its row-major offset
        bb = (Ja >> L) * nx + (Ia >> L)
        mk = (1 << L) - 1
        return first[bb] + (Ja & mk) * (1 << L) + (Ia & mk)

    nb = np.zeros((n, 8), dtype=np.int32)
    # per side: (di, dj) and the offsets of the two finer neighbours in slot order:
    # north west then east, east north then south, south east then west, west south then north
    sides = [((0, 1), [(0, 1), (1, 1)]), ((1, 0), [(1, 1), (1, 0)]),
             ((0, -1), [(1, 0), (0, 0)]), ((-1, 0), [(0, 0), (0, 1)])]
    for L in range(3):
        sel = np.nonzero(lev == L)[0]
        if sel.size == 0:
            continue
        Ic, Jc = I[sel], J[sel]
        for s_, ((di, dj), fine) in enumerate(sides):
            In, Jn = Ic + di, Jc + dj
            inside = (In >= 0) & (In < nx << L) & (Jn >= 0) & (Jn < ny << L)
            c, In, Jn = sel[inside], In[inside], Jn[inside]
            Lb = levels[Jn >> L, In >> L]
            same = Lb == L
            nb[c[same], 2 * s_] = ids(L, In[same], Jn[same]) + 1
            if L > 0:
                up = Lb == L - 1
                nb[c[up], 2 * s_] = ids(L - 1, In[up] >> 1, Jn[up] >> 1) + 1
            if L < 2:
                dn = Lb == L + 1
                for qq, (oa, ob) in enumerate(fine):
                    if di:
                        fi = 2 * In[dn] + (0 if di > 0 else 1)
                        fj = 2 * (Jn[dn] - dj) + ob
                    else:
                        fi = 2 * (In[dn] - di) + oa
                        fj = 2 * Jn[dn] + (0 if dj > 0 else 1)
                    nb[c[dn], 2 * s_ + qq] = ids(L + 1, fi, fj) + 1
    nfaces = np.maximum((nb[:, 0::2] > 0) + (nb[:, 1::2] > 0), 1).sum(axis=1)
    assert nfaces.max() <= 6, "a cell with finer neighbours on more than two sides (nmaxfaces = 6)"
    # the water-level string: the cells along the edge `side`, in order along it
    g = geom
    Lx, Ly = nx * dx, ny * dx
    if side == 1:
        sel = np.nonzero(np.isclose(g[:, 0] + 0.5 * g[:, 2], Lx))[0]
        order = sel[np.argsort(g[sel, 1], kind="stable")]
    elif side == 3:
        sel = np.nonzero(np.isclose(g[:, 0] - 0.5 * g[:, 2], 0.0))[0]
        order = sel[np.argsort(g[sel, 1], kind="stable")]
    elif side == 0:
        sel = np.nonzero(np.isclose(g[:, 1] + 0.5 * g[:, 3], Ly))[0]
        order = sel[np.argsort(g[sel, 0], kind="stable")]
    else:
        sel = np.nonzero(np.isclose(g[:, 1] - 0.5 * g[:, 3], 0.0))[0]
        order = sel[np.argsort(g[sel, 0], kind="stable")]
    return TelGrid(geom, nb, base.astype(np.int32), lev.astype(np.int32), order.astype(np.int32))


def band_levels(nx, ny, c1, c2):
    """Levels of a shore-parallel band: base columns [c1[0], c1[1]) at level 1
    around the columns [c2[0], c2[1]) at level 2."""
    lev = np.zeros((ny, nx), dtype=np.int32)
    lev[:, c1[0]:c1[1]] = 1
    lev[:, c2[0]:c2[1]] = 2
    return lev


def _smooth_at(rng, xn, yn, nmodes, peak):
    f = np.zeros_like(xn)
    for _ in range(nmodes):
        kx, ky = rng.integers(1, 4), rng.integers(1, 4)
        ph1, ph2 = rng.uniform(0, 2 * np.pi, 2)
        f += rng.normal() * np.sin(np.pi * kx * xn + ph1) * np.sin(np.pi * ky * yn + ph2)
    return f * (peak / max(np.abs(f).max(), 1e-30))


def tidal_tel(nx=2000, ny=1500, dx=20.0, seed=1, nsteps=2880, dt=30.0, band=None, radstress=True,
              levels=None) -> Case:
    """The tidal workload on a telescoping grid: the beach of ``tidal``
    (bed from +1 m at the west edge to -15 m at the east), wetting and drying,
    lognormal Manning (median 0.025, sigma_log 0.2), a smooth radiation-stress
    field up to 0.05 N/m^2, M2 tide on the east side, and a band of 10 m cells
    with a 5 m core along the shoreline (base columns band = ((a1, b1), (a2,
    b2)); by default 5.4-7.4 % and 5.8-7.0 % of the width, around the
    shoreline at 6.25 %); or any level map ``levels`` [ny, nx].  Fields are
    evaluated at the cell centres."""
    if levels is None:
        if band is None:
            band = ((int(round(0.054 * nx)), int(round(0.074 * nx))),
                    (int(round(0.058 * nx)), int(round(0.070 * nx))))
        levels = band_levels(nx, ny, band[0], band[1])
    tel = telescope(levels, dx, side=1)
    rng = np.random.default_rng(seed)
    x, y = tel.geom[:, 0], tel.geom[:, 1]
    xn, yn = x / (nx * dx), y / (ny * dx)
    depth = -1.0 + 16.0 * xn
    # Gaussian bumps in base-cell units (sigma 0.2 m)
    xi, yj = x / dx, y / dx
    for _ in range(40):
        ci, cj = rng.uniform(0, nx), rng.uniform(0, ny)
        r = rng.uniform(max(2, nx / 80), max(3, nx / 20))
        a = rng.normal(0.0, 0.2)
        depth = depth + a * np.exp(-((xi - ci) ** 2 + (yj - cj) ** 2) / (2 * r * r))
    zb = f32(-depth)
    n = len(x)
    mann = f32(0.025 * np.exp(0.2 * rng.standard_normal(n)))
    wx = wy = None
    noptset = 2
    if radstress:
        tx = _smooth_at(rng, xn, yn, 6, 0.05)
        ty = _smooth_at(rng, xn, yn, 6, 0.05)
        mag = np.sqrt(tx * tx + ty * ty)
        scale = np.where(mag > 0.05, 0.05 / np.maximum(mag, 1e-30), 1.0)
        wx = f32(tx * scale / RHOW)
        wy = f32(ty * scale / RHOW)
        noptset = 3
    return Case("tidal_tel", nx, ny, dx, dx, zb, mann, wx, wy, 1, [0.5], [M2_SPEED_DEG_HR], [90.0],
                dt, nsteps, noptset=noptset, tel=tel)


def patch_levels(nx, ny):
    """An L-shaped refined patch that touches the south edge: level 1 around a
    level-2 core; at its inner corner a coarse cell meets finer cells on two
    sides (six faces)."""
    lev = np.zeros((ny, nx), dtype=np.int32)
    a, b = nx // 10, nx // 10 + max(6, nx // 5)
    lev[0:ny // 2, a:b] = 1
    lev[0:ny // 4, a:b + max(4, nx // 8)] = 1
    lev[0:ny // 2 - 2, a + 2:b - 2] = 2
    return lev
