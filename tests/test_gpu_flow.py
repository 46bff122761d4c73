"""GPU parity of the implicit step against the FP64 oracle (race-free mode).

Water level and velocity must agree within 1e-6 m and 1e-6 m/s after the steps
run (the tolerance the north star states for the FP64 drop-in); wet/dry flags
must be identical.
"""
import numpy as np
import pytest

from cms2d_amd import _lib, cases
from cms2d_amd.flow import FlowStep
from pyoracle import Oracle
from test_reference_vectors import ALL

# The WSE_GRADIENTS run drives strong flow along the wet/dry edge, where the
# reference's OpenMP loops that read values other iterations write (the
# dry-to-wet start values and the eddy-viscosity fix-ups) give an answer that
# depends on their order: the race-free default differs from the reference's
# one-thread order by 9e-5 m/s there (4e-5 m/s with the curve on the north
# side), above the bounds.  With serial_order=1
# the GPU follows that order and matches the case's vectors
# (test_gpu_serial_order.py); the race-free default is checked against the
# race-free oracle (test_wse_gradient_matches_oracle).
# With a flux string on the north or south side the same order dependence
# moves velocities by 9e-6 and 4e-6 m/s (the oracle in race-free mode differs
# from the reference by the same; in the reference's order both match the
# double-precision runs to 5e-9 m/s, test_gpu_serial_order.py).
REF_CASES = [n for n in ALL if n not in ("tidal_wse_grad", "tidal_wse_grad_north", "river_north",
                                         "river_conveyance")]

TOL_ETA = 1e-6   # m
TOL_VEL = 1e-6   # m/s


def run_pair(case, nsteps, **kw):
    print(f"case {case.name} {case.nx}x{case.ny} steps {nsteps}", flush=True)
    o = Oracle(case, serial=0)
    g = FlowStep(case, **kw)
    for s in range(nsteps):
        rco, io = o.step(1)
        ig = g.step(1)
        assert rco == 0
        assert io["niter"] == ig["niter"], (s, io, ig)
    return o, g


def compare(o, g):
    wo, wg = o.get("iwet"), g.get("iwet")
    assert np.array_equal(wo, wg)
    wet = wo > 0
    err = {}
    for name, tol in (("eta", TOL_ETA), ("u", TOL_VEL), ("v", TOL_VEL)):
        a, b = o.get(name), g.get(name)
        e = np.abs(a - b)[wet].max() if wet.any() else 0.0
        err[name] = e
        assert e <= tol, (name, e)
    return err


@pytest.mark.gpu
def test_lake_at_rest(gpu_available):
    c = cases.lake_at_rest()
    g = FlowStep(c)
    for _ in range(5):
        g.step(1)
    assert np.abs(g.get("u")).max() == 0.0
    assert np.abs(g.get("v")).max() == 0.0
    assert np.abs(g.get("eta")).max() == 0.0


PRECONDS = [pytest.param(_lib.PRECOND_MG, id="mg"), pytest.param(_lib.PRECOND_JACOBI, id="jacobi")]


@pytest.mark.gpu
@pytest.mark.parametrize("precond", PRECONDS)
def test_basin_small_matches_oracle(gpu_available, precond):
    c = cases.basin(nx=40, ny=20, dx=10.0, seed=0)
    o, g = run_pair(c, 30, precond=precond)
    err = compare(o, g)
    print("max err", err)


@pytest.mark.gpu
@pytest.mark.parametrize("precond", PRECONDS)
def test_tidal_wetdry_radstress_matches_oracle(gpu_available, precond):
    c = cases.tidal(nx=48, ny=32, dx=20.0, seed=1)
    o, g = run_pair(c, 40, precond=precond)
    err = compare(o, g)
    print("max err", err)


@pytest.mark.gpu
def test_tidal_multigrid_levels_matches_oracle(gpu_available):
    """Odd sizes through four multigrid levels (403x243 -> 202x122 -> 101x61 -> 51x31)."""
    c = cases.tidal(nx=403, ny=243, dx=20.0, seed=4)
    o, g = run_pair(c, 6)
    err = compare(o, g)
    info = g.step(0)                    # counters of the last step
    print("max err", err, info)
    # the preconditioner's quality: a wrong cycle still converges, only slowly
    assert info["kry"][0] <= 15 * info["niter"], info


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(3, 3), (3, 40), (40, 3), (3, 5), (7, 3)],
                         ids=lambda s: f"{s[0]}x{s[1]}")
def test_small_shapes_match_oracle(gpu_available, shape):
    """The smallest grids accepted (three cells a side): hierarchies of one-
    and two-cell levels, tiles and Krylov strips mostly empty, one Krylov
    block."""
    c = cases.tidal(nx=shape[0], ny=shape[1], dx=20.0, seed=7)
    o, g = run_pair(c, 20)
    err = compare(o, g)
    print("max err", err)


@pytest.mark.parametrize("shape", [(1, 40), (40, 1), (2, 2), (1, 1)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_grids_under_three_cells_a_side_rejected(shape):
    """cms2d_flow_create refuses grids with fewer than three cells a side
    (CMS2D_ERR_ARG) before touching a device."""
    c = cases.tidal(nx=shape[0], ny=shape[1], dx=20.0, seed=7)
    with pytest.raises(_lib.Cms2dError) as e:
        FlowStep(c)
    assert e.value.code == -1                       # CMS2D_ERR_ARG


@pytest.mark.parametrize("field,value", [("friction", 4), ("friction", -1), ("turbulence", 5),
                                         ("coriolis_plane", 2), ("latitude", 91.0)])
def test_out_of_range_options_rejected(field, value):
    """cms2d_flow_create refuses option values outside the documented sets
    (CMS2D_ERR_ARG) before touching a device."""
    import ctypes as C
    zb = np.full(16, -2.0)
    mann = np.full(16, 0.02)
    d = _lib.FlowDesc()
    d.nx = d.ny = 4
    d.dx = d.dy = 10.0
    d.zb, d.mann = zb.ctypes.data_as(_lib.D), mann.ctypes.data_as(_lib.D)
    d.dt, d.hdry, d.kmax, d.rtol = 10.0, 0.05, 100, 1e-10
    setattr(d, field, value)
    h = C.c_void_p()
    assert _lib.lib().cms2d_flow_create(C.byref(d), C.byref(h)) == -1
    assert not h.value


@pytest.mark.gpu
@pytest.mark.parametrize("ndsch", [0, 1, 2, 3, 5, 6, 7, 8, 9])
def test_advection_schemes_match_oracle(gpu_available, ndsch):
    """Every ADVECTION_SCHEME the Cartesian path has (the default 4 is in the
    other tests): none, upwind, hybrid, power law, and upwind plus the NVD
    deferred corrections (HLPA, GAMMA, CUBISTA, ALVSMART, HOAB), on the
    wet/dry tidal beach."""
    c = cases.tidal(nx=48, ny=32, dx=20.0, seed=1)
    c.ndsch = ndsch
    o, g = run_pair(c, 40)
    err = compare(o, g)
    print("max err", err)


@pytest.mark.gpu
@pytest.mark.parametrize("plane", [0, 1])
def test_coriolis_matches_oracle(gpu_available, plane):
    """The Coriolis parameter per cell (f- and beta-plane at 30 N) and the flow
    it turns, on a basin 20 km across (beta (y - mean y) up to 2 % of f)."""
    c = cases.tidal(nx=100, ny=60, dx=200.0, seed=3)
    c.latitude, c.coriolis_plane = 30.0, plane
    o = Oracle(c, serial=0)
    g = FlowStep(c)
    a, b = o.get("fc"), g.get("fc")
    assert np.abs(a - b).max() <= 1e-15 * np.abs(a).max()
    print(f"coriolis plane {plane}: fc spread {a.max() - a.min():.2e}", flush=True)
    g.close()
    o, g = run_pair(c, 20)
    print("max err", compare(o, g))


@pytest.mark.gpu
@pytest.mark.parametrize("name", REF_CASES)
def test_gpu_matches_reference_vectors(gpu_available, name):
    """The HIP path against the compiled reference's own outputs."""
    from test_reference_vectors import TOL_ETA, TOL_VEL, compare_to_gold
    print(f"reference vectors {name}", flush=True)
    w = compare_to_gold(name, lambda c: FlowStep(c))
    print(name, w)
    assert w["eta"] <= TOL_ETA and w["u"] <= TOL_VEL and w["v"] <= TOL_VEL, w


@pytest.mark.gpu
@pytest.mark.parametrize("shape,steps", [((48, 32), 40), ((201, 121), 8)])
def test_gmres_matches_oracle(gpu_available, shape, steps):
    """The pp system by restarted GMRES (pgmres) with the multigrid
    preconditioner: same converged solutions as the oracle's BiCGStab."""
    nx, ny = shape
    c = cases.tidal(nx=nx, ny=ny, dx=20.0, seed=1 if nx < 100 else 4)
    o, g = run_pair(c, steps, solver=_lib.SOLVER_GMRES)
    err = compare(o, g)
    info = g.step(0)
    print("max err", err, info)
    # one restart cycle (20 steps) or two per solve; a broken Arnoldi process
    # would still converge through restarts, only far more slowly
    assert info["kry"][0] <= 2 * 20 * info["niter"], info


@pytest.mark.gpu
@pytest.mark.parametrize("precision,kind", [(0, "cartesian"), (1, "cartesian"), (0, "telescoping")])
def test_graph_replay_matches_direct_launches(gpu_available, monkeypatch, precision, kind):
    """BiCGStab iteration pairs replayed from a captured graph (the default
    launch mode) give bitwise the state and iteration counts of direct
    launches, on a Cartesian and a telescoping grid (whose pp pair carries
    the quadtree V-cycle)."""
    if kind == "cartesian":
        c = cases.tidal(nx=96, ny=64, dx=20.0, seed=5)
    else:
        c = cases.tidal_tel(nx=60, ny=40, dx=20.0, band=((2, 10), (4, 8)), seed=5)
    runs = []
    for graphs in ("1", "0"):
        monkeypatch.setenv("CMS2D_GRAPHS", graphs)
        print(f"graphs {graphs} precision {precision} {kind}, 12 steps", flush=True)
        g = FlowStep(c, precision=precision)
        kry = [g.step(1)["kry"] for _ in range(12)]
        runs.append((kry, [g.get(n) for n in ("eta", "u", "v", "p")]))
        g.close()
    assert runs[0][0] == runs[1][0]
    for a, b in zip(runs[0][1], runs[1][1]):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_multigrid_tile_widths_match(gpu_available, monkeypatch):
    """Every level-0 tile width the autotuner can pick gives bitwise the same
    step (halo values are recomputed by the same expressions), on a grid whose
    sizes are no multiple of any tile."""
    c = cases.tidal(nx=301, ny=203, dx=20.0, seed=6)
    runs = []
    for tile in ("64", "128", "256"):
        monkeypatch.setenv("CMS2D_MG_TILE", tile)
        print(f"tile {tile} tidal 301x203, 6 steps", flush=True)
        g = FlowStep(c)
        assert g.tuning()[0] == int(tile)
        kry = [g.step(1)["kry"] for _ in range(6)]
        runs.append((kry, [g.get(n) for n in ("eta", "u", "v")]))
        g.close()
    for r in runs[1:]:
        assert r[0] == runs[0][0]
        for a, b in zip(r[1], runs[0][1]):
            assert np.array_equal(a, b)
    monkeypatch.delenv("CMS2D_MG_TILE")
    g = FlowStep(c)
    assert g.tuning()[0] in (64, 128, 256)
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("law", [1, 3, 4])
def test_friction_laws_match_oracle(gpu_available, law):
    """Drag coefficient, roughness height and linear friction (fric_eval's
    mbedfric 1, 3, 4) on a 201 x 121 tidal case with wet/dry cells: state
    within the step tolerances, and the derived coefficients cfrict, coefman
    and z0 within 1e-6 relative per wet cell (functions of the depth, which
    agrees to about 1e-9 m; the linear law's cfrict divides by the previous
    current speed, so it is compared where that speed is above 1 mm/s)."""
    from test_reference_vectors import _tidal_fric
    c = _tidal_fric(law)()
    c2 = cases.tidal(nx=201, ny=121, dx=20.0, seed=4)
    rng = np.random.default_rng(7)
    n = c2.nx * c2.ny
    c2.mbedfric = law
    c2.mann = (np.float32(0.002 + 0.004 * rng.random(n)).astype(np.float64) if law == 1 else
               rng.integers(4, 41, size=n) / 4096.0 if law == 3 else np.full(n, float(np.float32(1e-4))))
    for case, steps in ((c, 20), (c2, 8)):
        o, g = run_pair(case, steps)
        err = compare(o, g)
        sel = o.get("iwet") > 0
        if law == 4:
            sel &= o.get("uv") > 1e-3
        assert sel.sum() > 50
        for f in ("cfrict", "coefman", "z0"):
            a, b = o.get(f)[sel], g.get(f)[sel]
            d = (np.abs(a - b) / np.abs(a)).max()
            assert d <= 1e-6, (f, d)
        print(f"friction law {law} {case.nx}x{case.ny}: {err}", flush=True)
        g.close()


@pytest.mark.gpu
def test_maxeta_matches_oracle(gpu_available):
    """The running maximum of the water level (maxeta) after 30 steps of the
    wet/dry tidal beach, within the water-level tolerance."""
    c = cases.tidal(nx=48, ny=32, dx=20.0, seed=1)
    o, g = run_pair(c, 30)
    compare(o, g)
    e = np.abs(o.get("maxeta") - g.get("maxeta")).max()
    print("maxeta err", e, flush=True)
    assert e <= TOL_ETA
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp32_refine", "gmres", "jacobi"])
@pytest.mark.parametrize("name", ["tidal", "basin_200x100"])
def test_gpu_forms_match_reference_vectors(gpu_available, form, name):
    """The FP32-refinement form, GMRES(20) and the Jacobi-preconditioned pp
    solve against the compiled reference's own outputs, to the same bounds as
    the default form."""
    from test_reference_vectors import TOL_ETA, TOL_VEL, compare_to_gold
    kw = {"fp32_refine": {"precision": _lib.PREC_FP32_REFINE}, "gmres": {"solver": _lib.SOLVER_GMRES},
          "jacobi": {"precond": _lib.PRECOND_JACOBI}}[form]
    print(f"reference vectors {name} {form}", flush=True)
    w = compare_to_gold(name, lambda c: FlowStep(c, **kw))
    print(name, form, w)
    assert w["eta"] <= TOL_ETA and w["u"] <= TOL_VEL and w["v"] <= TOL_VEL, w


@pytest.mark.parametrize("times,values", [(None, None), ([0.0, 0.2, 0.1], [0.0, 0.1, 0.2]),
                                          ([0.0, 0.1], [0.0, float("nan")]), ([], [])],
                         ids=["missing", "not_increasing", "nan", "empty"])
def test_bad_water_level_curves_rejected(times, values):
    """A CMS2D_BND_WSE_CURVE side without a usable curve is refused
    (CMS2D_ERR_ARG) before touching a device."""
    import ctypes as C
    zb = np.full(16, -2.0)
    mann = np.full(16, 0.02)
    d = _lib.FlowDesc()
    d.nx = d.ny = 4
    d.dx = d.dy = 10.0
    d.zb, d.mann = zb.ctypes.data_as(_lib.D), mann.ctypes.data_as(_lib.D)
    d.dt, d.hdry, d.kmax, d.rtol = 10.0, 0.05, 100, 1e-10
    d.side_type[1] = _lib.BND_WSE_CURVE
    keep = []
    if times is not None:
        t, v = np.array(times, dtype=np.float64), np.array(values, dtype=np.float64)
        keep += [t, v]
        d.wse_times, d.wse_values, d.wse_ntimes = t.ctypes.data_as(_lib.D), v.ctypes.data_as(_lib.D), len(t)
    h = C.c_void_p()
    assert _lib.lib().cms2d_flow_create(C.byref(d), C.byref(h)) == -1
    assert not h.value


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tidal_wse_grad", "tidal_wse_grad_north"])
def test_wse_gradient_matches_oracle(gpu_available, name):
    """The water-level curve with an along-side gradient (WSE_GRADIENTS) against
    the race-free oracle over 40 steps: on the east side varying along y, on
    the north side along x."""
    from test_reference_vectors import CASES
    c = CASES[name]()
    o, g = run_pair(c, 40)
    print("max err", compare(o, g), flush=True)
    g.close()
