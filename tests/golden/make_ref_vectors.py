#!/usr/bin/env python3
"""Writes golden vectors from the reference's own implicit step.

Runs oracle/_ref/ref_run (the reference's Fortran sources compiled unmodified
by oracle/build_ref.py, single precision as shipped, BiCGStab with ILUTP and
its 1e-3 inner tolerance) on seeded cases and stores eta, u, v and iwet after
selected steps in tests/golden/ref_<case>.npz; with --double the same from
oracle/_ref/ref_run8 (the reference built in double precision, build_ref.py
--double) into tests/golden/ref64_<case>.npz.  The reference runs on one
OpenMP thread (the run's thread count and precision are stored).  Regenerate
with: python tests/golden/make_ref_vectors.py [--double] [case ...]; with
--balance, the per-step volume balance and extremes of the double-precision
build (ref64_balance_<case>.npz).
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import refcase  # noqa: E402
from cms2d_amd import cases  # noqa: E402

CASES = {
    "basin": (lambda: cases.basin(nx=40, ny=20, dx=10.0, seed=0), 30, [1, 10, 20, 30]),
    "tidal": (lambda: cases.tidal(nx=48, ny=32, dx=20.0, seed=1), 40, [1, 10, 20, 40]),
}
# the NVD deferred corrections (ndsch 5 HLPA, 7 CUBISTA, 9 HOAB) on the tidal case
for _nd, _nm in ((5, "hlpa"), (7, "cubista"), (9, "hoab"), (0, "none"), (1, "upwind"), (2, "hybrid"),
                 (3, "powerlaw"), (6, "gamma"), (8, "alvsmart")):
    def _mk(nd=_nd):
        c = cases.tidal(nx=48, ny=32, dx=20.0, seed=1)
        c.ndsch = nd
        return c
    CASES[f"tidal_{_nm}"] = (_mk, 40, [1, 10, 20, 40])


# Coriolis at 45 degrees north, f-plane and beta-plane
for _pl, _nm in ((0, "fplane"), (1, "betaplane")):
    def _mk(pl=_pl):
        c = cases.tidal(nx=48, ny=32, dx=20.0, seed=1)
        c.latitude, c.coriolis_plane = 45.0, pl
        return c
    CASES[f"tidal_{_nm}"] = (_mk, 40, [1, 10, 20, 40])


# a constant 12 m/s wind from the south-west over the tidal case
def _mk_wind():
    c = cases.tidal(nx=48, ny=32, dx=20.0, seed=1)
    c.wind_speed, c.wind_dir = 12.0, 225.0
    return c


CASES["tidal_wind"] = (_mk_wind, 40, [1, 10, 20, 40])

# the other turbulence models of the Cartesian path
for _mt, _nm in ((0, "constant"), (1, "falconer"), (2, "parabolic"), (3, "subgridwu")):
    def _mk(mt=_mt):
        c = cases.tidal(nx=48, ny=32, dx=20.0, seed=1)
        c.mturbul = mt
        return c
    CASES[f"tidal_turb_{_nm}"] = (_mk, 40, [1, 10, 20, 40])




# the other bottom friction laws on the tidal case: a drag coefficient dataset,
# a roughness height dataset (z0 = k / 4096 m, so the height 30 z0 and its
# division by 30 are exact in single precision) and the linear law's constant
def _mk_fric(law):
    def mk():
        c = cases.tidal(nx=48, ny=32, dx=20.0, seed=1)
        rng = np.random.default_rng(7)
        n = c.nx * c.ny
        c.mbedfric = law
        if law == 1:
            c.mann = np.float32(0.002 + 0.004 * rng.random(n)).astype(np.float64)
        elif law == 3:
            c.mann = rng.integers(4, 41, size=n) / 4096.0
        else:
            c.mann = np.full(n, float(np.float32(1.0e-4)))
        return c
    return mk


for _law, _nm in ((1, "drag"), (3, "roughness"), (4, "linear")):
    CASES[f"tidal_fric_{_nm}"] = (_mk_fric(_law), 40, [1, 10, 20, 40])


# single water level boundaries on the tidal case instead of the tide: a
# five-point curve over the run with the default quadratic interpolation, the
# same with cubic interpolation, and a constant level (WSE_CONSTANT)
WSE_T = [0.0, 0.1, 0.2, 0.3, 0.4]
WSE_V = [0.0, 0.125, 0.0625, 0.25, 0.1875]


def _mk_wse(times, values, order):
    def mk():
        c = cases.tidal(nx=48, ny=32, dx=20.0, seed=1)
        c.wse_times, c.wse_values, c.wse_order = list(times), list(values), order
        return c
    return mk


CASES["tidal_wse_curve"] = (_mk_wse(WSE_T, WSE_V, 0), 40, [1, 10, 20, 40])
CASES["tidal_wse_cubic"] = (_mk_wse(WSE_T, WSE_V, 3), 40, [1, 10, 20, 40])
CASES["tidal_wse_const"] = (_mk_wse([0.0], [0.125], 0), 40, [1, 10, 20, 40])


def _mk_wse_noadj():
    c = _mk_wse(WSE_T, WSE_V, 0)()
    c.wse_adjust = False                            # WSE_ADJUSTMENT OFF
    return c


CASES["tidal_wse_noadj"] = (_mk_wse_noadj, 40, [1, 10, 20, 40])


def _mk_wse_grad():
    c = _mk_wse(WSE_T, WSE_V, 0)()
    c.wse_grad = (0.0, 2.0e-5)                      # WSE_GRADIENTS: 0.02 mm/m along the east side
    return c


CASES["tidal_wse_grad"] = (_mk_wse_grad, 40, [1, 10, 20, 40])


def _mk_wse_grad_north():
    c = _mk_wse(WSE_T, WSE_V, 0)()
    c.th_side = 0                                   # the curve on the north side,
    c.wse_grad = (2.0e-5, 0.0)                      # varying along x: WSE_GRADIENTS 2e-5 0
    return c


CASES["tidal_wse_grad_north"] = (_mk_wse_grad_north, 40, [1, 10, 20, 40])


# the tidal case without wall friction (USE_WALL_FRICTION_TERMS OFF), and the
# tidal case run by the reference's default MATRIX_SOLVER GMRES (its restarted
# GMRES with ILUT; the converged answer does not depend on the Krylov method)
def _mk_nowall():
    c = cases.tidal(nx=48, ny=32, dx=20.0, seed=1)
    c.wallfric = 0
    return c


CASES["tidal_nowallfric"] = (_mk_nowall, 40, [1, 10, 20, 40])
CASES["tidal_refgmres"] = (lambda: cases.tidal(nx=48, ny=32, dx=20.0, seed=1), 40, [1, 10, 20, 40],
                           {"solver": "GMRES"})


# the tide ramped in over a quarter of an hour (DURATION_RAMP, ramp_func)
def _mk_ramp():
    c = cases.tidal(nx=48, ny=32, dx=20.0, seed=1)
    c.rampdur = 0.25
    return c


CASES["tidal_ramp"] = (_mk_ramp, 40, [1, 10, 20, 40])


# the 200 x 100 basin at its full size (dx 10 m, dt 60 s), half an hour (longer
# runs drift from the FP64 answer by the reference's single precision and 1e-3
# inner solves: 2e-5 m/s after two hours)
CASES["basin_200x100"] = (lambda: cases.basin(nx=200, ny=100, dx=10.0, seed=0), 30, [1, 10, 20, 30])


# the options together: HOAB scheme, beta-plane Coriolis at 30 N, a 9 m/s wind
# from the north-east, the FALCONER model, roughness heights and the
# water-level curve on the open side
def _mk_combo():
    c = _mk_fric(3)()
    c.ndsch, c.latitude, c.coriolis_plane, c.mturbul = 9, 30.0, 1, 1
    c.wind_speed, c.wind_dir = 9.0, 45.0
    c.wse_times, c.wse_values, c.wse_order = list(WSE_T), list(WSE_V), 0
    return c


CASES["tidal_combo"] = (_mk_combo, 40, [1, 10, 20, 40])


# telescoping grids (read_tel): a shore-parallel band of 10 m cells around a
# 5 m core across the beach, and an L-shaped patch on the south edge with
# six-face cells; the tidal physics (wet/dry, radiation stress, M2 tide)
CASES["tel_band"] = (lambda: cases.tidal_tel(nx=40, ny=24, dx=20.0, band=((2, 10), (4, 8)), seed=1),
                     30, [1, 10, 20, 30])
CASES["tel_patch"] = (lambda: cases.tidal_tel(nx=40, ny=24, dx=20.0, seed=2,
                                              levels=cases.patch_levels(40, 24)), 30, [1, 10, 20, 30])


# VARIABLE_TIMESTEP ON: the tidal case at dt = 3600 s, whose sixth step ends
# its outer iterations unconverged (iconv 3), so timing halves the next step
# and doubles it back two steps later (CMS_Flow.f90:180-205)
def _mk_dtvar():
    c = cases.tidal(nx=48, ny=32, dx=20.0, seed=1, dt=3600.0)
    c.dtvar = 1
    return c


CASES["tidal_dtvar"] = (_mk_dtvar, 12, [1, 6, 7, 8, 9, 12])

# flux (river) strings (FLUX_BEGIN blocks): a constant inflow on the west side
# with the tide on the east, on the north side, a Lagrange curve, a shoal that
# dries part of the string, and the CONVEYANCE card on the south side
RIVER = {
    "river": dict(),
    "river_north": dict(q_side=0),
    "river_curve": dict(q_curve=([0.0, 0.005, 0.01, 0.3], [0.0, 30.0, 60.0, 45.0]), q_order=2),
    "river_shoal": dict(shoal=True, q=25.0),
    "river_conveyance": dict(q_cmvel=0.8, q_side=2),
    "river_angle": dict(q_angle=75.0),           # INFLOW_ANGLE: 15 degrees off the normal
    "river_rating": dict(q_rating=([1.0, 2.0, 2.5, 4.0], [10.0, 30.0, 45.0, 80.0])),   # RATING_CURVE
}
for _nm, _kw in RIVER.items():
    CASES[_nm] = ((lambda kw=_kw: cases.river(**kw)), 40, [1, 10, 20, 40])
# the flux string on a telescoping grid (constant inflow; a rating curve)
CASES["tel_river"] = (lambda: cases.river_tel(), 40, [1, 10, 20, 40])
CASES["tel_river_rating"] = (lambda: cases.river_tel(q_rating=([1.0, 2.0, 2.5, 4.0], [10.0, 30.0, 45.0, 80.0])),
                             40, [1, 10, 20, 40])


# flow_balance and flow_step_stat of every step (WATER_VOLUME_BALANCE ON):
# ref64_balance_<case>.npz holds "bal" [nsteps, 12] (cumulative storage,
# efflux, flux and norm of flow_balance; pxtrm, idp, uxtrm, idu, vxtrm, idv,
# crmax, idcr of flow_step_stat, cells 1-based); written with --balance
BALANCE = {
    "tidal": (lambda: cases.tidal(nx=48, ny=32, dx=20.0, seed=1), 40),
    "basin": (lambda: cases.basin(nx=40, ny=20, dx=10.0, seed=0), 40),
    "square": (lambda: cases.square(nx=40, ny=24, dx=20.0, seed=3), 40),
    "tel_band": (lambda: cases.tidal_tel(nx=40, ny=24, dx=20.0, band=((2, 10), (4, 8)), seed=1), 40),
    "tel_patch": (lambda: cases.tidal_tel(nx=40, ny=24, dx=20.0, seed=2, levels=cases.patch_levels(40, 24)), 40),
    "tidal_dtvar": (_mk_dtvar, 12),
    "river": (lambda: cases.river(), 40),
    "river_shoal": (lambda: cases.river(shoal=True, q=25.0), 40),
}


# flow statistics (FLOW_STATISTICS 0 0.05 0 hrs: windows of 6 steps of 30 s)
# of the tidal case: ref64_stats_tidal.npz holds every stat_def field after
# step 5 (open window) and step 11 (the second window, after a reset); written
# with --stats
def write_stats():
    c = cases.tidal(nx=48, ny=32, dx=20.0, seed=1)
    with tempfile.TemporaryDirectory() as d:
        res = refcase.run(c, d + "/", 11, exe=refcase.REF_RUN8, stats=(0.0, 0.05, 0.0))
    out = {"steps": np.array([5, 11]), "window_s": np.array([0.0, 180.0])}
    for s in (5, 11):
        out[f"count_{s}"] = np.array(res[s - 1]["statcount"])
        for k, v in res[s - 1]["stats"].items():
            out[f"{k}_{s}"] = v
    np.savez_compressed(os.path.join(HERE, "ref64_stats_tidal.npz"), **out)
    print("wrote ref64_stats_tidal.npz")


def write_balance(names):
    exe = refcase.REF_RUN8
    for name, (mk, nsteps) in BALANCE.items():
        if names and name not in names:
            continue
        with tempfile.TemporaryDirectory() as d:
            res = refcase.run(mk(), d + "/", nsteps, exe=exe, balance=True)
        bal = np.stack([r["bal"] for r in res])
        np.savez_compressed(os.path.join(HERE, f"ref64_balance_{name}.npz"), bal=bal, omp_threads=np.array(1),
                            ikind=np.array(8))
        print(f"wrote ref64_balance_{name}.npz ({nsteps} steps)")


def only(names):
    return {k: v for k, v in CASES.items() if k in names}


def main():
    args = sys.argv[1:]
    double = "--double" in args
    args = [a for a in args if a != "--double"]
    if "--stats" in args:
        write_stats()
        return 0
    if "--balance" in args:
        if not refcase.available(refcase.REF_RUN8):
            print(f"{refcase.REF_RUN8} not built (python oracle/build_ref.py --double)", file=sys.stderr)
            return 1
        write_balance([a for a in args if a != "--balance"])
        return 0
    exe = refcase.REF_RUN8 if double else refcase.REF_RUN
    prefix = "ref64_" if double else "ref_"
    if not refcase.available(exe):
        print(f"{exe} not built (python oracle/build_ref.py{' --double' if double else ''})", file=sys.stderr)
        return 1
    todo = only(args) if args else CASES
    for name, (mk, nsteps, keep, *opt) in todo.items():
        c = mk()
        with tempfile.TemporaryDirectory() as d:
            res = refcase.run(c, d + "/", nsteps, exe=exe, **(opt[0] if opt else {}))
        out = {"steps": np.array(keep), "omp_threads": np.array(1), "ikind": np.array(8 if double else 4)}
        for s in keep:
            for f in ("eta", "u", "v", "iwet"):
                out[f"{f}_{s}"] = res[s - 1][f]
        np.savez_compressed(os.path.join(HERE, f"{prefix}{name}.npz"), **out)
        print(f"wrote {prefix}{name}.npz ({nsteps} steps, kept {keep})")
    return 0


if __name__ == "__main__":
    sys.exit(main())
