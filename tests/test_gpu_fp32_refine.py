"""The FP32 form: pressure-correction BiCGStab and its SpMV in FP32 inside
FP64 iterative refinement (FP64 residual, FP64 update of pp, FP64 stop test at
the FP64 solve's threshold), everything else FP64.  Checked against the FP64
oracle at the drop-in's tolerance (1e-6 m water level, 1e-6 m/s velocity), and
the domain volume against the oracle's."""
import numpy as np
import pytest

from cms2d_amd import _lib, cases
from cms2d_amd.flow import FlowStep
from pyoracle import Oracle

TOL_ETA = 1e-6   # m
TOL_VEL = 1e-6   # m/s


@pytest.mark.gpu
@pytest.mark.parametrize("shape,steps", [((48, 32), 40), ((201, 121), 8)])
def test_fp32_refine_matches_fp64_oracle(gpu_available, shape, steps):
    nx, ny = shape
    c = cases.tidal(nx=nx, ny=ny, dx=20.0, seed=1)
    print(f"fp32 refine tidal {nx}x{ny} steps {steps}", flush=True)
    o = Oracle(c, serial=0)
    g = FlowStep(c, precision=_lib.PREC_FP32_REFINE)
    for s in range(steps):
        assert o.step(1)[0] == 0
        info = g.step(1)
        assert info["iconv"] in (1, 2)
    w = o.get("iwet")
    assert np.array_equal(w, g.get("iwet"))
    wet = w > 0
    err = {}
    for name, tol in (("eta", TOL_ETA), ("u", TOL_VEL), ("v", TOL_VEL)):
        e = np.abs(o.get(name) - g.get(name))[wet].max()
        err[name] = e
        assert e <= tol, (name, e)
    area = c.dx * c.dy
    dvol = abs(area * (o.get("h").sum() - g.get("h").sum()))
    print("max err", err, "volume difference m^3", dvol, "kry", info["kry"])
    assert dvol <= 1e-6 * area * wet.sum()


@pytest.mark.gpu
@pytest.mark.parametrize("band", ["band", "patch"])
def test_fp32_refine_telescoping_matches_fp64_oracle(gpu_available, band):
    """The FP32 form on a telescoping grid (FP32 coefficients, vectors and
    quadtree V-cycle in the inner solve; FP64 refinement), against the FP64
    oracle on the same cell list, and its volume."""
    if band == "band":
        c = cases.tidal_tel(nx=40, ny=24, dx=20.0, band=((2, 10), (4, 8)), seed=1)
    else:
        c = cases.tidal_tel(nx=40, ny=24, dx=20.0, seed=2, levels=cases.patch_levels(40, 24))
    print(f"fp32 refine telescoping {band}, {c.ncells} cells", flush=True)
    o = Oracle(c, serial=0)
    g = FlowStep(c, precision=_lib.PREC_FP32_REFINE)
    for _ in range(12):
        assert o.step(1)[0] == 0
        info = g.step(1)
        assert info["iconv"] in (1, 2)
    w = o.get("iwet")
    assert np.array_equal(w, g.get("iwet"))
    wet = w > 0
    err = {}
    for name, tol in (("eta", TOL_ETA), ("u", TOL_VEL), ("v", TOL_VEL)):
        e = np.abs(o.get(name) - g.get(name))[wet].max()
        err[name] = e
        assert e <= tol, (name, e)
    area = np.asarray(c.tel.geom)[:, 2] * np.asarray(c.tel.geom)[:, 3]
    dvol = abs((area * (o.get("h") - g.get("h"))).sum())
    print("max err", err, "volume difference m^3", dvol, "kry", info["kry"], flush=True)
    assert dvol <= 1e-6 * area[wet].sum()
    g.close()
