"""Domain mass balance: over one step the change of the water volume equals
the net inflow through the open boundary (the continuity equation solved by
the pressure correction, source/flow/flow_imp.F90:164-206; the boundary sum
of flow_balance, :2260-2372).  Wetting and drying and the outer-iteration
tolerance leave a small residue: at most 0.29 % of the boundary flow over
120 steps of the 200x150 tidal case in the oracle, so the tests allow 0.5 %.
A dropped flux term, a flux of the wrong sign or a face taken twice breaks
the balance by the size of the flow itself."""
import os
import sys

import numpy as np
import pytest

from cms2d_amd import cases

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))

TOL_REL = 5e-3       # of the net boundary flow


def boundary_inflow(f, c):
    """Net inflow (m^3/s) from the face fluxes [ncells, 4] (N, E, S, W),
    outward positive, already multiplied by the face length."""
    f = f.reshape(c.ny, c.nx, 4)
    return -(f[-1, :, 0].sum() + f[:, -1, 1].sum() + f[0, :, 2].sum() + f[:, 0, 3].sum())


def check_balance(model, c, spin, steps):
    tel = getattr(c, "tel", None)
    if tel is not None:
        area = tel.areas()
        edge = tel.edge_faces()
        inflow = lambda f: -f[edge].sum()  # noqa: E731
    else:
        area = c.dx * c.dy
        inflow = lambda f: boundary_inflow(f, c)  # noqa: E731
    for _ in range(spin):
        model.step(1)
    worst = 0.0
    for _ in range(steps):
        v0 = (model.get("h") * area).sum()
        model.step(1)
        v1 = (model.get("h") * area).sum()
        q = inflow(model.get_face("flux"))
        err = abs((v1 - v0) / c.dt - q)
        worst = max(worst, err / abs(q))
        assert err <= TOL_REL * abs(q), ((v1 - v0) / c.dt, q)
    return worst


def test_oracle_mass_balance():
    from pyoracle import Oracle
    c = cases.tidal(nx=200, ny=150, dx=20.0, seed=1, dt=30.0)
    check_balance(Oracle(c, serial=0), c, spin=20, steps=20)


def test_oracle_mass_balance_telescoping():
    """The balance on a telescoping grid (edge faces from the .tel slots),
    where the skewness corrections of the joint cells enter the fluxes."""
    from pyoracle import Oracle
    c = cases.tidal_tel(nx=60, ny=40, dx=20.0, band=((2, 10), (4, 8)), seed=1)
    check_balance(Oracle(c, serial=0), c, spin=20, steps=10)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_gpu_mass_balance_telescoping_full_size(gpu_available, precision):
    """The 3.61 M-cell telescoping tidal case of the benchmark (FP64 form and
    the FP32 refinement form)."""
    from cms2d_amd.flow import FlowStep
    c = cases.tidal_tel()
    print(f"mass balance telescoping tidal, {c.ncells} cells, precision {precision}", flush=True)
    g = FlowStep(c, precision=precision)
    worst = check_balance(g, c, spin=30, steps=5)
    print(f"worst relative imbalance {worst:.2e}", flush=True)
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_gpu_mass_balance_full_size(gpu_available, precision):
    """The 2000x1500 tidal case (3M cells, radiation stress on)."""
    from cms2d_amd.flow import FlowStep
    c = cases.tidal()
    print(f"mass balance 2000x1500 tidal, precision {precision}", flush=True)
    g = FlowStep(c, precision=precision)
    worst = check_balance(g, c, spin=30, steps=5)
    print(f"worst relative imbalance {worst:.2e}", flush=True)
    g.close()
