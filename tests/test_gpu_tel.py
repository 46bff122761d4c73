"""The implicit step on telescoping grids (cms2d_flow_create_tel) against the
FP64 oracle, step by step, and against the compiled reference's own runs on
the same .tel cases (tests/golden/ref_tel_*.npz)."""
import os

import numpy as np
import pytest

from cms2d_amd import cases
from cms2d_amd.flow import FlowStep
from pyoracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
TOL_ETA, TOL_VEL = 2e-7, 5e-6          # the reference's single precision (test_reference_vectors)

TEL = {
    "tel_band": lambda: cases.tidal_tel(nx=40, ny=24, dx=20.0, band=((2, 10), (4, 8)), seed=1),
    "tel_patch": lambda: cases.tidal_tel(nx=40, ny=24, dx=20.0, seed=2, levels=cases.patch_levels(40, 24)),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(TEL))
@pytest.mark.parametrize("precond", ["mg", "jacobi"])
def test_tel_step_matches_oracle(gpu_available, name, precond):
    from cms2d_amd import _lib
    c = TEL[name]()
    print(f"tel {name} precond {precond} cells {c.ncells}", flush=True)
    pc = _lib.PRECOND_MG if precond == "mg" else _lib.PRECOND_JACOBI
    g = FlowStep(c, precond=pc)
    o = Oracle(c, serial=0)
    worst = 0.0
    for s in range(12):
        info = g.step(1)
        rc, oi = o.step(1)
        assert rc == 0
        assert np.array_equal(g.get("iwet"), o.get("iwet")), s
        w = o.get("iwet") > 0
        de = np.abs(g.get("eta") - o.get("eta"))[w].max()
        du = max(np.abs(g.get("u") - o.get("u")).max(), np.abs(g.get("v") - o.get("v")).max())
        worst = max(worst, de, du)
        assert de <= 1e-9 and du <= 1e-9, (s, de, du, info, oi)
    # face fluxes of the final state, every directed face of every cell
    fo, fg = o.get_face("flux"), g.get_face("flux")
    ncf, _, _, _ = o.topology()
    for i in range(c.ncells):
        assert np.abs(fg[i, :ncf[i]] - fo[i, :ncf[i]]).max() <= 1e-9 * max(1.0, np.abs(fo).max())
    print(name, precond, "max difference to the oracle", worst, "pp iterations", info["kry"], flush=True)
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(TEL))
def test_tel_step_matches_reference_vectors(gpu_available, name):
    gold = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    c = TEL[name]()
    print(f"tel {name} against the reference's run", flush=True)
    g = FlowStep(c)
    step = 0
    for s in gold["steps"]:
        while step < s:
            g.step(1)
            step += 1
        iw = gold[f"iwet_{s}"]
        assert np.array_equal(g.get("iwet"), iw), s
        wet = iw > 0
        for f, tol in (("eta", TOL_ETA), ("u", TOL_VEL), ("v", TOL_VEL)):
            d = np.abs(g.get(f) - gold[f"{f}_{s}"])[wet].max()
            assert d <= tol, (s, f, d)
    g.close()


@pytest.mark.gpu
def test_tel_uniform_grid_equals_cartesian_flow(gpu_available):
    """A telescoping cell list whose cells are all base cells is the
    rectangle: the cell-list step agrees with the Cartesian step."""
    import dataclasses
    c = cases.tidal(nx=32, ny=20, dx=20.0, seed=1)
    t = cases.telescope(np.zeros((20, 32), dtype=int), 20.0, side=1)
    a, b = FlowStep(c), FlowStep(dataclasses.replace(c, tel=t))
    for _ in range(8):
        a.step(1)
        b.step(1)
    assert np.array_equal(a.get("iwet"), b.get("iwet"))
    for f in ("eta", "u", "v"):
        assert np.abs(a.get(f) - b.get(f)).max() <= 1e-10, f
    a.close()
    b.close()
