"""Flux (river) strings on the HIP path: bndriver_vel's conveyance split of the
inflow (boundaries.F90:4259-4356), bnd_vel_flux (bnd_vel.f90:2-32), bndflux
(boundaries.F90:4359-4414) and the string's ghosts (flow_imp.F90:866-874,
984-995), against the race-free FP64 oracle (whose river code matches the
compiled reference's runs, tests/golden/ref*_river*.npz) at the drop-in's
1e-6 m / 1e-6 m/s and, to show the kernels follow it expression by
expression, at 1e-9; the inflow the faces carry; the FP32 form and GMRES; a
checkpoint across the flux curve; what is rejected."""
import numpy as np
import pytest

from cms2d_amd import _lib, cases
from cms2d_amd.flow import FlowStep
from pyoracle import Oracle
from test_gpu_flow import compare, run_pair
from test_reference_vectors import RIVER

pytestmark = pytest.mark.gpu

TIGHT = 1e-9


@pytest.mark.parametrize("name", list(RIVER))
def test_river_matches_oracle(gpu_available, name):
    c = cases.river(**RIVER[name])
    o, g = run_pair(c, 40)
    err = compare(o, g)
    print(name, err, flush=True)
    assert max(err.values()) <= TIGHT, err
    fo, fg = o.get_face("flux"), g.get_face("flux")
    assert np.abs(fo - fg).max() <= TIGHT * max(1.0, np.abs(fo).max())
    g.close()


def _inflow(g, c):
    f = g.get_face("flux").reshape(c.ny, c.nx, 4)
    side = c.q_side
    edge = [f[-1, :, 0], f[:, -1, 1], f[0, :, 2], f[:, 0, 3]][side]
    return -edge.sum()


def test_river_faces_carry_the_inflow(gpu_available):
    """After each step the string's faces carry ramp x Q (constant) or ramp x
    the curve's Lagrange value at the step's time (linear by default)."""
    c = cases.river(q=40.0)
    print("river inflow, constant", flush=True)
    g = FlowStep(c)
    for _ in range(5):
        g.step(1)
        assert _inflow(g, c) == pytest.approx(40.0, rel=1e-12)
    g.close()
    times, vals = [0.0, 0.01, 0.02], [0.0, 30.0, 60.0]
    c = cases.river(q_curve=(times, vals))
    print("river inflow, curve", flush=True)
    g = FlowStep(c)
    for n in range(1, 5):
        g.step(1)
        t = n * c.dt / 3600.0
        assert _inflow(g, c) == pytest.approx(np.interp(t, times, vals), rel=1e-12)
    g.close()


@pytest.mark.parametrize("form", ["fp32_refine", "gmres"])
def test_river_forms_match_oracle(gpu_available, form):
    c = cases.river(**RIVER["river_curve"])
    kw = {"precision": _lib.PREC_FP32_REFINE} if form == "fp32_refine" else {"solver": _lib.SOLVER_GMRES}
    o, g = run_pair(c, 20, **kw)
    err = compare(o, g)
    print(form, err, flush=True)
    g.close()


def test_river_checkpoint_resumes_bitwise(gpu_available):
    c = cases.river(**RIVER["river_curve"])
    print("river checkpoint", flush=True)
    f = FlowStep(c)
    f.step(5)
    ck = f.checkpoint()
    f.step(5)
    ref = [f.get(k) for k in ("eta", "u", "v")]
    f.restore(ck)
    f.step(5)
    for a, k in zip(ref, ("eta", "u", "v")):
        assert np.array_equal(a, f.get(k)), k
    f.close()


def test_river_rejects():
    c = cases.river()
    c.mbedfric = 1                     # drag law: n changes every friction evaluation
    with pytest.raises(RuntimeError, match="Manning"):
        FlowStep(c)
    t = cases.river_tel()
    t.tel.flux_cells = t.tel.flux_cells[:1]          # a one-cell string
    with pytest.raises(RuntimeError, match="2..8192"):
        FlowStep(t)


def test_river_adjoint_rejected(gpu_available):
    c = cases.river()
    f = FlowStep(c)
    with pytest.raises(RuntimeError, match="flux sides"):
        f.adjoint(2, [5], obs=np.zeros((2, 1)))
    f.close()


@pytest.mark.parametrize("name", ["tel_river", "tel_river_rating"])
def test_river_telescoping_matches_oracle(gpu_available, name):
    """The flux string on a telescoping grid (the band near the mouth), constant
    and by a rating curve, against the oracle on the same cell list."""
    from test_reference_vectors import CASES
    c = CASES[name]()
    print(f"{name}: {c.ncells} cells", flush=True)
    o = Oracle(c, serial=0)
    g = FlowStep(c)
    for _ in range(20):
        assert o.step(1)[0] == 0
        g.step(1)
    w = o.get("iwet")
    assert np.array_equal(w, g.get("iwet"))
    err = {f: np.abs(o.get(f) - g.get(f))[w > 0].max() for f in ("eta", "u", "v")}
    print(name, err, flush=True)
    assert max(err.values()) <= 1e-9, err
    g.close()
