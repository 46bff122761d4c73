"""Variable time step (VARIABLE_TIMESTEP, source/flow/flow.F90:107): a
divergent outer iteration sends the step back to the previous time level with
half the time step (flow_recalculate, source/flow/flow_imp.F90:1133-1242) and
timing doubles it back once the pp residual is small (source/main/
CMS_Flow.f90:180-205).  The case is a 2 m dam break in a closed basin at
dt = 300 s, which diverges at the full step; set() makes it a consistent
initial condition (both time levels, face fluxes)."""
import numpy as np
import pytest

from cms2d_amd import _lib, cases
from cms2d_amd.flow import FlowStep
from pyoracle import Oracle

TOL_ETA = 1e-6   # m
TOL_VEL = 1e-6   # m/s


def dam_break(dt=300.0, jump=2.0):
    c = cases.basin(nx=48, ny=32, dx=10.0, seed=0, dt=dt)
    eta = np.zeros((c.ny, c.nx))
    eta[:, : c.nx // 3] = jump
    return c, 9.81 * eta.ravel()


def test_oracle_recovers_with_variable_step():
    c, p = dam_break()
    o = Oracle(c, serial=0)
    o.set("p", p)
    rc, _ = o.step(1)
    assert rc != 0                                   # divergent at the full step
    o = Oracle(c, serial=0, dtvar=1)
    o.set("p", p)
    times = []
    for _ in range(12):
        rc, info = o.step(1)
        assert rc == 0
        times.append(o.time)
    dts = np.diff([0.0] + times)
    assert dts.min() < 300.0                         # the step was cut ...
    assert np.all(np.isin(dts, [300.0 / 2 ** k for k in range(9)]))   # ... by powers of two
    assert dts[-1] > dts.min()                       # ... and grew back


@pytest.mark.gpu
@pytest.mark.parametrize("curve", [False, True], ids=["tide", "wse_curve"])
def test_gpu_variable_step_matches_oracle(gpu_available, curve):
    """With a water-level curve on the open side (quadratic interpolation over
    uneven times), redone steps move the time back inside the curve's windows,
    so the window search state is exercised too."""
    c, p = dam_break()
    if curve:
        c.wse_times = [0.0, 0.05, 0.1, 0.2, 0.3, 0.5, 1.0]
        c.wse_values = [0.0, 0.2, -0.1, 0.3, 0.1, 0.4, 0.2]
    print(f"dam break 48x32 dt 300 curve={curve}", flush=True)
    g = FlowStep(c)
    g.set("p", p)
    with pytest.raises(_lib.Cms2dError) as e:
        g.step(1)
    assert e.value.code == -3                        # CMS2D_ERR_DIVERGED
    g.close()
    o = Oracle(c, serial=0, dtvar=1)
    o.set("p", p)
    g = FlowStep(c, dtvar=1)
    g.set("p", p)
    assert np.array_equal(o.get("iwet"), g.get("iwet"))
    for s in range(12):
        rco, io = o.step(1)
        ig = g.step(1)
        assert rco == 0
        assert abs(o.time - g.time) < 1e-9, (s, o.time, g.time)
        assert io["niter"] == ig["niter"], (s, io, ig)
    wet = o.get("iwet") > 0
    assert np.array_equal(o.get("iwet"), g.get("iwet"))
    for name, tol in (("eta", TOL_ETA), ("u", TOL_VEL), ("v", TOL_VEL)):
        err = np.abs(o.get(name) - g.get(name))[wet].max()
        print(name, err)
        assert err <= tol, (name, err)


@pytest.mark.gpu
def test_gpu_initial_condition_matches_oracle(gpu_available):
    """set() as an initial condition (a small dam break that converges at dt)."""
    c, p = dam_break(dt=60.0, jump=0.3)
    u = np.full(c.nx * c.ny, 0.05)
    o = Oracle(c, serial=0)
    g = FlowStep(c)
    for f in (o, g):
        f.set("p", p)
        f.set("u", u)
    for _ in range(5):
        assert o.step(1)[0] == 0
        g.step(1)
    wet = o.get("iwet") > 0
    assert np.array_equal(o.get("iwet"), g.get("iwet"))
    for name, tol in (("eta", TOL_ETA), ("u", TOL_VEL), ("v", TOL_VEL)):
        err = np.abs(o.get(name) - g.get(name))[wet].max()
        assert err <= tol, (name, err)
