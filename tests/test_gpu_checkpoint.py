"""Exact checkpoint/resume (cms2d_flow_checkpoint / cms2d_flow_restore): a flow
restored into a fresh flow object continues bitwise as the original, through
wetting and drying, the FP32 form and the variable time step (whose control
state lives on the host)."""
import numpy as np
import pytest

from cms2d_amd import _lib, cases
from cms2d_amd.flow import FlowStep

FIELDS = ("eta", "u", "v", "p", "h", "iwet", "pp", "vis")


def dam_break():
    c = cases.basin(nx=48, ny=32, dx=10.0, seed=0, dt=300.0)
    eta = np.zeros((c.ny, c.nx))
    eta[:, : c.nx // 3] = 2.0
    return c, 9.81 * eta.ravel()


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "fp32", "dtvar", "dtvar_curve"])
def test_resume_is_bitwise(gpu_available, form):
    kw = {"precision": 1} if form == "fp32" else {"dtvar": 1} if form.startswith("dtvar") else {}
    if form.startswith("dtvar"):
        c, p0 = dam_break()
        if form == "dtvar_curve":       # water-level curve on the open side (its window state resumes too)
            c.wse_times = [0.0, 0.05, 0.1, 0.2, 0.3, 0.5, 1.0]
            c.wse_values = [0.0, 0.2, -0.1, 0.3, 0.1, 0.4, 0.2]
    else:
        c, p0 = cases.tidal(nx=96, ny=64, dx=20.0, seed=9), None
    print(f"checkpoint {form} {c.nx}x{c.ny}", flush=True)
    a = FlowStep(c, **kw)
    if p0 is not None:
        a.set("p", p0)
    for _ in range(6):
        a.step(1)
    ck = a.checkpoint()
    ia = [a.step(1) for _ in range(6)]
    b = FlowStep(c, **kw)
    b.restore(ck)
    ib = [b.step(1) for _ in range(6)]
    for x, y in zip(ia, ib):
        assert (x["niter"], x["kry"], x["dt"]) == (y["niter"], y["kry"], y["dt"])
    assert a.time == b.time
    for name in FIELDS:
        assert np.array_equal(a.get(name), b.get(name)), name
    if form.startswith("dtvar"):
        assert len({i["dt"] for i in ia}) >= 1 and min(i["dt"] for i in ia) < c.dt   # the control state was live
    a.close()
    b.close()


@pytest.mark.gpu
def test_restore_rejects_other_grids(gpu_available):
    a = FlowStep(cases.tidal(nx=48, ny=32, dx=20.0, seed=9))
    ck = a.checkpoint()
    b = FlowStep(cases.tidal(nx=40, ny=32, dx=20.0, seed=9))
    for bad in (ck, ck[:100], b"x" * len(ck)):
        with pytest.raises(_lib.Cms2dError) as e:
            b.restore(bad)
        assert e.value.code == -1                      # CMS2D_ERR_ARG
    a.restore(ck)                                      # its own checkpoint is accepted
    a.close()
    b.close()
