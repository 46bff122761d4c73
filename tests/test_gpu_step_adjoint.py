"""Discrete adjoint of the implicit time loop (cms2d_flow_adjoint): gradients
of a gauge water-level misfit with respect to the Manning n and the bed
elevation of every cell.  The reference has no adjoint; the gradients are
pinned by what fixes them independently of the code under test:

* central finite differences of the same forward (cell by cell
  and along random directions of all cells at once: the dot-product identity
  <J v, w> = <v, J^T w> with w the loss's weights),
* central finite differences of the FP64 oracle's forward (the CPU tangent),
* the checkpoint interval, which must not change the result,
* the autograd wrapper against the entry point, and torch.autograd.gradcheck.

Solves run to rtol 1e-13 here: a solve stopped at 1e-10 changes its output
by ~1e-10 relative when a perturbation changes its iteration count, which is
larger than a central difference of a 1e-10 gradient can absorb.  The loss
difference is formed as sum (e+ - e-)(e+ + e- - 2 obs), which equals L+ - L-
without the cancellation of two nearly equal sums.  Cell differences use
h = 1e-5 (0.04 % of n; 1e-5 m of bed), those along random directions
h = 1e-6; a difference whose one-sided halves disagree by more than 0.1 % (a
step across a discrete switch of the forward run) is retaken with h/10, then
10 h.  An entry far below the gradient's largest moves the gauges by a few
ulps of eta within h, so the cell checks take entries of at least a quarter
of the largest (the direction checks cover every cell)."""
import copy
import os
import sys

import numpy as np
import pytest

from cms2d_amd import cases

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))

pytestmark = pytest.mark.gpu

TOL = dict(rtol=1e-13, atol=0.0)
H = {"mann": 1e-5, "zb": 1e-5}
REL_CELL = 1e-4      # the adjoint-vs-difference bound of a gradient entry
H_DIR = 1e-6         # step along a random direction of every cell (a sum of many entries)
REL_DIR = 1e-4       # ... the bound of that directional derivative
FLOOR = 0.25         # cell checks take entries of at least this fraction of the largest


def _case(name):
    if name == "basin":
        c = cases.basin(nx=16, ny=8, dx=10.0, seed=0)
        return c, 20, np.array([3 * 16 + 5, 5 * 16 + 12]), [3 * 16 + 5, 40, 100, 7]
    c = cases.square(nx=40, ny=24, dx=20.0, seed=3)    # beach: wetting and drying
    return c, 6, np.array([10 * 40 + 20, 5 * 40 + 30, 20 * 40 + 35]), [10 * 40 + 20, 10 * 40 + 21, 300, 700]


def _gauges(c, nsteps, gauges, mann=None, zb=None):
    from cms2d_amd.flow import FlowStep
    c2 = copy.copy(c)
    if mann is not None:
        c2.mann = mann
    if zb is not None:
        c2.zb = zb
    f = FlowStep(c2, **TOL)
    eta = np.zeros((nsteps, len(gauges)))
    for n in range(nsteps):
        f.step(1)
        eta[n] = f.get("eta")[gauges]
    f.close()
    return eta


def _adjoint(c, nsteps, gauges, obs, every):
    from cms2d_amd.flow import FlowStep
    f = FlowStep(c, **TOL)
    out = f.adjoint(nsteps, gauges, obs=obs, every=every)
    f.close()
    return out


def _difference(c, nsteps, gauges, obs, eta0, field, base, direction, h0):
    """Central difference of the loss along `direction` of `field`, at the
    first step of h0, h0/10, 10 h0 whose two one-sided halves agree to 0.1 %:
    a step that crosses a discrete switch of the forward run (a clipping
    branch, a solve's iteration count) shows up as halves that disagree."""
    best = None
    for h in (h0, h0 / 10, h0 * 10):
        ep = _gauges(c, nsteps, gauges, **{field: base + h * direction})
        em = _gauges(c, nsteps, gauges, **{field: base - h * direction})
        fwd = ((ep - eta0) * (ep + eta0 - 2.0 * obs)).sum() / h
        bwd = ((eta0 - em) * (eta0 + em - 2.0 * obs)).sum() / h
        cen = 0.5 * (fwd + bwd)
        spread = abs(fwd - bwd) / max(abs(cen), 1e-300)
        if spread <= 1e-3:
            return cen
        if best is None or spread < best[0]:
            best = (spread, cen)
    return best[1]


@pytest.mark.parametrize("name", ["basin", "square"])
def test_adjoint_matches_central_differences(gpu_available, name):
    c, nsteps, gauges, cells = _case(name)
    print(f"step adjoint vs central differences, {name} {c.nx}x{c.ny}, {nsteps} steps", flush=True)
    rng = np.random.default_rng(5)
    eta0 = _gauges(c, nsteps, gauges)
    obs = eta0 + 0.01 * rng.standard_normal(eta0.shape)
    eta, loss, gm, gz = _adjoint(c, nsteps, gauges, obs, every=2)
    assert np.array_equal(eta, eta0)
    assert loss == pytest.approx(((eta0 - obs) ** 2).sum(), rel=1e-14)
    for field, grad, base in (("mann", gm, c.mann), ("zb", gz, c.zb)):
        scale = np.abs(grad).max()
        assert scale > 0.0
        # the listed cells where the entry is resolved, and the three largest
        top = [int(k) for k in np.argsort(-np.abs(grad))[:3]]
        for k in [k for k in cells if abs(grad[k]) >= FLOOR * scale] + top:
            e = np.zeros_like(base)
            e[k] = 1.0
            fd = _difference(c, nsteps, gauges, obs, eta0, field, base, e, H[field])
            print(f"  d/d{field}[{k}] adjoint {grad[k]: .9e} difference {fd: .9e}", flush=True)
            assert abs(grad[k] - fd) <= REL_CELL * abs(fd)
        v = rng.standard_normal(base.shape)
        fd = _difference(c, nsteps, gauges, obs, eta0, field, base, v, H_DIR)
        ad = float(grad @ v)
        print(f"  random direction of {field}: adjoint {ad: .9e} difference {fd: .9e}", flush=True)
        assert abs(ad - fd) <= REL_DIR * abs(fd)


def test_adjoint_checkpoint_interval_and_state(gpu_available):
    """Every checkpoint interval recomputes the same forward: the gradients
    agree to rounding (the reverse kernels' atomic sums), and the flow is
    back at its initial state after the call."""
    from cms2d_amd.flow import FlowStep
    c, nsteps, gauges, _ = _case("square")
    print("step adjoint: checkpoint intervals 1, 4, 6", flush=True)
    obs = np.zeros((nsteps, len(gauges)))
    ref = None
    for every in (1, 4, nsteps):
        eta, loss, gm, gz = _adjoint(c, nsteps, gauges, obs, every)
        if ref is None:
            ref = (eta, loss, gm, gz)
            continue
        assert np.array_equal(eta, ref[0]) and loss == ref[1]
        assert np.abs(gm - ref[2]).max() <= 1e-12 * np.abs(ref[2]).max()
        assert np.abs(gz - ref[3]).max() <= 1e-12 * np.abs(ref[3]).max()
    f = FlowStep(c, **TOL)
    f.adjoint(nsteps, gauges, obs=obs, every=2)
    f.step(2)
    g = FlowStep(c, **TOL)
    g.step(2)
    assert np.array_equal(f.get("eta"), g.get("eta")) and np.array_equal(f.get("u"), g.get("u"))
    f.close()
    g.close()


def test_adjoint_against_oracle_differences(gpu_available):
    """The FP64 CPU oracle's central differences (its forward restates the
    reference's step; ILU(0)-BiCGStab to rtol 1e-13) at the gauge cell and a
    cell beside it."""
    from pyoracle import Oracle
    c, nsteps, gauges, _ = _case("square")
    print("step adjoint vs oracle central differences, square 40x24", flush=True)
    obs = np.zeros((nsteps, len(gauges)))
    _, _, gm, gz = _adjoint(c, nsteps, gauges, obs, every=3)

    def oracle_gauges(**fields):
        c2 = copy.copy(c)
        for k, v in fields.items():
            setattr(c2, k, v)
        o = Oracle(c2, serial=0, rtol=1e-13, atol=0.0)
        eta = np.zeros((nsteps, len(gauges)))
        for n in range(nsteps):
            rc, _ = o.step(1)
            assert rc == 0
            eta[n] = o.get("eta")[gauges]
        return eta

    for field, grad, base in (("mann", gm, c.mann), ("zb", gz, c.zb)):
        h = H[field]
        for k in (gauges[0], gauges[0] + 1):
            p, m = base.copy(), base.copy()
            p[k] += h
            m[k] -= h
            ep, em = oracle_gauges(**{field: p}), oracle_gauges(**{field: m})
            fd = ((ep - em) * (ep + em)).sum() / (2 * h)
            print(f"  d/d{field}[{k}] adjoint {grad[k]: .9e} oracle difference {fd: .9e}", flush=True)
            assert abs(grad[k] - fd) <= REL_CELL * abs(fd)


def test_gauge_run_autograd(gpu_available):
    """gauge_run's backward is the entry point's gradient; gradcheck (one
    random direction) on a 6x4 basin over two steps."""
    import torch
    from cms2d_amd.adjoint import gauge_run
    c, nsteps, gauges, _ = _case("basin")
    print("gauge_run autograd", flush=True)
    mann = torch.tensor(c.mann, dtype=torch.float64, requires_grad=True)
    zb = torch.tensor(c.zb, dtype=torch.float64, requires_grad=True)
    eta = gauge_run(mann, zb, c, nsteps, gauges, every=2, **TOL)
    obs = torch.zeros_like(eta)
    loss = ((eta - obs) ** 2).sum()
    loss.backward()
    _, L, gm, gz = _adjoint(c, nsteps, gauges, obs.numpy(), every=2)
    assert loss.item() == pytest.approx(L, rel=1e-14)
    assert np.abs(mann.grad.numpy() - gm).max() <= 1e-12 * np.abs(gm).max()
    assert np.abs(zb.grad.numpy() - gz).max() <= 1e-12 * np.abs(gz).max()
    small = cases.basin(nx=6, ny=4, dx=10.0, seed=1)
    m = torch.tensor(small.mann, dtype=torch.float64, requires_grad=True)
    z = torch.tensor(small.zb, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: gauge_run(a, b, small, 2, [9, 14], every=1, **TOL), (m, z),
                                    eps=1e-6, atol=1e-8, rtol=1e-4, fast_mode=True)


def test_adjoint_rejects_unsupported(gpu_available):
    from cms2d_amd import _lib
    from cms2d_amd.flow import FlowStep
    c, nsteps, gauges, _ = _case("basin")
    obs = np.zeros((nsteps, len(gauges)))
    f = FlowStep(c)
    f.step(1)
    with pytest.raises(RuntimeError, match="initial state"):
        f.adjoint(nsteps, gauges, obs=obs)
    f.close()
    f = FlowStep(c, precision=_lib.PREC_FP32_REFINE)
    with pytest.raises(RuntimeError, match="FP64 form"):
        f.adjoint(nsteps, gauges, obs=obs)
    f.close()
    c5 = copy.copy(c)
    c5.ndsch = 5
    f = FlowStep(c5)
    with pytest.raises(RuntimeError, match="ndsch"):
        f.adjoint(nsteps, gauges, obs=obs)
    f.close()
    f = FlowStep(c)
    with pytest.raises(RuntimeError, match="gauge cell"):
        f.adjoint(nsteps, [c.ncells], obs=np.zeros((nsteps, 1)))
    f.close()
