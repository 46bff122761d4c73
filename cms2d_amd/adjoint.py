"""Gradients through the implicit step as torch.autograd.Functions.

`gauge_run(mann, zb, case, nsteps, gauges)` returns the water levels at the
gauge cells after each of nsteps steps; its backward is the discrete adjoint
of the time loop (cms2d_flow_adjoint), giving dL/dn and dL/dzb per cell.
`solve_fn(acoef, ss, sp, phi_old, cell2cell, ncells, relax)` returns phi on
every cell (the solved active cells, the ghost values of phi_old); backward
gives dL/dacoef, dL/dss, dL/dsp and dL/dphi_old from one transposed solve.
Tensors are FP64 on the host: the entry points take host buffers and run on
the GPU (torch is only the autograd bookkeeping here)."""
from __future__ import annotations

import numpy as np
import torch

from .flow import FlowStep, solve, solve_adjoint


class Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, acoef, ss, sp, phi_old, cell2cell, ncells, relax, rtol, atol):
        a = acoef.detach().cpu().numpy().astype(np.float64)
        c2c = np.asarray(cell2cell, dtype=np.int32)
        ssn = ss.detach().cpu().numpy().astype(np.float64).copy()
        spn = sp.detach().cpu().numpy().astype(np.float64)
        phi = phi_old.detach().cpu().numpy().astype(np.float64).copy()
        solve(a, c2c, np.ones(len(phi), dtype=np.int32), ssn, spn, phi, relax, ncells=ncells, rtol=rtol,
              atol=atol)
        out = torch.from_numpy(phi)
        ctx.save_for_backward(acoef.detach(), sp.detach(), phi_old.detach(), out)
        ctx.c2c, ctx.ncells, ctx.relax, ctx.rtol, ctx.atol = c2c, ncells, relax, rtol, atol
        return out

    @staticmethod
    def backward(ctx, gphi):
        acoef, sp, phi_old, phi = ctx.saved_tensors
        n = ctx.ncells
        g = gphi.detach().cpu().numpy().astype(np.float64)
        gss, gsp, ga, gpo, _ = solve_adjoint(acoef.numpy(), ctx.c2c, sp.numpy(), phi_old.numpy(), phi.numpy(),
                                             ctx.relax, g[:n], ncells=n, rtol=ctx.rtol, atol=ctx.atol)
        # phi's ghost entries are phi_old's: their incoming gradient passes straight through
        gpo[n:] += g[n:]
        t = torch.from_numpy
        return t(ga), t(gss), t(gsp), t(gpo), None, None, None, None, None


def solve_fn(acoef, ss, sp, phi_old, cell2cell, ncells, relax=1.0, rtol=1e-13, atol=0.0):
    return Solve.apply(acoef, ss, sp, phi_old, cell2cell, ncells, relax, rtol, atol)


class GaugeRun(torch.autograd.Function):
    """Water levels at gauge cells over nsteps implicit steps, as a function
    of the per-cell Manning n and bed elevation; backward is one discrete
    adjoint sweep of the time loop (cms2d_flow_adjoint, checkpoints every
    `every` steps).  mann, zb: FP64 [ncells] tensors; returns eta [nsteps,
    ngauge].  `case` supplies every other input of the flow."""

    @staticmethod
    def forward(ctx, mann, zb, case, nsteps, gauges, every, flow_kw):
        c = _with_fields(case, mann, zb)
        f = FlowStep(c, **flow_kw)
        g = np.asarray(gauges, dtype=np.int64)
        eta = np.zeros((nsteps, len(g)))
        for n in range(nsteps):
            f.step(1)
            eta[n] = f.get("eta")[g]
        f.close()
        ctx.case, ctx.nsteps, ctx.gauges, ctx.every, ctx.flow_kw = c, nsteps, g, every, flow_kw
        return torch.from_numpy(eta)

    @staticmethod
    def backward(ctx, geta):
        f = FlowStep(ctx.case, **ctx.flow_kw)
        _, _, gm, gz = f.adjoint(ctx.nsteps, ctx.gauges, geta=geta.detach().cpu().numpy().astype(np.float64),
                                 every=ctx.every)
        f.close()
        return torch.from_numpy(gm), torch.from_numpy(gz), None, None, None, None, None


def _with_fields(case, mann, zb):
    import copy
    c = copy.copy(case)
    c.mann = mann.detach().cpu().numpy().astype(np.float64).copy()
    c.zb = zb.detach().cpu().numpy().astype(np.float64).copy()
    return c


def gauge_run(mann, zb, case, nsteps, gauges, every=60, **flow_kw):
    return GaugeRun.apply(mann, zb, case, nsteps, gauges, every, flow_kw)
