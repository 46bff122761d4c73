"""Gradients through the linear solves: solve (cms2d_solve) and its reverse
(cms2d_solve_adjoint) as a torch.autograd.Function.

`solve_fn(acoef, ss, sp, phi_old, cell2cell, ncells, relax)` returns phi on
every cell (the solved active cells, the ghost values of phi_old); backward
gives dL/dacoef, dL/dss, dL/dsp and dL/dphi_old from one transposed solve.
Tensors are FP64 on the host: the two entry points take host buffers and run
on the GPU (torch is only the autograd bookkeeping here)."""
from __future__ import annotations

import numpy as np
import torch

from .flow import solve, solve_adjoint


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
