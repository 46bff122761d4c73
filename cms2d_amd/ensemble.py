"""Batched ensemble of independent domains (front end of cms2d_ensemble_step).

Each member is a FlowStep (its own device state and stream); ``step`` advances
every member by n time steps with the members spread over host threads so
that small domains share the device.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import check
from .flow import FlowStep


class Ensemble:
    def __init__(self, cases, *, nthreads=8, **kw):
        self.members = [FlowStep(c, **kw) for c in cases]
        self.nthreads = nthreads
        self.ncells = sum(m.ncells for m in self.members)
        self._handles = (C.c_void_p * len(self.members))(*[m._h.value for m in self.members])

    def step(self, n=1):
        infos = (_lib.StepInfo * len(self.members))()
        check(_lib.lib().cms2d_ensemble_step(self._handles, len(self.members), n, self.nthreads, infos))
        return [{"niter": i.niter, "iconv": i.iconv, "kry": list(i.kry_iters), "nwet": i.nwet, "dt": i.dt}
                for i in infos]

    def close(self):
        for m in self.members:
            m.close()
        self.members = []
