"""Python front end of the MI355X implicit step (thin layer over the C ABI).

Mirrors the reference's entry points: ``FlowStep.step`` advances the time loop
the way source/main/CMS_Flow.f90:43-87 does (timing, flow_imp, flow_update),
and ``solve`` is the reference's solve(acoef,ss,sp,res,phi,n)
(source/solvers/solvers.f90:205-304) on the general cell list.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import D, I, check

SIDE = {"N": 0, "E": 1, "S": 2, "W": 3}


def _dp(a):
    return a.ctypes.data_as(D)


class FlowStep:
    """Device-resident implicit flow model on a uniform Cartesian grid, or on a
    telescoping grid when the case carries one (case.tel)."""

    def __init__(self, case, *, rtol=1e-10, atol=1e-12, kmax=4000, maxit=0,
                 precision=_lib.PREC_FP64, device=0, precond=None, solver=None, dtvar=0,
                 serial_order=0):
        L = _lib.lib()
        self.case = case
        keep = []

        def arr(a):
            if a is None:
                return D()
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return _dp(a)

        d = _lib.FlowDesc()
        d.nx, d.ny, d.dx, d.dy = case.nx, case.ny, case.dx, case.dy
        d.zb, d.mann = arr(case.zb), arr(case.mann)
        d.wavestrx = arr(case.wavestrx)
        d.wavestry = arr(case.wavestry)
        for k in range(4):
            d.side_type[k] = _lib.BND_WALL
        if case.th_side >= 0:
            if case.wse_times is not None:
                d.side_type[case.th_side] = _lib.BND_WSE_CURVE
                d.wse_times, d.wse_values = arr(case.wse_times), arr(case.wse_values)
                d.wse_ntimes, d.wse_order = len(case.wse_times), case.wse_order
                d.wse_adjust_off = int(not case.wse_adjust)
                d.wse_gradx, d.wse_grady = case.wse_grad
            else:
                d.side_type[case.th_side] = _lib.BND_WSE_HARMONIC
        if getattr(case, "q_side", -1) >= 0:
            # the flux (river) string: total inflow or its curve, conveyance
            d.side_type[case.q_side] = _lib.BND_FLUX
            d.flux_q, d.flux_order, d.flux_conveyance = case.q_const, case.q_order, case.q_cmvel
            if getattr(case, "q_angle", None) is not None:
                d.flux_angle_on, d.flux_angle = 1, case.q_angle
            d.flux_rating = int(getattr(case, "q_rating", False))
            if case.q_times is not None:
                d.flux_times, d.flux_values = arr(case.q_times), arr(case.q_values)
                d.flux_ntimes = len(case.q_times)
        d.ntc = len(case.amp)
        for k in range(d.ntc):
            d.amp[k], d.speed[k], d.phase[k] = case.amp[k], case.speed_rad_hr[k], case.phase_rad[k]
        d.dt, d.rampdur, d.hdry = case.dt, case.rampdur, float(np.float32(case.hdry))
        d.noptset, d.wallfric, d.ndsch, d.maxit = case.noptset, case.wallfric, case.ndsch, maxit
        d.rtol, d.atol, d.kmax, d.precision, d.device = rtol, atol, kmax, precision, device
        d.precond = _lib.PRECOND_MG if precond is None else precond
        d.solver = _lib.SOLVER_BICGSTAB if solver is None else solver
        d.dtvar = dtvar
        d.serial_order = serial_order
        d.latitude, d.coriolis_plane = case.latitude, case.coriolis_plane
        d.turbulence = {5: 0, 0: 1, 1: 2, 2: 3, 3: 4}[case.mturbul]   # CMS2D_TURB_* from mturbul
        d.friction = {2: 0, 1: 1, 3: 2, 4: 3}[case.mbedfric]           # CMS2D_FRIC_* from mbedfric
        h = C.c_void_p()
        tel = getattr(case, "tel", None)
        self.nf = 4
        if tel is not None:
            # a telescoping grid: the .tel arrays and the water-level cellstring
            geom = np.ascontiguousarray(tel.geom, dtype=np.float64)
            nb = np.ascontiguousarray(tel.nb, dtype=np.int32)
            bnd = np.ascontiguousarray(tel.bnd_cells, dtype=np.int32)
            keep += [geom, nb, bnd]
            tg = _lib.TelGrid()
            tg.ncells, tg.geom, tg.nb = len(geom), _dp(geom), nb.ctypes.data_as(I)
            tg.bnd_cells, tg.nbnd = bnd.ctypes.data_as(I), len(bnd)
            if getattr(case, "q_side", -1) >= 0:
                qc = np.ascontiguousarray(tel.flux_cells, dtype=np.int32)
                keep.append(qc)
                tg.flux_cells, tg.nflux = qc.ctypes.data_as(I), len(qc)
            check(L.cms2d_flow_create_tel(C.byref(d), C.byref(tg), C.byref(h)))
            self.nf = 6
        else:
            check(L.cms2d_flow_create(C.byref(d), C.byref(h)))
        self._h = h
        if case.wind_speed:
            self.set_wind(*case.wind_vector())
        self.ncells = case.ncells

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().cms2d_flow_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        info = _lib.StepInfo()
        check(_lib.lib().cms2d_flow_step(self._h, n, C.byref(info)))
        return {"niter": info.niter, "iconv": info.iconv, "rmom": list(info.rmom),
                "kry": list(info.kry_iters), "nwet": info.nwet, "kernel_ms": info.kernel_ms,
                "dt": info.dt}

    def balance(self):
        """flow_balance and flow_step_stat of the last step (cms2d_flow_balance):
        the step's and the cumulative water volumes [m^3] (storage, efflux
        through the water-level string, flux through every cell face, balance,
        net, percentage errors) and the extreme water level, u, v and Courant
        number over the wet cells with their cells."""
        b = _lib.Balance()
        check(_lib.lib().cms2d_flow_balance(self._h, C.byref(b)))
        return {name: getattr(b, name) for name, _ in _lib.Balance._fields_}

    def statistics(self, t_start, t_end):
        """Flow statistics over windows [t_start, t_end] in seconds
        (cms2d_flow_statistics, the FLOW_STATISTICS card)."""
        check(_lib.lib().cms2d_flow_statistics(self._h, float(t_start), float(t_end)))

    def stat(self, name):
        """(accumulator of `name` per cell, steps counted in the open window)."""
        out, n = np.empty(self.ncells), C.c_int()
        check(_lib.lib().cms2d_flow_stat(self._h, name.encode(), _dp(out), C.byref(n)))
        return out, n.value

    def get(self, name):
        out = np.empty(self.ncells)
        check(_lib.lib().cms2d_flow_get(self._h, name.encode(), _dp(out)))
        return out

    def set(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        check(_lib.lib().cms2d_flow_set(self._h, name.encode(), _dp(a)))

    def get_face(self, name):
        out = np.empty(self.nf * self.ncells)
        check(_lib.lib().cms2d_flow_get_face(self._h, name.encode(), _dp(out)))
        return out.reshape(self.ncells, self.nf)

    @property
    def time(self):
        return _lib.lib().cms2d_flow_time(self._h)

    def set_waves(self, whgt, wper, wlen, wunitx, wunity):
        """Radiation-stress forcing from wave fields on the flow grid
        (cms2d_flow_set_waves); arrays over the active cells, row-major."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (whgt, wper, wlen, wunitx, wunity)]
        for a in arrs:
            assert a.size == self.ncells
        check(_lib.lib().cms2d_flow_set_waves(self._h, *[_dp(a) for a in arrs]))

    def set_wind(self, wndx, wndy):
        """Spatially constant 10-m wind vector [m/s] in grid coordinates
        (cms2d_flow_set_wind)."""
        check(_lib.lib().cms2d_flow_set_wind(self._h, float(wndx), float(wndy)))

    def checkpoint(self):
        """Exact checkpoint of the device state (bytes), cms2d_flow_checkpoint."""
        n = C.c_size_t(0)
        check(_lib.lib().cms2d_flow_checkpoint(self._h, None, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(_lib.lib().cms2d_flow_checkpoint(self._h, buf, C.byref(n)))
        return buf.raw[:n.value]

    def restore(self, data):
        """Continues from a checkpoint of a flow on the same grid."""
        buf = C.create_string_buffer(bytes(data), len(data))
        check(_lib.lib().cms2d_flow_restore(self._h, buf, len(data)))

    def tuning(self):
        """(level-0 multigrid tile width, graph replay on) picked at creation."""
        t, gr = C.c_int(), C.c_int()
        check(_lib.lib().cms2d_flow_tuning(self._h, C.byref(t), C.byref(gr)))
        return t.value, bool(gr.value)

    def adjoint(self, nsteps, gauges, *, obs=None, geta=None, every=60):
        """Discrete adjoint of nsteps steps from the initial state
        (cms2d_flow_adjoint): gauges are row-major cell indices; with obs
        [nsteps, ngauge] the loss is sum (eta - obs)^2, else geta [nsteps,
        ngauge] is dL/deta.  Returns (eta [nsteps, ngauge], loss or None,
        dL/dn [ncells], dL/dzb [ncells]); the flow is back at its initial
        state afterwards."""
        cells = np.ascontiguousarray(gauges, dtype=np.int32)
        ng = len(cells)
        eta = np.zeros((nsteps, ng))
        gm, gz = np.zeros(self.ncells), np.zeros(self.ncells)
        loss = C.c_double(0.0)
        o = None if obs is None else np.ascontiguousarray(obs, dtype=np.float64).reshape(nsteps, ng)
        ge = None if geta is None else np.ascontiguousarray(geta, dtype=np.float64).reshape(nsteps, ng)
        check(_lib.lib().cms2d_flow_adjoint(self._h, nsteps, every, ng, cells.ctypes.data_as(I),
                                            D() if o is None else _dp(o), D() if ge is None else _dp(ge),
                                            _dp(eta), C.cast(C.byref(loss), D), _dp(gm), _dp(gz)))
        return eta, (loss.value if o is not None else None), gm, gz

    def kernel_stats(self, name="spmv_pp"):
        ms, n = C.c_double(), C.c_long()
        check(_lib.lib().cms2d_flow_kernel_stats(self._h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value


def solve(acoef, cell2cell, iwet, ss, sp, phi, relax, *, ncells, rtol=1e-10, atol=1e-12,
          kmax=4000, noalloc=False):
    """solve(acoef,ss,sp,res,phi,n) on the reference's cell list.

    acoef, cell2cell: [ncellsD, nmaxfaces] (4 faces N,E,S,W on Cartesian grids,
    up to 6 in any order on telescoping grids), cell2cell 0-based with -1 for
    no face; iwet, ss, sp, phi: [ncellsD].  ss and phi are updated in place.
    noalloc: through cms2d_solve_noalloc (no device allocation; systems up to
    its static scratch size).  Returns (res, rmom, iterations)."""
    acoef = np.ascontiguousarray(acoef, dtype=np.float64)
    c2c = np.ascontiguousarray(cell2cell, dtype=np.int32)
    iw = np.ascontiguousarray(iwet, dtype=np.int32)
    assert ss.dtype == np.float64 and ss.flags.c_contiguous
    assert phi.dtype == np.float64 and phi.flags.c_contiguous
    spa = np.ascontiguousarray(sp, dtype=np.float64)
    ncellsD = acoef.shape[0]
    assert c2c.shape == acoef.shape
    res = np.zeros(ncellsD)
    rmom = C.c_double()
    iters = C.c_int()
    fn = _lib.lib().cms2d_solve_noalloc if noalloc else _lib.lib().cms2d_solve
    check(fn(ncells, ncellsD, acoef.shape[1], _dp(acoef), c2c.ctypes.data_as(I),
             iw.ctypes.data_as(I), _dp(ss), _dp(spa), _dp(res), _dp(phi),
             relax, rtol, atol, kmax, C.byref(rmom), C.byref(iters)))
    return res, rmom.value, iters.value


def solve_adjoint(acoef, cell2cell, sp, phi_old, phi, relax, gphi, *, ncells, rtol=1e-10, atol=1e-12,
                  kmax=4000):
    """The reverse of solve (cms2d_solve_adjoint): for phi = solve(acoef, ss,
    sp, phi_old, relax) and gphi = dL/dphi on the active cells, returns
    (dL/dss, dL/dsp, dL/dacoef, dL/dphi_old, iterations of the transposed
    solve), shaped like the inputs."""
    acoef = np.ascontiguousarray(acoef, dtype=np.float64)
    c2c = np.ascontiguousarray(cell2cell, dtype=np.int32)
    ncellsD, nf = acoef.shape
    assert c2c.shape == acoef.shape
    spa, po, ph = (np.ascontiguousarray(v, dtype=np.float64) for v in (sp, phi_old, phi))
    gp = np.ascontiguousarray(gphi, dtype=np.float64)
    assert spa.shape == po.shape == ph.shape == (ncellsD,) and gp.shape[0] >= ncells
    gss, gsp, gpo = np.zeros(ncellsD), np.zeros(ncellsD), np.zeros(ncellsD)
    ga = np.zeros((ncellsD, nf))
    iters = C.c_int()
    check(_lib.lib().cms2d_solve_adjoint(ncells, ncellsD, nf, _dp(acoef), c2c.ctypes.data_as(I), _dp(spa),
                                         _dp(po), _dp(ph), relax, _dp(gp), rtol, atol, kmax, _dp(gss), _dp(gsp),
                                         _dp(ga), _dp(gpo), C.byref(iters)))
    return gss, gsp, ga, gpo, iters.value
