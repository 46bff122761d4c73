"""ctypes front end of the FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by cms2d_amd/.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)


class CmsCase(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
        ("zb", D), ("mann", D), ("wavestrx", D), ("wavestry", D),
        ("th_side", C.c_int), ("ntc", C.c_int),
        ("amp", C.c_double * 8), ("speed", C.c_double * 8), ("phase", C.c_double * 8),
        ("dt", C.c_double), ("rampdur", C.c_double), ("hdry", C.c_double),
        ("noptset", C.c_int), ("wallfric", C.c_int), ("ndsch", C.c_int), ("maxit", C.c_int),
        ("solver_mode", C.c_int), ("rtol", C.c_double), ("atol", C.c_double), ("kmax", C.c_int),
        ("serial", C.c_int), ("dtvar", C.c_int), ("latitude", C.c_double), ("coriolis_plane", C.c_int),
        ("mturbul", C.c_int), ("mbedfric", C.c_int),
        ("wse_times", D), ("wse_values", D), ("wse_ntimes", C.c_int), ("wse_order", C.c_int),
        ("wse_adjust_off", C.c_int), ("wse_gradx", C.c_double), ("wse_grady", C.c_double),
        ("tel_n", C.c_int), ("tel_geom", D), ("tel_nb", I), ("tel_bnd", I), ("tel_nbnd", C.c_int),
        ("noskew", C.c_int),
        ("q_side", C.c_int), ("q_const", C.c_double), ("q_times", D), ("q_values", D),
        ("q_ntimes", C.c_int), ("q_order", C.c_int), ("q_cmvel", C.c_double),
        ("q_angle_on", C.c_int), ("q_angle", C.c_double), ("q_rating", C.c_int),
        ("tel_qcells", I), ("tel_nq", C.c_int),
    ]


class StepInfo(C.Structure):
    _fields_ = [("niter", C.c_int), ("iconv", C.c_int), ("rmom", C.c_double * 3),
                ("kry_iters", C.c_int * 3), ("nwet", C.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.cms_oracle_create.restype = C.c_void_p
        L.cms_oracle_create.argtypes = [C.POINTER(CmsCase)]
        L.cms_oracle_destroy.argtypes = [C.c_void_p]
        L.cms_oracle_step.argtypes = [C.c_void_p, C.c_int, C.POINTER(StepInfo)]
        L.cms_oracle_get.argtypes = [C.c_void_p, C.c_char_p, D]
        L.cms_oracle_set.argtypes = [C.c_void_p, C.c_char_p, D]
        L.cms_oracle_get_face.argtypes = [C.c_void_p, C.c_char_p, D]
        L.cms_oracle_ncells.argtypes = [C.c_void_p]
        L.cms_oracle_ncellsD.argtypes = [C.c_void_p]
        L.cms_oracle_nf.argtypes = [C.c_void_p]
        L.cms_oracle_topology.argtypes = [C.c_void_p, I, I, I, I]
        L.cms_oracle_time.restype = C.c_double
        L.cms_oracle_time.argtypes = [C.c_void_p]
        L.cms_oracle_volume.argtypes = [C.c_void_p, D, D]
        for fn, na in (("cms_oracle_schmcoef", 3), ("cms_oracle_wall_coef", 2),
                       ("cms_oracle_man2drag", 4), ("cms_oracle_drag2length", 2),
                       ("cms_oracle_drag2man", 4), ("cms_oracle_length2drag", 2)):
            f = getattr(L, fn)
            f.restype = C.c_double
            f.argtypes = ([C.c_int] if fn == "cms_oracle_schmcoef" else []) + \
                [C.c_double] * (na - (1 if fn == "cms_oracle_schmcoef" else 0))
        L.cms_oracle_plagr.restype = C.c_double
        L.cms_oracle_plagr.argtypes = [C.c_int, D, D, C.c_int, C.c_double, C.POINTER(C.c_int)]
        L.cms_oracle_grad.argtypes = [C.c_void_p, D, D, D]
        L.cms_oracle_stats.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.cms_oracle_stat_get.argtypes = [C.c_void_p, C.c_char_p, D, C.POINTER(C.c_int)]
        L.cms_oracle_set_waves.argtypes = [C.c_void_p, D, D, D, D, D]
        L.cms_oracle_set_wind.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.cms_oracle_wind_drag_hsu.restype = C.c_double
        L.cms_oracle_wind_drag_hsu.argtypes = [C.c_double]
        L.cms_oracle_nvd_gamma.restype = C.c_double
        L.cms_oracle_nvd_gamma.argtypes = [C.c_int, C.c_double]
        L.cms_oracle_defcor.argtypes = [C.c_void_p, D, D]
        L.cms_oracle_solve.restype = C.c_int
        L.cms_oracle_solve.argtypes = [C.c_int, C.c_int, C.c_int, D, I, I, D, D, D, D, C.c_double,
                                       C.c_double, C.c_double, C.c_int, D]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(D)


class Oracle:
    """FP64 restatement run on a Case (see cms2d_amd.cases)."""

    def __init__(self, case, *, serial=0, solver_mode=0, rtol=1e-10, atol=1e-12, kmax=4000,
                 maxit=0, dtvar=0):
        L = lib()
        self.case = case
        self._keep = []

        def arr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return _dp(a)

        c = CmsCase()
        c.nx, c.ny, c.dx, c.dy = case.nx, case.ny, case.dx, case.dy
        c.zb, c.mann = arr(case.zb), arr(case.mann)
        c.wavestrx = arr(case.wavestrx) if case.wavestrx is not None else D()
        c.wavestry = arr(case.wavestry) if case.wavestry is not None else D()
        c.th_side = case.th_side
        c.ntc = len(case.amp)
        for k in range(c.ntc):
            c.amp[k], c.speed[k], c.phase[k] = case.amp[k], case.speed_rad_hr[k], case.phase_rad[k]
        c.dt, c.rampdur, c.hdry = case.dt, case.rampdur, float(np.float32(case.hdry))
        c.noptset, c.wallfric, c.ndsch = case.noptset, case.wallfric, case.ndsch
        c.maxit, c.solver_mode, c.rtol, c.atol, c.kmax = maxit, solver_mode, rtol, atol, kmax
        c.serial = serial
        c.dtvar = dtvar
        c.latitude, c.coriolis_plane = case.latitude, case.coriolis_plane
        c.mturbul = case.mturbul
        c.mbedfric = getattr(case, "mbedfric", 2)
        if getattr(case, "wse_times", None) is not None:
            c.wse_times, c.wse_values = arr(case.wse_times), arr(case.wse_values)
            c.wse_ntimes, c.wse_order = len(case.wse_times), case.wse_order
            c.wse_adjust_off = int(not getattr(case, "wse_adjust", True))
            c.wse_gradx, c.wse_grady = getattr(case, "wse_grad", (0.0, 0.0))
        tel = getattr(case, "tel", None)
        if tel is not None:
            if case.ndsch >= 5:
                raise ValueError("NVD schemes are not restated on telescoping grids")
            geom = np.ascontiguousarray(tel.geom, dtype=np.float64)
            nb = np.ascontiguousarray(tel.nb, dtype=np.int32)
            bnd = np.ascontiguousarray(tel.bnd_cells, dtype=np.int32)
            self._keep += [geom, nb, bnd]
            c.tel_n, c.tel_geom, c.tel_nb = len(geom), _dp(geom), nb.ctypes.data_as(I)
            c.tel_bnd, c.tel_nbnd = bnd.ctypes.data_as(I), len(bnd)
        c.noskew = int(getattr(case, "noskew", 0))
        c.q_side = getattr(case, "q_side", -1)
        if c.q_side >= 0:
            if tel is not None:
                qc = np.ascontiguousarray(tel.flux_cells, dtype=np.int32)
                self._keep.append(qc)
                c.tel_qcells, c.tel_nq = qc.ctypes.data_as(I), len(qc)
            c.q_const = case.q_const
            if case.q_times is not None:
                c.q_times, c.q_values = arr(case.q_times), arr(case.q_values)
                c.q_ntimes, c.q_order = len(case.q_times), case.q_order
            c.q_cmvel = case.q_cmvel
            if getattr(case, "q_angle", None) is not None:
                c.q_angle_on, c.q_angle = 1, case.q_angle
            c.q_rating = int(getattr(case, "q_rating", False))
        self._c = c
        self.h = L.cms_oracle_create(C.byref(c))
        self.ncells = L.cms_oracle_ncells(self.h)
        self.ncellsD = L.cms_oracle_ncellsD(self.h)
        self.nf = L.cms_oracle_nf(self.h)
        if getattr(case, "wind_speed", 0.0):
            self.set_wind(*case.wind_vector())

    def __del__(self):
        if getattr(self, "h", None):
            lib().cms_oracle_destroy(self.h)
            self.h = None

    def step(self, n=1):
        info = StepInfo()
        rc = lib().cms_oracle_step(self.h, n, C.byref(info))
        return rc, {"niter": info.niter, "iconv": info.iconv, "rmom": list(info.rmom),
                    "kry": list(info.kry_iters), "nwet": info.nwet}

    def get(self, name):
        out = np.empty(self.ncells)
        if lib().cms_oracle_get(self.h, name.encode(), _dp(out)) != 0:
            raise KeyError(name)
        return out

    def set(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        if lib().cms_oracle_set(self.h, name.encode(), _dp(a)) != 0:
            raise KeyError(name)

    def get_face(self, name):
        out = np.empty(self.nf * self.ncells)
        if lib().cms_oracle_get_face(self.h, name.encode(), _dp(out)) != 0:
            raise KeyError(name)
        return out.reshape(self.ncells, self.nf)

    def stats(self, t1, t2):
        """Flow statistics over windows [t1, t2] in seconds (FLOW_STATISTICS)."""
        if lib().cms_oracle_stats(self.h, float(t1), float(t2)) != 0:
            raise ValueError("statistics window")

    def stat(self, name):
        """(accumulator of `name` per cell, steps counted in the open window)."""
        out, n = np.empty(self.ncells), C.c_int()
        if lib().cms_oracle_stat_get(self.h, name.encode(), _dp(out), C.byref(n)) != 0:
            raise KeyError(name)
        return out, n.value

    def topology(self):
        """(ncface [n], cell2cell [nD, nf], idirface [nD, nf], isjoint [n])."""
        n, nd, nf = self.ncells, self.ncellsD, self.nf
        ncf, c2c = np.empty(n, np.int32), np.empty(nd * nf, np.int32)
        idr, jt = np.empty(nd * nf, np.int32), np.empty(n, np.int32)
        lib().cms_oracle_topology(self.h, ncf.ctypes.data_as(I), c2c.ctypes.data_as(I),
                                  idr.ctypes.data_as(I), jt.ctypes.data_as(I))
        return ncf, c2c.reshape(nd, nf), idr.reshape(nd, nf), jt

    def grad(self, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        gx, gy = np.empty(self.ncells), np.empty(self.ncells)
        lib().cms_oracle_grad(self.h, _dp(a), _dp(gx), _dp(gy))
        return gx, gy

    def set_waves(self, whgt, wper, wlen, wunitx, wunity):
        """Radiation-stress forcing from wave fields (active cells); False on
        the reference's error (NaN or a component above 10)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (whgt, wper, wlen, wunitx, wunity)]
        return lib().cms_oracle_set_waves(self.h, *[_dp(a) for a in arrs]) == 0

    def set_wind(self, wndx, wndy):
        """Spatially constant 10-m wind vector [m/s] in grid coordinates."""
        lib().cms_oracle_set_wind(self.h, float(wndx), float(wndy))

    def defcor(self):
        """Deferred-correction sums (dsu, dsv) of the current state."""
        du, dv = np.empty(self.ncells), np.empty(self.ncells)
        lib().cms_oracle_defcor(self.h, _dp(du), _dp(dv))
        return du, dv

    def volume(self):
        v, b = C.c_double(), C.c_double()
        lib().cms_oracle_volume(self.h, C.byref(v), C.byref(b))
        return v.value, b.value

    @property
    def time(self):
        return lib().cms_oracle_time(self.h)


GRAV = 9.81
GRAVINV = 1.0 / 9.81      # flow_def gravinv


class VolumeBalance:
    """flow_balance (source/flow/flow_imp.F90:2260-2372; ntsch 1, no rain or
    evaporation) and flow_step_stat (source/flow/flow.F90:1912-2008) restated
    over the exported fields of a model (this oracle or the HIP flow: get and
    get_face with the same names).  Call before() ahead of a step and after()
    behind it.  Volumes in m^3; extremes over the wet cells, cells 0-based;
    flow_step_stat keeps the first cell whose key is strictly larger than the
    best so far, from key 0 at the first cell."""

    def __init__(self, case):
        tel = getattr(case, "tel", None)
        self.dt = case.dt
        if tel is not None:
            g = np.asarray(tel.geom)
            self.area, self.dx, self.dy = g[:, 2] * g[:, 3], g[:, 2], g[:, 3]
            # the strings' faces: the domain-edge faces of their cells (the
            # others among them are walls with zero flux)
            cells = list(np.asarray(tel.bnd_cells))
            if getattr(case, "q_side", -1) >= 0:
                cells += [c for c in np.asarray(tel.flux_cells) if c not in set(cells)]
            edge = tel.edge_faces()
            self.str_faces = [(c, k) for c in cells for k in np.nonzero(edge[c])[0]]
        else:
            n = case.nx * case.ny
            self.area = np.full(n, case.dx * case.dy)
            self.dx, self.dy = np.full(n, case.dx), np.full(n, case.dy)
            # every forcing string: the flux side and the water-level side
            self.str_faces = []
            j, i = np.mgrid[0:case.ny, 0:case.nx]
            for side in (getattr(case, "q_side", -1), case.th_side):     # N, E, S, W; -1: none
                if side >= 0:
                    on = [j == case.ny - 1, i == case.nx - 1, j == 0, i == 0][side]
                    self.str_faces += [(c, side) for c in np.nonzero(on.ravel())[0]]
        self.cum = dict(storage=0.0, efflux=0.0, flux=0.0, norm=0.0)
        self.h1 = None

    def before(self, model):
        self.h1 = model.get("h")

    def after(self, model, dt=None):
        """The step's values and the running totals; dt: the step's time step
        when it varies (VARIABLE_TIMESTEP)."""
        dt = self.dt if dt is None else dt
        h, f = model.get("h"), model.get_face("flux")
        stg = float((self.area * (h - self.h1)).sum())
        flux = float((dt * f.sum(axis=1)).sum())
        bnd = float(sum(dt * f[c, k] for c, k in self.str_faces))
        c = self.cum
        c["storage"] += stg
        c["flux"] += flux
        c["efflux"] += bnd
        c["norm"] = max(c["norm"], abs(c["storage"]))
        out = dict(dt=dt, storage=stg, efflux=bnd, flux=flux, balance=stg + bnd, net=stg + flux,
                   error_pct=100.0 * (stg + bnd) / max(abs(stg) + abs(bnd), 1e-10),
                   cum_storage=c["storage"], cum_efflux=c["efflux"], cum_flux=c["flux"],
                   cum_balance=c["storage"] + c["efflux"], cum_net=c["storage"] + c["flux"],
                   cum_error_pct=100.0 * (c["storage"] + c["efflux"]) / max(c["norm"], 1e-10))
        wet = model.get("iwet") == 1
        p, u, v = model.get("p"), model.get("u"), model.get("v")
        cel = np.sqrt(GRAV * np.where(wet, h, 0.0))
        cr = dt * np.maximum((np.abs(u) + cel) / self.dx, (np.abs(v) + cel) / self.dy)
        for name, key, val in (("wse", np.abs(p), p * GRAVINV), ("u", np.abs(u), u), ("v", np.abs(v), v),
                               ("courant", cr, cr)):
            k = np.where(wet, key, -1.0)
            i = int(np.argmax(k))
            if k[i] > 0.0:
                out[name], out[name + "_cell"] = float(val[i]), i
            else:
                out[name], out[name + "_cell"] = 0.0, 0
        return out
