"""ctypes binding of libcms2d_hip.so, the C ABI declared in include/cms2d_amd.h.

This is the binding a Python caller of the reference would add (the reference
itself binds the same entry points from Fortran through ISO_C_BINDING; see
cms2d_amd/fortran/cms2d_hip_mod.f90).  Loading fails loudly when the HIP library
has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libcms2d_hip.so")

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)

BND_WALL = 0
BND_WSE_HARMONIC = 1
BND_WSE_CURVE = 2
BND_FLUX = 3
PREC_FP64 = 0
PREC_FP32_REFINE = 1
PRECOND_MG = 0
PRECOND_JACOBI = 1
TURB_SUBGRID, TURB_CONSTANT, TURB_FALCONER, TURB_PARABOLIC, TURB_SUBGRID_WU = 0, 1, 2, 3, 4
FRIC_MANNING, FRIC_DRAG, FRIC_ROUGHNESS, FRIC_LINEAR = 0, 1, 2, 3
SOLVER_BICGSTAB = 0
SOLVER_GMRES = 1

ERR = {0: "ok", -1: "invalid argument", -2: "HIP error", -3: "diverged", -4: "no device"}

# every symbol include/cms2d_amd.h declares
EXPORTS = [
    "cms2d_last_error", "cms2d_device_count", "cms2d_flow_create", "cms2d_flow_destroy",
    "cms2d_flow_step", "cms2d_flow_get", "cms2d_flow_get_face", "cms2d_flow_set",
    "cms2d_flow_time", "cms2d_flow_kernel_stats", "cms2d_flow_tuning", "cms2d_flow_set_waves",
    "cms2d_flow_checkpoint", "cms2d_flow_restore", "cms2d_flow_set_wind",
    "cms2d_ensemble_step", "cms2d_solve", "cms2d_solve_noalloc", "cms2d_flow_create_tel",
    "cms2d_solve_adjoint",
    "cms2d_flow_adjoint", "cms2d_flow_balance", "cms2d_flow_statistics", "cms2d_flow_stat",
]


class FlowDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
        ("zb", D), ("mann", D), ("wavestrx", D), ("wavestry", D),
        ("side_type", C.c_int * 4), ("ntc", C.c_int),
        ("amp", C.c_double * 8), ("speed", C.c_double * 8), ("phase", C.c_double * 8),
        ("dt", C.c_double), ("rampdur", C.c_double), ("hdry", C.c_double),
        ("noptset", C.c_int), ("wallfric", C.c_int), ("ndsch", C.c_int), ("maxit", C.c_int),
        ("rtol", C.c_double), ("atol", C.c_double), ("kmax", C.c_int), ("precision", C.c_int),
        ("device", C.c_int), ("precond", C.c_int), ("solver", C.c_int), ("dtvar", C.c_int),
        ("latitude", C.c_double), ("coriolis_plane", C.c_int), ("turbulence", C.c_int),
        ("friction", C.c_int),
        ("wse_times", D), ("wse_values", D), ("wse_ntimes", C.c_int), ("wse_order", C.c_int),
        ("wse_adjust_off", C.c_int), ("wse_gradx", C.c_double), ("wse_grady", C.c_double),
        ("serial_order", C.c_int),
        ("flux_q", C.c_double), ("flux_times", D), ("flux_values", D), ("flux_ntimes", C.c_int),
        ("flux_order", C.c_int), ("flux_conveyance", C.c_double), ("flux_angle_on", C.c_int),
        ("flux_angle", C.c_double), ("flux_rating", C.c_int),
    ]


class TelGrid(C.Structure):
    _fields_ = [("ncells", C.c_int), ("geom", D), ("nb", I), ("bnd_cells", I), ("nbnd", C.c_int),
                ("flux_cells", I), ("nflux", C.c_int)]


class StepInfo(C.Structure):
    _fields_ = [("niter", C.c_int), ("iconv", C.c_int), ("rmom", C.c_double * 3),
                ("kry_iters", C.c_int * 3), ("nwet", C.c_int), ("kernel_ms", C.c_double),
                ("dt", C.c_double)]


class Balance(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "storage", "efflux", "flux", "balance", "net", "error_pct",
        "cum_storage", "cum_efflux", "cum_flux", "cum_balance", "cum_net", "cum_error_pct",
        "wse", "u", "v", "courant")] + [
        (n, C.c_int) for n in ("wse_cell", "u_cell", "v_cell", "courant_cell", "steps")]


_LIB = None


class Cms2dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cms2d error {code} ({ERR.get(code, '?')}): {msg}")
        self.code = code


def lib():
    """Loads the HIP library (raises if it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"HIP library not built: {LIBPATH} (run python -m cms2d_amd.build)")
    L = C.CDLL(LIBPATH)
    L.cms2d_last_error.restype = C.c_char_p
    L.cms2d_device_count.argtypes = [I]
    L.cms2d_flow_create.argtypes = [C.POINTER(FlowDesc), C.POINTER(C.c_void_p)]
    L.cms2d_flow_destroy.argtypes = [C.c_void_p]
    L.cms2d_flow_create_tel.argtypes = [C.POINTER(FlowDesc), C.POINTER(TelGrid), C.POINTER(C.c_void_p)]
    L.cms2d_flow_step.argtypes = [C.c_void_p, C.c_int, C.POINTER(StepInfo)]
    L.cms2d_flow_get.argtypes = [C.c_void_p, C.c_char_p, D]
    L.cms2d_flow_get_face.argtypes = [C.c_void_p, C.c_char_p, D]
    L.cms2d_flow_set.argtypes = [C.c_void_p, C.c_char_p, D]
    L.cms2d_flow_time.restype = C.c_double
    L.cms2d_flow_time.argtypes = [C.c_void_p]
    L.cms2d_flow_kernel_stats.argtypes = [C.c_void_p, C.c_char_p, D, C.POINTER(C.c_long)]
    L.cms2d_flow_tuning.argtypes = [C.c_void_p, I, I]
    L.cms2d_flow_set_waves.argtypes = [C.c_void_p, D, D, D, D, D]
    L.cms2d_flow_set_wind.argtypes = [C.c_void_p, C.c_double, C.c_double]
    L.cms2d_flow_checkpoint.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_size_t)]
    L.cms2d_flow_restore.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.cms2d_ensemble_step.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int,
                                      C.POINTER(StepInfo)]
    L.cms2d_solve.argtypes = [C.c_int, C.c_int, C.c_int, D, I, I, D, D, D, D, C.c_double, C.c_double,
                              C.c_double, C.c_int, D, I]
    L.cms2d_solve_noalloc.argtypes = L.cms2d_solve.argtypes
    L.cms2d_flow_balance.argtypes = [C.c_void_p, C.POINTER(Balance)]
    L.cms2d_flow_statistics.argtypes = [C.c_void_p, C.c_double, C.c_double]
    L.cms2d_flow_stat.argtypes = [C.c_void_p, C.c_char_p, D, C.POINTER(C.c_int)]
    L.cms2d_flow_adjoint.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, I, D, D, D, D, D, D]
    L.cms2d_solve_adjoint.argtypes = [C.c_int, C.c_int, C.c_int, D, I, D, D, D, C.c_double, D, C.c_double,
                                      C.c_double, C.c_int, D, D, D, D, I]
    _LIB = L
    return L


def check(rc):
    if rc != 0:
        raise Cms2dError(rc, lib().cms2d_last_error().decode())


def device_count():
    n = C.c_int(0)
    lib().cms2d_device_count(C.byref(n))
    return n.value
