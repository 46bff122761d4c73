"""The C-ABI library loads on a CPU-only host and exports every entry point
include/cms2d_amd.h declares; without a device the entry points fail cleanly."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "cms2d_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cms2d_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from cms2d_amd import _lib
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert sorted(_lib.EXPORTS) == names


def test_errors_without_device():
    from cms2d_amd import _lib
    if _lib.device_count() > 0:
        pytest.skip("a device is present")
    from cms2d_amd import cases
    from cms2d_amd.flow import FlowStep
    with pytest.raises(_lib.Cms2dError) as e:
        FlowStep(cases.lake_at_rest())
    assert e.value.code == -4


def test_invalid_arguments_rejected():
    from cms2d_amd import _lib
    L = _lib.lib()
    d = _lib.FlowDesc()
    d.nx = 1
    h = C.c_void_p()
    assert L.cms2d_flow_create(C.byref(d), C.byref(h)) == -1
    assert h.value is None
    out = np.zeros(4)
    assert L.cms2d_solve(0, 0, 4, out.ctypes.data_as(_lib.D), None, None, None, None, None, None,
                         1.0, 1e-10, 0.0, 10, None, None) == -1


@pytest.mark.parametrize("field,value", [("precision", 7), ("precond", 9), ("solver", 4)])
def test_unknown_options_rejected_before_device(field, value):
    """Option codes are validated (CMS2D_ERR_ARG) before any device is touched."""
    from cms2d_amd import _lib, cases
    from cms2d_amd.flow import FlowStep
    with pytest.raises(_lib.Cms2dError) as e:
        FlowStep(cases.lake_at_rest(), **{field: value})
    assert e.value.code == -1


def test_gmres_requires_multigrid_fp64():
    from cms2d_amd import _lib, cases
    from cms2d_amd.flow import FlowStep
    for kw in ({"precond": _lib.PRECOND_JACOBI}, {"precision": _lib.PREC_FP32_REFINE}):
        with pytest.raises(_lib.Cms2dError) as e:
            FlowStep(cases.lake_at_rest(), solver=_lib.SOLVER_GMRES, **kw)
        assert e.value.code == -1


def test_noalloc_solve_refuses_systems_beyond_its_scratch():
    """cms2d_solve_noalloc checks the scratch it needs before any device use."""
    import ctypes as C

    import numpy as np

    from cms2d_amd.flow import solve
    n = 600_000
    nd = n + 4
    c2c = np.full((nd, 4), -1, dtype=np.int32)
    acoef = np.zeros((nd, 4))
    iwet = np.ones(nd, dtype=np.int32)
    ss, sp, phi = np.zeros(nd), -np.ones(nd), np.zeros(nd)
    from cms2d_amd import _lib
    with pytest.raises(_lib.Cms2dError, match="scratch"):
        solve(acoef, c2c, iwet, ss, sp, phi, 1.0, ncells=n, noalloc=True)


def _desc(case, keep):
    from cms2d_amd import _lib
    d = _lib.FlowDesc()
    d.nx, d.ny, d.dx, d.dy = case.nx, case.ny, case.dx, case.dy
    for a in (case.zb, case.mann):
        keep.append(np.ascontiguousarray(a, dtype=np.float64))
    d.zb, d.mann = keep[-2].ctypes.data_as(_lib.D), keep[-1].ctypes.data_as(_lib.D)
    d.ntc, d.amp[0], d.speed[0] = 1, 0.5, 0.5
    d.dt, d.hdry, d.ndsch, d.rtol, d.atol, d.kmax = 30.0, 0.05, 4, 1e-10, 1e-12, 100
    return d


@pytest.mark.parametrize("sides", [(1, 1, 0, 0), (1, 0, 0, 2), (2, 0, 2, 0)])
def test_more_than_one_water_level_side_refused(sides):
    """Each water-level string is forced on its own: the side's set-up and the
    corner cells where two forced sides meet have no pinned counterpart, so a
    description with two or more of them is refused before any device use."""
    from cms2d_amd import _lib, cases
    L, keep = _lib.lib(), []
    d = _desc(cases.lake_at_rest(), keep)
    t, v = np.array([0.0, 1.0]), np.array([0.0, 0.1])
    d.wse_times, d.wse_values = t.ctypes.data_as(_lib.D), v.ctypes.data_as(_lib.D)
    d.wse_ntimes = 2
    for k in range(4):
        d.side_type[k] = sides[k]
    h = C.c_void_p()
    assert L.cms2d_flow_create(C.byref(d), C.byref(h)) == -1
    assert b"one water-level side" in L.cms2d_last_error()


def test_long_water_level_side_and_bad_gradient_refused():
    from cms2d_amd import _lib, cases
    L, keep = _lib.lib(), []
    c = cases.lake_at_rest(nx=9000, ny=3)
    d = _desc(c, keep)
    d.side_type[0] = _lib.BND_WSE_HARMONIC
    h = C.c_void_p()
    assert L.cms2d_flow_create(C.byref(d), C.byref(h)) == -1
    assert b"8192" in L.cms2d_last_error()
    d = _desc(cases.lake_at_rest(), keep)
    t, v = np.array([0.0, 1.0]), np.array([0.0, 0.1])
    d.wse_times, d.wse_values = t.ctypes.data_as(_lib.D), v.ctypes.data_as(_lib.D)
    d.wse_ntimes = 2
    d.side_type[1] = _lib.BND_WSE_CURVE
    d.wse_grady = float("nan")
    assert L.cms2d_flow_create(C.byref(d), C.byref(h)) == -1
    assert b"gradients" in L.cms2d_last_error()


def test_telescoping_grid_checked_before_device():
    """Malformed .tel arrays are refused with CMS2D_ERR_ARG before any device
    use; a well-formed grid reaches the device probe."""
    import dataclasses

    from cms2d_amd import _lib, cases
    from cms2d_amd.flow import FlowStep
    c = cases.tidal_tel(nx=12, ny=8, dx=20.0, band=((2, 6), (3, 5)), seed=1)
    bad = dataclasses.replace(c.tel, nb=np.where(c.tel.nb > 0, c.tel.nb + 10_000, 0).astype(np.int32))
    with pytest.raises(_lib.Cms2dError) as e:
        FlowStep(dataclasses.replace(c, tel=bad))
    assert e.value.code == -1
    nb = c.tel.nb.copy()
    nb[0, 0] = 0                              # one side of a pair without its reciprocal
    nb[0, 2] = 3
    with pytest.raises(_lib.Cms2dError) as e:
        FlowStep(dataclasses.replace(c, tel=dataclasses.replace(c.tel, nb=nb)))
    assert e.value.code == -1
    with pytest.raises(_lib.Cms2dError) as e:
        FlowStep(dataclasses.replace(c, ndsch=7))
    assert e.value.code == -1
    with pytest.raises(_lib.Cms2dError) as e:
        FlowStep(c)
    assert e.value.code == -4                 # no device here
