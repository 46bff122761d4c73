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
