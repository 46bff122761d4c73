"""The Fortran host layer (ISO_C_BINDING over the C ABI)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "cms2d_amd", "fortran")
RUN = os.path.join(FDIR, "flow_run")


def test_fortran_layer_built_and_struct_layout_matches_c():
    from cms2d_amd import _lib
    assert os.path.exists(RUN), "cms2d_amd/fortran/flow_run not built"
    assert os.path.exists(os.path.join(FDIR, "libcms2d_fortran.so"))
    out = subprocess.run([RUN, "--sizeof"], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == C.sizeof(_lib.FlowDesc)
    assert int(out[1]) == C.sizeof(_lib.StepInfo)


@pytest.mark.gpu
def test_fortran_time_loop_matches_python_front_end(gpu_available, tmp_path):
    from cms2d_amd import cases
    from cms2d_amd.flow import FlowStep
    from fortran_case import write
    c = cases.tidal(nx=40, ny=24, dx=20.0, seed=2)
    cf, of = tmp_path / "case.bin", tmp_path / "out.bin"
    write(c, cf)
    print("fortran flow_run tidal 40x24, 8 steps", flush=True)
    r = subprocess.run([RUN, str(cf), "8", str(of)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    res = np.fromfile(of, dtype=np.float64).reshape(3, -1)
    g = FlowStep(c)
    g.step(8)
    for k, name in enumerate(("eta", "u", "v")):
        assert np.array_equal(res[k], g.get(name)), name
