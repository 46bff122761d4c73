"""The Fortran host layer (ISO_C_BINDING over the C ABI)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "cms2d_amd", "fortran")
RUN = os.path.join(FDIR, "flow_run")
SOLVE_RUN = os.path.join(FDIR, "solve_run")


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


def write_system(path, n, nd, c2c, acoef, iwet, ss, sp, phi, relax):
    """The system file of solve_run: reference shapes, Fortran (column-major)
    order, cell2cell 1-based with 0 for no face."""
    nf = c2c.shape[1]
    with open(path, "wb") as f:
        np.array([n, nd, nf], dtype=np.int32).tofile(f)
        np.array([relax], dtype=np.float64).tofile(f)
        np.ascontiguousarray(c2c + 1, dtype=np.int32).tofile(f)      # [nd, nf] C order = (nf, nd) Fortran
        np.ascontiguousarray(iwet, dtype=np.int32).tofile(f)
        np.ascontiguousarray(acoef, dtype=np.float64).tofile(f)
        for a in (ss, sp, phi):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)


@pytest.mark.gpu
@pytest.mark.parametrize("relax", [1.0, 0.8])
def test_fortran_solve_caller_on_telescoping_list(gpu_available, tmp_path, relax):
    """cms2d_solve_ref called from Fortran with acoef(6,ncellsD) and 1-based
    cell2cell(6,ncellsD) gives the Python front end's result bit for bit."""
    from cell_list import tel_cell_list
    from cms2d_amd.flow import solve
    rng = np.random.default_rng(3)
    refined = np.zeros((20, 28), bool)
    refined[4:14, 0:9] = True
    refined[12:17, 8:15] = True
    n, nd, c2c, _, _ = tel_cell_list(refined)
    acoef = np.where(c2c >= 0, rng.uniform(0.2, 2.0, c2c.shape), 0.0)
    acoef[n:] = 0.0
    sp = np.zeros(nd)
    sp[:n] = -rng.uniform(0.05, 0.5, n)
    ss = np.zeros(nd)
    ss[:n] = rng.normal(size=n)
    phi = rng.normal(scale=0.1, size=nd)
    iwet = np.ones(nd, dtype=np.int32)
    sf, of = tmp_path / "sys.bin", tmp_path / "out.bin"
    write_system(sf, n, nd, c2c, acoef, iwet, ss, sp, phi, relax)
    print(f"fortran solve_run telescoping list {n} cells, relax {relax}", flush=True)
    r = subprocess.run([SOLVE_RUN, str(sf), str(of)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(of, dtype=np.float64)
    phi_f, res_f, ss_f, rmom_f = out[:nd], out[nd:2 * nd], out[2 * nd:3 * nd], out[3 * nd]
    ss_p, phi_p = ss.copy(), phi.copy()
    res_p, rmom_p, _ = solve(acoef, c2c, iwet, ss_p, sp, phi_p, relax, ncells=n)
    assert np.array_equal(phi_f, phi_p)
    assert np.array_equal(res_f[:n], res_p[:n])
    assert np.array_equal(ss_f[:n], ss_p[:n])
    assert rmom_f == rmom_p
