"""cms2d_solve (drop-in for the reference's solve) on a general cell list."""
import ctypes as C

import numpy as np
import pytest

from cell_list import cart_cell_list, dense_system
from cms2d_amd.flow import solve
from pyoracle import D, I, lib as olib


def make_system(seed, nx=24, ny=18, land=0.15):
    rng = np.random.default_rng(seed)
    act = rng.random((ny, nx)) > land
    n, nd, c2c = cart_cell_list(act)
    acoef = np.zeros((nd, 4))
    acoef[:n] = rng.uniform(0.2, 2.0, (n, 4))
    sp = np.zeros(nd)
    sp[:n] = -rng.uniform(0.05, 0.5, n)
    ss = np.zeros(nd)
    ss[:n] = rng.normal(size=n)
    phi = np.zeros(nd)
    phi[:n] = rng.normal(scale=0.1, size=n)
    phi[n:] = rng.normal(size=nd - n)
    iwet = np.ones(nd, dtype=np.int32)
    return n, nd, c2c, acoef, sp, ss, phi, iwet


@pytest.mark.gpu
@pytest.mark.parametrize("seed,relax", [(0, 1.0), (1, 0.8), (2, 1.0)])
def test_solve_matches_dense_and_oracle(gpu_available, seed, relax):
    n, nd, c2c, acoef, sp, ss, phi, iwet = make_system(seed)
    print(f"solve seed {seed} relax {relax} n {n}", flush=True)
    A, b = dense_system(n, acoef, c2c, sp, ss, phi, relax)
    x = np.linalg.solve(A, b + 1e-15 * 0)
    ss_g, phi_g = ss.copy(), phi.copy()
    res_g, rmom_g, it_g = solve(acoef, c2c, iwet, ss_g, sp, phi_g, relax, ncells=n)
    ss_o, phi_o, res_o = ss.copy(), phi.copy(), np.zeros(nd)
    rm = C.c_double()
    olib().cms_oracle_solve(n, nd, acoef.ctypes.data_as(D), c2c.ctypes.data_as(I),
                            iwet.ctypes.data_as(I), ss_o.ctypes.data_as(D), sp.ctypes.data_as(D),
                            res_o.ctypes.data_as(D), phi_o.ctypes.data_as(D), relax, 1e-10, 1e-12,
                            4000, C.byref(rm))
    scale = np.abs(x).max()
    assert np.abs(phi_g[:n] - x).max() <= 1e-8 * scale
    assert np.abs(phi_o[:n] - x).max() <= 1e-8 * scale
    assert np.array_equal(phi_g[n:], phi[n:])                 # ghosts untouched
    assert np.abs(rmom_g - rm.value) <= 1e-12 * max(rm.value, 1e-30)
    assert np.abs(res_g[:n] - res_o[:n]).max() <= 1e-12 * np.abs(res_o).max()
    assert np.abs(ss_g[:n] - ss_o[:n]).max() <= 1e-13 * np.abs(ss_o).max()
    assert it_g > 0
