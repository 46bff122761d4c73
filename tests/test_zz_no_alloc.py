"""cms2d_solve_noalloc (no device allocation; static device scratch) against
cms2d_solve on the same systems: two calls in a row, the second smaller, so a
call that left stale scratch behind or sized a part from the previous call
would show.  Both entry points run the same kernels on the same fixed grid, so
the results are bitwise equal."""
import numpy as np
import pytest

from cell_list import tel_cell_list
from cms2d_amd.flow import solve
from test_gpu_solve import make_system


def _tel_system(seed):
    rng = np.random.default_rng(seed)
    refined = np.zeros((8, 12), dtype=bool)
    refined[2:5, 3:8] = True
    n, nd, c2c, _, _ = tel_cell_list(refined)
    nf = c2c.shape[1]
    acoef = np.where(c2c >= 0, rng.uniform(0.2, 2.0, (nd, nf)), 0.0)
    acoef[n:] = 0.0
    sp = np.zeros(nd)
    sp[:n] = -rng.uniform(0.05, 0.5, n)
    ss = np.zeros(nd)
    ss[:n] = rng.normal(size=n)
    phi = np.zeros(nd)
    phi[n:] = rng.normal(size=nd - n)
    iwet = np.ones(nd, dtype=np.int32)
    return n, nd, c2c, acoef, sp, ss, phi, iwet


def _both(n, nd, c2c, acoef, sp, ss, phi, iwet, relax):
    ss_a, phi_a = ss.copy(), phi.copy()
    res_a, rm_a, it_a = solve(acoef, c2c, iwet, ss_a, sp, phi_a, relax, ncells=n)
    ss_b, phi_b = ss.copy(), phi.copy()
    res_b, rm_b, it_b = solve(acoef, c2c, iwet, ss_b, sp, phi_b, relax, ncells=n, noalloc=True)
    assert it_b == it_a > 0
    assert rm_b == rm_a
    assert np.array_equal(phi_b, phi_a)
    assert np.array_equal(ss_b[:n], ss_a[:n])
    assert np.array_equal(res_b[:n], res_a[:n])


@pytest.mark.gpu
def test_noalloc_solve_matches_solve_on_two_calls(gpu_available):
    big = make_system(5, nx=160, ny=120)
    small = make_system(6, nx=24, ny=18)
    print(f"noalloc: cartesian n={big[0]} then n={small[0]}", flush=True)
    _both(*big, relax=0.8)
    _both(*small, relax=1.0)


@pytest.mark.gpu
def test_noalloc_solve_telescoping_after_cartesian(gpu_available):
    cart = make_system(7, nx=64, ny=48)
    tel = _tel_system(8)
    print(f"noalloc: cartesian n={cart[0]} then telescoping n={tel[0]}", flush=True)
    _both(*cart, relax=1.0)
    _both(*tel, relax=1.0)
