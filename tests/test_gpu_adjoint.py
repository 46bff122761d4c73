"""The reverse of solve (cms2d_solve_adjoint) and the torch.autograd.Function
built on it (cms2d_amd.adjoint).  The adjoint has no counterpart in the
reference; it is pinned by what the mathematics fixes: the transposed solve
against a dense solve of A^T, the dot-product identity <A^-1 v, w> =
<v, A^-T w>, and central finite differences (h = 1e-6) of the forward GPU
solve for every kind of input, on Cartesian and telescoping cell lists."""
import numpy as np
import pytest

from cell_list import dense_system, tel_cell_list
from cms2d_amd.flow import solve, solve_adjoint
from test_gpu_solve import make_system

RTOL = 1e-14          # solves converge to rounding, so differences see only h


def tel_system(seed):
    rng = np.random.default_rng(seed)
    refined = np.zeros((6, 9), dtype=bool)
    refined[1:4, 2:6] = True
    n, nd, c2c, _, _ = tel_cell_list(refined)
    nf = c2c.shape[1]
    acoef = np.where(c2c >= 0, rng.uniform(0.2, 2.0, (nd, nf)), 0.0)
    acoef[n:] = 0.0
    sp = np.zeros(nd)
    sp[:n] = -rng.uniform(0.05, 0.5, n)
    ss = np.zeros(nd)
    ss[:n] = rng.normal(size=n)
    phi = np.zeros(nd)
    phi[:n] = rng.normal(scale=0.1, size=n)
    phi[n:] = rng.normal(size=nd - n)
    return n, nd, c2c, acoef, sp, ss, phi, np.ones(nd, dtype=np.int32)


SYSTEMS = {"cartesian": lambda s: make_system(s, nx=12, ny=9), "telescoping": tel_system}


def forward(n, c2c, acoef, sp, ss, phi_old, relax):
    s, p = ss.copy(), phi_old.copy()
    solve(acoef, c2c, np.ones(len(p), dtype=np.int32), s, sp, p, relax, ncells=n, rtol=RTOL, atol=0.0)
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("kind", list(SYSTEMS))
@pytest.mark.parametrize("relax", [1.0, 0.8])
def test_transposed_solve_matches_dense(gpu_available, kind, relax):
    n, nd, c2c, acoef, sp, ss, phi, _ = SYSTEMS[kind](3)
    print(f"adjoint {kind} relax {relax} n {n}", flush=True)
    A, _ = dense_system(n, acoef, c2c, sp, ss, phi, relax)
    g = np.random.default_rng(4).normal(size=n)
    lam = np.linalg.solve(A.T, g)
    x = forward(n, c2c, acoef, sp, ss, phi, relax)
    gss, gsp, ga, gpo, it = solve_adjoint(acoef, c2c, sp, phi, x, relax, g, ncells=n, rtol=RTOL, atol=0.0)
    assert it > 0
    assert np.abs(gss[:n] - lam).max() <= 1e-10 * np.abs(lam).max()
    assert not gss[n:].any() and not gsp[n:].any() and not ga[n:].any()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", list(SYSTEMS))
def test_dot_product_identity(gpu_available, kind):
    """<J v, w> = <v, J^T w> for J = d phi / d ss (relax 1: phi = A^-1 (ss + ghost terms))."""
    n, nd, c2c, acoef, sp, _, phi, _ = SYSTEMS[kind](5)
    rng = np.random.default_rng(6)
    v = np.zeros(nd)
    v[:n] = rng.normal(size=n)
    w = rng.normal(size=n)
    p0 = np.zeros(nd)                       # ghost values 0: the map is linear in ss
    jv = forward(n, c2c, acoef, sp, v, p0, 1.0)[:n]
    jtw = solve_adjoint(acoef, c2c, sp, p0, np.concatenate([jv, p0[n:]]), 1.0, w, ncells=n, rtol=RTOL,
                        atol=0.0)[0][:n]
    lhs, rhs = jv @ w, v[:n] @ jtw
    print(kind, lhs, rhs, flush=True)
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), np.linalg.norm(jv) * np.linalg.norm(w))


@pytest.mark.gpu
@pytest.mark.parametrize("kind", list(SYSTEMS))
@pytest.mark.parametrize("relax", [1.0, 0.8])
def test_gradients_match_central_differences(gpu_available, kind, relax):
    """L = 1/2 |phi - d|^2 over the active cells; 32 random inputs of each
    kind (coefficients of active faces, including faces to ghosts; sp, ss and
    phi_old of active cells; ghost values) against central differences."""
    n, nd, c2c, acoef, sp, ss, phi, _ = SYSTEMS[kind](7)
    rng = np.random.default_rng(8)
    d = rng.normal(scale=0.1, size=n)

    def loss(a=acoef, s_p=sp, s_s=ss, po=phi):
        x = forward(n, c2c, a, s_p, s_s, po, relax)[:n]
        return 0.5 * np.sum((x - d) ** 2)

    x = forward(n, c2c, acoef, sp, ss, phi, relax)
    gss, gsp, ga, gpo, _ = solve_adjoint(acoef, c2c, sp, phi, x, relax, x[:n] - d, ncells=n, rtol=RTOL,
                                         atol=0.0)
    worst = 0.0

    def check(arr, idx, grad, name):
        nonlocal worst
        for i in idx:
            h = 1e-6 * max(1.0, abs(arr.flat[i]))
            ap, am = arr.copy(), arr.copy()
            ap.flat[i] += h
            am.flat[i] -= h
            kw = {name: ap}
            lp = loss(**kw)
            kw = {name: am}
            lm = loss(**kw)
            fd = (lp - lm) / (2 * h)
            err = abs(fd - grad.flat[i]) / max(abs(grad.flat[i]), 1e-3)
            worst = max(worst, err)
            assert err <= 1e-4, (name, i, fd, grad.flat[i])

    faces = np.flatnonzero((c2c[:n] >= 0).ravel())
    check(acoef, rng.choice(faces, 32, replace=False), ga, "a")
    check(sp, rng.choice(n, 32, replace=False), gsp, "s_p")
    check(ss, rng.choice(n, 32, replace=False), gss, "s_s")
    check(phi, rng.choice(n, 32, replace=False), gpo, "po")
    check(phi, np.arange(n, nd)[:32], gpo, "po")
    print(kind, relax, "worst relative difference", worst, flush=True)


@pytest.mark.gpu
def test_autograd_function_gradcheck(gpu_available):
    """torch.autograd.gradcheck of the Function (every input element)."""
    import torch

    from cms2d_amd.adjoint import solve_fn
    n, nd, c2c, acoef, sp, ss, phi, _ = tel_system(9)
    t = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in (acoef, ss, sp, phi)]
    fn = lambda a, s, p, po: solve_fn(a, s, p, po, c2c, n, 0.8)  # noqa: E731
    assert torch.autograd.gradcheck(fn, t, eps=1e-6, atol=1e-6, rtol=1e-4)
