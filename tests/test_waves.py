"""Wave forcing from wave fields: wave_rad_stress_fl
(source/waves/wave_rad_stress_fl.f90:2-59) in the oracle (CPU) and through
cms2d_flow_set_waves (GPU).

The oracle is pinned by closed forms: on a flat bed with the wave height
linear in x, central differences are exact for the quadratic stresses, so the
interior forcing is -(g/16)(2n - 1/2) 2 H dH/dx; with an oblique direction the
y forcing is -(g/16) n uy^2 2 H dH/dx, the reference's Sxy = E n uy uy (:31)
rather than the textbook E n ux uy, which the test tells apart."""
import numpy as np
import pytest

from cms2d_amd import _lib, cases
from pyoracle import Oracle

G = 9.81


def flat_case():
    c = cases.lake_at_rest(nx=40, ny=30)
    c.zb = np.full(c.nx * c.ny, -3.0, dtype=np.float32)
    c.noptset = 3
    return c


def linear_waves(c, theta_deg=0.0, b=0.01):
    jj, ii = np.divmod(np.arange(c.ncells), c.nx)
    x = (ii + 0.5) * c.dx
    H = 0.5 + b * x
    th = np.deg2rad(theta_deg)
    return (H, np.full(c.ncells, 8.0), np.full(c.ncells, 60.0), np.full(c.ncells, np.cos(th)),
            np.full(c.ncells, np.sin(th))), ii, jj


def n_group(h, L):
    kh = min(2.0 * np.pi / L * h, 10.0)
    return 0.5 + kh / np.sinh(2.0 * kh)


@pytest.mark.parametrize("theta", [0.0, 30.0])
def test_oracle_radiation_stress_closed_form(theta):
    c = flat_case()
    o = Oracle(c, serial=0)
    w, ii, jj = linear_waves(c, theta)
    assert o.set_waves(*w)
    fx, fy = o.get("wavestrx"), o.get("wavestry")
    inner = (ii > 0) & (ii < c.nx - 1) & (jj > 0) & (jj < c.ny - 1)
    H, ux, uy = w[0], w[3][0], w[4][0]
    cn = n_group(3.0, 60.0)
    dE = G / 16.0 * 2.0 * H * 0.01                     # dE/dx
    fx_exact = -dE * (cn * (1.0 + ux * ux) - 0.5)
    fy_exact = -dE * cn * uy * uy                       # Sxy = E n uy uy
    scale = np.abs(fx_exact).max()
    assert np.abs(fx - fx_exact)[inner].max() <= 1e-13 * scale
    assert np.abs(fy - fy_exact)[inner].max() <= 1e-13 * scale
    if theta:
        fy_textbook = -dE * cn * ux * uy
        assert np.abs(fy - fy_textbook)[inner].max() > 1e-3 * scale


def test_oracle_radiation_stress_error_keeps_forcing():
    c = flat_case()
    o = Oracle(c, serial=0)
    w, _, _ = linear_waves(c)
    assert o.set_waves(*w)
    before = o.get("wavestrx").copy()
    huge = (w[0] * 200.0,) + w[1:]
    assert not o.set_waves(*huge)                       # |forcing| > 10: the reference stops
    assert np.array_equal(o.get("wavestrx"), before)


def smooth_waves(c, seed=3):
    rng = np.random.default_rng(seed)
    jj, ii = np.divmod(np.arange(c.ncells), c.nx)
    x, y = ii / c.nx, jj / c.ny
    H = 0.4 + 0.3 * x + 0.1 * np.sin(2 * np.pi * (y + rng.random())) * x
    L = 40.0 + 30.0 * x
    th = np.deg2rad(180.0 + 20.0 * np.sin(2 * np.pi * y))   # waves travelling shoreward (-x)
    return H, np.full(c.ncells, 7.0), L, np.cos(th), np.sin(th)


@pytest.mark.gpu
def test_gpu_wave_forcing_matches_oracle(gpu_available):
    from cms2d_amd.flow import FlowStep
    from test_gpu_flow import TOL_ETA, TOL_VEL, compare
    c = cases.tidal(nx=96, ny=64, dx=20.0, seed=8)        # wave mode (noptset 3), wet/dry beach
    o = Oracle(c, serial=0)
    g = FlowStep(c)
    for _ in range(3):
        assert o.step(1)[0] == 0
        g.step(1)
    w = smooth_waves(c)
    print("wave forcing tidal 96x64", flush=True)
    assert o.set_waves(*w)
    g.set_waves(*w)
    for name in ("wavestrx", "wavestry"):
        a, b = o.get(name), g.get(name)
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max(), name
    for _ in range(8):
        assert o.step(1)[0] == 0
        g.step(1)
    print("max err", compare(o, g), TOL_ETA, TOL_VEL)


@pytest.mark.gpu
def test_gpu_wave_forcing_errors(gpu_available):
    from cms2d_amd.flow import FlowStep
    c = cases.tidal(nx=48, ny=32, dx=20.0, seed=8)
    g = FlowStep(c)
    w = smooth_waves(c)
    g.set_waves(*w)
    before = g.get("wavestrx")
    with pytest.raises(_lib.Cms2dError) as e:
        g.set_waves(w[0] * 200.0, *w[1:])
    assert e.value.code == -1
    assert np.array_equal(g.get("wavestrx"), before)
    g.close()
    c2 = cases.basin(nx=48, ny=32)                        # noptset 2: no wave mode
    g2 = FlowStep(c2)
    with pytest.raises(_lib.Cms2dError):
        g2.set_waves(*smooth_waves(c2))
    g2.close()
