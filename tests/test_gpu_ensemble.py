"""Batched ensemble (cms2d_ensemble_step): members advanced side by side on
host threads must give bitwise the states they reach when advanced alone, and
one member is checked against the FP64 oracle."""
import numpy as np
import pytest

from cms2d_amd import _lib, cases
from cms2d_amd.ensemble import Ensemble
from cms2d_amd.flow import FlowStep
from pyoracle import Oracle


def test_member_generator_ranges():
    ms = cases.ensemble_members(n=40, seed=3)
    assert len(ms) == 40
    for c in ms:
        assert 64 <= c.nx <= 512 and 64 <= c.ny <= 512
        assert 0.2 <= c.amp[0] <= 1.0
        assert 0.015 * np.exp(-1.5) < np.median(c.mann) < 0.035 * np.exp(1.5)
    assert len({(c.nx, c.ny) for c in ms}) > 30


@pytest.mark.gpu
def test_ensemble_equals_members_alone(gpu_available):
    ms = cases.ensemble_members(n=6, seed=5, side_min=40, side_max=120)
    print("ensemble members", [(c.nx, c.ny) for c in ms], flush=True)
    e = Ensemble(ms, nthreads=4, precision=_lib.PREC_FP32_REFINE)
    alone = [FlowStep(c, precision=_lib.PREC_FP32_REFINE) for c in ms]
    for _ in range(3):
        infos = e.step(2)
        for g in alone:
            g.step(2)
    for m, g in zip(e.members, alone):
        for name in ("eta", "u", "v", "iwet"):
            assert np.array_equal(m.get(name), g.get(name)), name
    assert all(i["iconv"] in (1, 2) for i in infos)
    # one member against the oracle
    o = Oracle(ms[0], serial=0)
    o.step(6)
    wet = o.get("iwet") > 0
    assert np.array_equal(o.get("iwet"), e.members[0].get("iwet"))
    err = np.abs(o.get("eta") - e.members[0].get("eta"))[wet].max()
    print("member 0 vs oracle eta", err)
    assert err <= 1e-6
    e.close()


@pytest.mark.gpu
def test_ensemble_at_config_size(gpu_available):
    """256 members with sides uniform in 64-512 (about 20 M cells), the FP32
    form at dt 60 s: every member converges, stays finite and keeps its water
    volume against its boundary inflow; the two smallest members match the
    FP64 oracle run alone."""
    ms = cases.ensemble_members(n=256, seed=11)
    print("ensemble 256 members", sum(c.nx * c.ny for c in ms), "cells", flush=True)
    e = Ensemble(ms, nthreads=8, precision=_lib.PREC_FP32_REFINE)
    e.step(2)                                       # spin-up
    vol0 = [m.get("h").sum() * c.dx * c.dy for m, c in zip(e.members, ms)]
    infos = e.step(1)
    assert all(i["iconv"] in (1, 2) for i in infos)
    worst = 0.0
    for m, c, v0 in zip(e.members, ms, vol0):
        for name in ("eta", "u", "v"):
            assert np.isfinite(m.get(name)).all(), name
        v1 = m.get("h").sum() * c.dx * c.dy
        f = m.get_face("flux").reshape(c.ny, c.nx, 4)
        q = -(f[-1, :, 0].sum() + f[:, -1, 1].sum() + f[0, :, 2].sum() + f[:, 0, 3].sum())
        err = abs((v1 - v0) / c.dt - q)
        worst = max(worst, err / max(abs(q), 1e-9))
        assert err <= 5e-3 * max(abs(q), 1e-6 * v1 / c.dt), (c.nx, c.ny, (v1 - v0) / c.dt, q)
    print("worst relative imbalance", worst, flush=True)
    order = np.argsort([c.nx * c.ny for c in ms])[:2]
    for k in order:
        o = Oracle(ms[k], serial=0)
        o.step(3)
        wet = o.get("iwet") > 0
        assert np.array_equal(o.get("iwet"), e.members[k].get("iwet"))
        err = np.abs(o.get("eta") - e.members[k].get("eta"))[wet].max()
        print("member", k, ms[k].nx, ms[k].ny, "vs oracle eta", err, flush=True)
        assert err <= 1e-6
    e.close()
