"""Ponding (flow_pond, source/flow/flow.F90:1807-1876) on the GPU against the
oracle: wet pits on land that no forcing boundary reaches are dried; a wet
channel connected to the sea stays wet.  The pits sit across the 32x32 tiles
of the device labelling so that unions at tile edges are exercised."""
import numpy as np
import pytest

from cms2d_amd import cases
from cms2d_amd.flow import FlowStep
from pyoracle import Oracle


def pond_case():
    c = cases.tidal(nx=150, ny=90, dx=20.0, seed=7, radstress=False)
    zb = c.zb.reshape(c.ny, c.nx).copy()
    zb[:, :60] = np.maximum(zb[:, :60], 1.0)           # dry land strip on the west
    zb[10:20, 5:15] = -1.5                              # isolated pits
    zb[28:37, 25:40] = -1.5                             # straddles tile rows/columns 32
    zb[60:64, 40:44] = -1.5
    zb[70:72, 20:100] = -1.5                            # channel reaching the sea: stays wet
    c.zb = np.asarray(zb, np.float32).astype(np.float64).ravel()
    c.name = "pond"
    return c


def test_oracle_dries_isolated_pits():
    c = pond_case()
    o = Oracle(c, serial=0)
    w = o.get("iwet").reshape(c.ny, c.nx)
    assert w[10:20, 5:15].sum() == 0 and w[28:37, 25:40].sum() == 0 and w[60:64, 40:44].sum() == 0
    assert w[70:72, 20:100].all()


@pytest.mark.gpu
def test_gpu_pond_matches_oracle(gpu_available):
    c = pond_case()
    print(f"case {c.name} {c.nx}x{c.ny}", flush=True)
    o = Oracle(c, serial=0)
    g = FlowStep(c)
    assert np.array_equal(o.get("iwet"), g.get("iwet"))
    for _ in range(3):
        o.step(1)
        g.step(1)
        assert np.array_equal(o.get("iwet"), g.get("iwet"))
    e = np.abs(o.get("eta") - g.get("eta"))[o.get("iwet") > 0].max()
    assert e <= 1e-6, e


def tel_pond_case():
    """The same land, pits and channel on a telescoping grid whose refined
    band (base columns 5-13, 5 m core 7-11) runs through the first pit."""
    c = cases.tidal_tel(nx=150, ny=90, dx=20.0, band=((5, 14), (7, 12)), seed=7, radstress=False)
    g = np.asarray(c.tel.geom, dtype=np.float64)
    bi, bj = np.floor(g[:, 0] / c.dx).astype(int), np.floor(g[:, 1] / c.dx).astype(int)
    zb = c.zb.copy()
    land = bi < 60
    zb[land] = np.maximum(zb[land], 1.0)
    for (j0, j1, i0, i1) in [(10, 20, 5, 15), (28, 37, 25, 40), (60, 64, 40, 44), (70, 72, 20, 100)]:
        zb[(bj >= j0) & (bj < j1) & (bi >= i0) & (bi < i1)] = -1.5
    c.zb = np.asarray(zb, np.float32).astype(np.float64)
    c.name = "tel_pond"
    return c, bi, bj


def test_oracle_dries_isolated_pits_on_telescoping_grid():
    c, bi, bj = tel_pond_case()
    w = Oracle(c, serial=0).get("iwet")
    pit = lambda j0, j1, i0, i1: (bj >= j0) & (bj < j1) & (bi >= i0) & (bi < i1)  # noqa: E731
    assert w[pit(10, 20, 5, 15)].sum() == 0 and w[pit(28, 37, 25, 40)].sum() == 0
    assert w[pit(60, 64, 40, 44)].sum() == 0 and w[pit(70, 72, 20, 100)].all()


@pytest.mark.gpu
def test_gpu_pond_matches_oracle_on_telescoping_grid(gpu_available):
    """The cell-list labelling (tiles of 32 x 32 base cells in LDS, unions
    across tile edges) against the oracle's sweeps."""
    c, _, _ = tel_pond_case()
    print(f"case {c.name} {c.ncells} cells", flush=True)
    o = Oracle(c, serial=0)
    g = FlowStep(c)
    assert np.array_equal(o.get("iwet"), g.get("iwet"))
    for _ in range(3):
        o.step(1)
        g.step(1)
        assert np.array_equal(o.get("iwet"), g.get("iwet"))
    e = np.abs(o.get("eta") - g.get("eta"))[o.get("iwet") > 0].max()
    assert e <= 1e-6, e
    g.close()
