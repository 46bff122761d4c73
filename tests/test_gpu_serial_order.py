"""The HIP path with serial_order=1: the three reference loops whose
iterations read values that earlier iterations write (the dry-to-wet start
values, flow.F90:1477-1491; the eddy-viscosity fix-ups next to dry cells and
at corners, flow_eddyvis.f90:168-212) give the result of the reference run on
one thread, which is how the golden vectors were made.  The race-free default
differs from that order where flow runs strongly along the wet/dry edge (the
WSE_GRADIENTS case, 9e-5 m/s) and with a flux string on the north or south
side (9e-6 and 4e-6 m/s; the race-free oracle differs from the reference by the
same), so those cases are checked against the reference here."""
import numpy as np
import pytest

from cms2d_amd.flow import FlowStep
from pyoracle import Oracle
from test_reference_vectors import (ALL, CASES, RIVER, TEL, TOL64_ETA, TOL64_VEL, TOL_ETA, TOL_VEL,
                                    compare_to_gold)

SERIAL = [n for n in ALL if n not in TEL]     # serial_order is built for Cartesian grids


@pytest.mark.gpu
@pytest.mark.parametrize("name", SERIAL)
def test_serial_order_matches_reference_vectors(gpu_available, name):
    print(f"serial order, reference vectors {name}", flush=True)
    w = compare_to_gold(name, lambda c: FlowStep(c, serial_order=1))
    print(name, w)
    assert w["eta"] <= TOL_ETA and w["u"] <= TOL_VEL and w["v"] <= TOL_VEL, w


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tidal_wse_grad", "tidal_wse_grad_north", "tidal", "tidal_combo", *RIVER])
def test_serial_order_matches_double_precision_reference(gpu_available, name):
    print(f"serial order, double-precision reference {name}", flush=True)
    w = compare_to_gold(name, lambda c: FlowStep(c, serial_order=1), prefix="ref64_")
    print(name, w)
    assert w["eta"] <= TOL64_ETA and w["u"] <= TOL64_VEL and w["v"] <= TOL64_VEL, w


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tidal_wse_grad", "tidal", "tidal_combo"])
def test_serial_order_matches_serial_oracle(gpu_available, name):
    """Step by step against the oracle in the reference's order."""
    c = CASES[name]()
    print(f"serial order against the serial oracle {name}", flush=True)
    o, g = Oracle(c, serial=1), FlowStep(c, serial_order=1)
    worst = 0.0
    for s in range(40):
        rc, _ = o.step(1)
        g.step(1)
        assert rc == 0
        assert np.array_equal(g.get("iwet"), o.get("iwet")), s
        w = o.get("iwet") > 0
        d = max(np.abs(g.get(f) - o.get(f))[w].max() for f in ("eta", "u", "v"))
        worst = max(worst, d)
        assert d <= 1e-9, (s, d)
    print(name, "max difference", worst)
    g.close()


@pytest.mark.gpu
def test_serial_order_differs_from_race_free_where_order_matters(gpu_available):
    """On the WSE_GRADIENTS case the two orders give different answers, and
    only the serial one is within the bounds of the reference's vectors."""
    race = compare_to_gold("tidal_wse_grad", lambda c: FlowStep(c), prefix="ref64_")
    ser = compare_to_gold("tidal_wse_grad", lambda c: FlowStep(c, serial_order=1), prefix="ref64_")
    print("race-free", race, "serial", ser)
    assert max(race["u"], race["v"]) > 10 * max(ser["u"], ser["v"])
