"""Flow statistics on the HIP path (cms2d_flow_statistics / cms2d_flow_stat)
against the oracle's restatement step by step (1e-9 of each field's largest
magnitude, at least 1: the state's agreement), and against the
reference's own accumulators at the oracle's bound."""
import numpy as np
import pytest

from cms2d_amd.flow import FlowStep
from pyoracle import Oracle
from test_stats_oracle import GOLD, STATS, check_bounds, compare_stats, stats_case

pytestmark = pytest.mark.gpu


def test_statistics_match_oracle_and_reference(gpu_available):
    c = stats_case()
    g = np.load(GOLD)
    print("flow statistics, tidal 48x32, windows of 180 s", flush=True)
    o, f = Oracle(c, serial=0), FlowStep(c)
    o.stats(0.0, 180.0)
    f.statistics(0.0, 180.0)
    for n in range(1, 12):
        assert o.step(1)[0] == 0
        f.step(1)
        for k in STATS:
            a, na = f.stat(k)
            b, nb = o.stat(k)
            assert na == nb
            assert np.abs(a - b).max() <= 1e-9 * max(1.0, np.abs(b).max()), (n, k)
    fs = FlowStep(c, serial_order=1)
    fs.statistics(0.0, 180.0)
    for n in range(1, 12):
        fs.step(1)
        if n in (5, 11):
            w = compare_stats(fs, g, n)
            print(n, sorted(w.items(), key=lambda x: -x[1])[:4], flush=True)
            check_bounds(w)
    f.close()
    fs.close()


def test_statistics_rejects():
    c = stats_case()
    f = FlowStep(c)
    with pytest.raises(RuntimeError, match="not enabled"):
        f.stat("etamax")
    with pytest.raises(RuntimeError, match="t_end > t_start"):
        f.statistics(10.0, 10.0)
    f.statistics(0.0, 60.0)
    with pytest.raises(RuntimeError, match="unknown statistic"):
        f.stat("nope")
    f.close()
