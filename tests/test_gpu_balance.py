"""cms2d_flow_balance: the volume balance (flow_balance,
source/flow/flow_imp.F90:2260-2372) and the step extremes (flow_step_stat,
source/flow/flow.F90:1912-2008) that the device evaluates after every step.

* Against the restatement over the HIP flow's own fields (oracle/pyoracle.py
  VolumeBalance, itself pinned to the reference's output): per-step volumes to
  rounding (1e-12 of the sum of the magnitudes of their terms: the device sums
  in another order), cumulative sums and percentages from those exactly,
  extremes and their cells exactly (Courant number to 1e-15).
* Against the reference's own balance (tests/golden/ref64_balance_*.npz) at the
  bounds the reference's 1e-3 inner solves leave (as for the oracle).
* A checkpoint restores the running sums; the variable time step counts each
  accepted step once, with its own time step."""
import os

import numpy as np
import pytest

from cms2d_amd import _lib
from cms2d_amd.flow import FlowStep
from pyoracle import VolumeBalance
from test_balance_oracle import (BAL_CASES, DTVAR_ETA, DTVAR_VEL, GOLD, TOL_ETA, TOL_VEL, TOL_VOL, VOLS,
                                 compare_balance, courant_tol, dtvar_case)

pytestmark = pytest.mark.gpu

STEP_VOLS = ("storage", "efflux", "flux")


def run_gpu(c, nsteps, **kw):
    f = FlowStep(c, **kw)
    vb = VolumeBalance(c)
    rows, mine = [], []
    for _ in range(nsteps):
        vb.before(f)
        h1 = vb.h1
        info = f.step(1)
        b = f.balance()
        r = vb.after(f, dt=info["dt"])
        # magnitudes of the terms each volume sums
        face = f.get_face("flux")
        mag = {"storage": float((vb.area * np.abs(f.get("h") - h1)).sum()),
               "flux": info["dt"] * float(np.abs(face).sum()),
               "efflux": info["dt"] * float(sum(abs(face[c_, k]) for c_, k in vb.str_faces))}
        rows.append(b)
        mine.append((r, mag))
    f.close()
    return rows, mine


def check_against_restatement(rows, mine):
    cum = {k: 0.0 for k in STEP_VOLS}
    norm = 0.0
    for n, (b, (r, mag)) in enumerate(zip(rows, mine)):
        assert b["steps"] == n + 1
        for k in STEP_VOLS:
            assert abs(b[k] - r[k]) <= 1e-12 * mag[k] + 1e-300, (n, k, b[k], r[k], mag[k])
            cum[k] += b[k]
            assert b["cum_" + k] == cum[k], (n, k)
        norm = max(norm, abs(cum["storage"]))
        assert b["balance"] == b["storage"] + b["efflux"] and b["net"] == b["storage"] + b["flux"]
        assert b["error_pct"] == 100.0 * b["balance"] / max(abs(b["storage"]) + abs(b["efflux"]), 1e-10)
        assert b["cum_error_pct"] == 100.0 * b["cum_balance"] / max(norm, 1e-10)
        for k in ("wse", "u", "v"):
            assert b[k + "_cell"] == r[k + "_cell"] and b[k] == r[k], (n, k, b[k], r[k])
        assert b["courant_cell"] == r["courant_cell"], n
        assert abs(b["courant"] - r["courant"]) <= 1e-15 * r["courant"], n


@pytest.mark.parametrize("name", list(BAL_CASES))
def test_balance_matches_restatement_and_reference(gpu_available, name):
    c = BAL_CASES[name]()
    gold = np.load(os.path.join(GOLD, f"ref64_balance_{name}.npz"))["bal"]
    print(f"volume balance {name}, {c.ncells} cells, {len(gold)} steps", flush=True)
    rows, mine = run_gpu(c, len(gold))
    check_against_restatement(rows, mine)
    w = compare_balance(name, rows, gold)
    print(" against the reference:", w, "cumulative error %", rows[-1]["cum_error_pct"], flush=True)
    assert w["vol"] <= TOL_VOL and w["wse"] <= TOL_ETA, w
    assert w["u"] <= TOL_VEL and w["v"] <= TOL_VEL and w["courant"] <= courant_tol(c), w


def test_balance_variable_time_step(gpu_available):
    """VARIABLE_TIMESTEP halves a step and recomputes it: the balance counts the
    accepted step once, with its time step (the bounds of the case's
    unconverged steps, as for the oracle)."""
    c = dtvar_case()
    gold = np.load(os.path.join(GOLD, "ref64_balance_tidal_dtvar.npz"))["bal"]
    print(f"volume balance, variable time step, {len(gold)} steps", flush=True)
    rows, mine = run_gpu(c, len(gold), dtvar=1)
    assert [r["dt"] for r, _ in mine[5:9]] == [3600.0, 1800.0, 1800.0, 3600.0]
    check_against_restatement(rows, mine)
    w = compare_balance("tidal_dtvar", rows, gold, vols=VOLS[:1])
    print(" against the reference:", w, flush=True)
    assert w["vol"] <= TOL_VOL and w["wse"] <= DTVAR_ETA and w["u"] <= DTVAR_VEL and w["v"] <= DTVAR_VEL, w
    assert w["courant"] <= courant_tol(c, DTVAR_VEL), w


def test_balance_fp32_form(gpu_available):
    """The FP32 refinement form: the same balance from its own fields."""
    c = BAL_CASES["tidal"]()
    print("volume balance, FP32 form", flush=True)
    rows, mine = run_gpu(c, 10, precision=_lib.PREC_FP32_REFINE)
    check_against_restatement(rows, mine)


def test_balance_checkpoint(gpu_available):
    c = BAL_CASES["square"]()
    print("volume balance across a checkpoint", flush=True)
    f = FlowStep(c)
    f.step(6)
    ck = f.checkpoint()
    f.step(4)
    b1 = f.balance()
    f.restore(ck)
    assert f.balance()["steps"] == 6
    f.step(4)
    assert f.balance() == b1
    f.close()
