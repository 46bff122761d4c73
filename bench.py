#!/usr/bin/env python3
"""Throughput of the MI355X implicit CMS-Flow step (implicit cell-steps/s).

Workload: the tidal basin case -- 2000 x 1500 Cartesian cells of 20 m (3.0 M
cells), beach with wetting and drying (hdry 0.05 m), lognormal Manning,
radiation-stress forcing up to 0.05 N/m^2, M2 tide on the east side,
dt = 30 s, FP64 SIMPLEC with BiCGStab solves to rtol 1e-10.  A "step" is one
implicit time step (flow_imp + flow_update) over the whole grid, with every
input resident in HBM.

Prints one JSON line.  With --gpus N > 1 it exits: every box here has one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one launch of the pp-system SpMV (k_bicg_spmv, MODE 1,
# multigrid-preconditioned) per active cell: read aN, aE, aS, aW, ap, r0
# (FP64) and z = M^-1 p (FP32), write v (FP64).
SPMV_BYTES_PER_CELL = 7 * 8 + 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=2000)
    ap.add_argument("--ny", type=int, default=1500)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-refinement form")
    ap.add_argument("--cpu-sample", default="500x375x2",
                    help="oracle sample: NXxNYxSTEPS of the same workload definition")
    return ap.parse_args()


def cpu_baseline(sample):
    """Times the FP64 oracle (single thread) on a bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle  # checker library, used here only as the CPU baseline
    from cms2d_amd import cases
    nx, ny, steps = (int(v) for v in sample.split("x"))
    c = cases.tidal(nx=nx, ny=ny, dx=20.0, seed=1)
    o = Oracle(c, serial=1)
    o.step(1)                     # cold-start step, untimed
    t0 = time.perf_counter()
    rc, _ = o.step(steps)
    dt = time.perf_counter() - t0
    if rc != 0:
        return None
    return {"value": nx * ny * steps / dt, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle (ILU0-BiCGStab, rtol 1e-10, one thread) on the tidal workload "
                      f"generated at {nx}x{ny} cells (dx 20 m, dt 30 s), {steps} steps after one "
                      f"untimed cold-start step, {dt:.1f} s"}


def fp32_form(a, case, g64, value64):
    """The FP32 form (FP32 pressure-correction BiCGStab + SpMV inside FP64
    refinement) on the same workload and step count, timed the same way, and
    its state against the FP64 form's after the same steps."""
    from cms2d_amd import _lib
    from cms2d_amd.flow import FlowStep
    g = FlowStep(case, precision=_lib.PREC_FP32_REFINE)
    for _ in range(a.warmup):
        g.step(1)
    kry = 0.0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        kry += g.step(1)["kry"][0]
    el = time.perf_counter() - t0
    v32 = case.ncells * a.steps / el
    wet = g64.get("iwet") > 0
    wdiff = int((g.get("iwet") != g64.get("iwet")).sum())
    err = {}
    for name in ("eta", "u", "v"):
        d = (g.get(name) - g64.get(name))[wet]
        err[name + "_max"] = float(np.abs(d).max()) if d.size else 0.0
        err[name + "_rms"] = float(np.sqrt(np.mean(d * d))) if d.size else 0.0
    area = case.dx * case.dy
    dvol = float(area * (g.get("h").sum() - g64.get("h").sum()))
    nstep = a.warmup + a.steps
    g.close()
    return {"value": v32, "unit": "cell-steps/s", "ms_per_step": el * 1e3 / a.steps,
            "speedup_vs_fp64": v32 / value64, "pp_inner_krylov_iters_per_step": kry / a.steps,
            "error_vs_fp64_form": err, "wet_cells_differ": wdiff,
            "volume_difference_m3_after_steps": dvol, "steps_compared": nstep,
            "volume_drift_m_per_step": abs(dvol) / (area * case.ncells) / nstep}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: --gpus {a.gpus} not supported: this build runs on one GPU "
              f"(multi-GPU sharding is designed, not built)", file=sys.stderr)
        sys.exit(2)
    from cms2d_amd import cases
    from cms2d_amd.flow import FlowStep
    case = cases.tidal(nx=a.nx, ny=a.ny, dx=20.0, seed=1)
    ncells = case.ncells
    g = FlowStep(case)
    for _ in range(a.warmup):
        g.step(1)
    ms0, n0 = g.kernel_stats("spmv_pp")
    kry = np.zeros(3)
    outer = 0
    t0 = time.perf_counter()
    kernel_ms = 0.0
    for _ in range(a.steps):
        info = g.step(1)
        kernel_ms += info["kernel_ms"]
        kry += np.array(info["kry"])
        outer += info["niter"]
    elapsed = time.perf_counter() - t0
    ms1, n1 = g.kernel_stats("spmv_pp")
    tot_ms = ms1 * n1 - ms0 * n0
    nl = n1 - n0
    avg_ms = tot_ms / nl if nl else float("nan")
    achieved = SPMV_BYTES_PER_CELL * ncells / (avg_ms * 1e-3) / 1e9 if nl else None
    value = ncells * a.steps / elapsed
    out = {
        "metric": "implicit cell-steps/s",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"tidal basin {a.nx}x{a.ny} uniform Cartesian dx=20m (the 2-level "
                        f"telescoping band of the reference case is not built yet), wet/dry, "
                        f"radiation stress, M2 tide, dt=30s, FP64 BiCGStab rtol 1e-10",
            "cells": ncells, "dt_s": case.dt,
            "outer_iters_per_step": outer / a.steps,
            "krylov_iters_per_step": {"pp": kry[0] / a.steps, "u": kry[1] / a.steps,
                                      "v": kry[2] / a.steps},
        },
        "roofline": {
            "kernel": "k_bicg_spmv (pp system, MODE 1: v = A z, z from the multigrid cycle)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": None,
            "avg_launch_ms": avg_ms,
            "launches": nl,
            "bytes_per_launch": SPMV_BYTES_PER_CELL * ncells,
        },
        "cpu_baseline": None,
    }
    if not a.no_fp32:
        out["fp32_refine"] = fp32_form(a, case, g, value)
    g.close()
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
