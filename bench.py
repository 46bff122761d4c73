#!/usr/bin/env python3
"""Throughput of the MI355X implicit CMS-Flow step (implicit cell-steps/s).

Workload: the tidal case on a telescoping grid -- 2000 x 1500 base cells of
20 m with a shore-parallel band of 10 m cells around a 5 m core across the
beach (3.61 M cells), wetting and drying (hdry 0.05 m), lognormal Manning,
radiation-stress forcing up to 0.05 N/m^2, M2 tide on the east side,
dt = 30 s, FP64 SIMPLEC with BiCGStab solves to rtol 1e-10.  The same physics
on the uniform 2000 x 1500 grid is reported beside it.  A "step" is one
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
# Kernel groups timed live (per launch, stream events stamped by the dispatch;
# mg_coarse is the span between events recorded on the stream around a
# V-cycle's levels below level 0) with their algorithmic bytes per launch:
#   spmv_pp   k_bicg_spmv MODE 1 (v = A z) and MODE 2 (t = A zs): read
#             aN,aE,aS,aW,ap,r0 (FP64) and z (FP32), write v (FP64): 60 B per
#             cell; on the cell list the face metadata word, per face the
#             neighbour index and coefficient, ap, r0, z, v: 36 + 12 F B
#   mg_down0  k_mg_down on level 0: read b (FP64, FP32 on telescoping grids),
#             aN..aW, 1/ap (FP32), write z (FP32) and a quarter-size coarse b:
#             8 + 5*4 + 4 + 1 = 33 B per base cell (29 with FP32 b)
#   mg_up0    k_mg_up on level 0: read b, aN..aW, 1/ap, z and a quarter-size
#             coarse z, write z: 8 + 6*4 + 1 + 4 = 37 B per base cell (33)
#   mg_coarse the levels below level 0 of one V-cycle: down and up sweeps of
#             level l read and write 29 + 33 = 62 B per level-l cell
#   tel_restrict    residual r - A Z summed into the base cells: per cell the
#             list entry, metadata, r, base cell and its range (28 B), on cells
#             next to refined cells also ap, Z and per face index, coefficient
#             and Z (12 + 16 F B), 4 B per base cell written
#   tel_fjac  damped Jacobi sweep on the finest cells: list entry, metadata,
#             r, ap, Z in and out, per face index and coefficient: 36 + 12 F B
#   tel_cs_restrict residual summed into the 10 m squares: per cell of level
#             >= 1 its entry, square, metadata, r, ap, Z and per face index,
#             coefficient and Z (36 + 16 F B), per square 16 B
# HBM bytes per active cell of one launch as measured by rocprofv3 counters
# (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, factors from a copy kernel of known
# size) on the uniform 2000 x 1500 workload; the commands and the table are in
# profiles/traffic_tidal_2000x1500.md.
TRAFFIC_PER_CELL = {"spmv_pp": (166.8 + 24.5) * 1e6 / 3.0e6,
                    "mg_down0": (114.3 + 15.2) * 1e6 / 3.0e6,
                    "mg_up0": (128.1 + 16.9) * 1e6 / 3.0e6}
# The same per launch on the telescoping workload at 2000 x 1500
# (profiles/traffic_tel_2000x1500.md).
TEL_TRAFFIC = {"spmv_pp": 341.6e6, "mg_down0": 111.8e6, "mg_up0": 128.3e6, "tel_restrict": 174.9e6,
               "tel_cs_restrict": 74.0e6, "tel_fjac": 58.3e6}
KERNEL_DESC = {
    "spmv_pp": "k_bicg_spmv MODE 1 and 2 (pp system: v = A z, t = A zs, z from the preconditioner)",
    "mg_down0": "k_mg_down (multigrid level 0: red-black pre-smoothing + restriction)",
    "mg_up0": "k_mg_up (multigrid level 0: correction + red-black post-smoothing)",
    "mg_coarse": "multigrid levels 1.. of a V-cycle (k_mg_down_pc, k_mg_coarsest, k_mg_up_pc; span)",
    "tel_restrict": "k_tl_mg_restrict (residual summed into the base cells)",
    "tel_fjac": "k_tl_cs_fjac (damped Jacobi sweep on the 5 m cells)",
    "tel_cs_restrict": "k_tl_cs_restrict (residual summed into the 10 m squares)",
}


def mg_level_cells(nx, ny):
    """Cells of the multigrid levels below level 0 (cell-centred 2x2
    coarsening to at most 4096 cells, at least four levels)."""
    out = []
    while True:
        nx, ny = (nx + 1) // 2, (ny + 1) // 2
        out.append(nx * ny)
        if nx * ny <= 4096 and len(out) >= 3:
            return out


def cartesian_groups(case):
    n = case.nx * case.ny
    coarse = 62 * sum(mg_level_cells(case.nx, case.ny)[:-1])
    return {"spmv_pp": 60 * n, "mg_down0": 33 * n, "mg_up0": 37 * n, "mg_coarse": coarse}


def tel_groups(case):
    """Algorithmic bytes per launch of the timed groups on a telescoping grid."""
    t = case.tel
    nb = np.asarray(t.nb)
    lev = np.asarray(t.level)
    n = len(nb)
    nbase = case.nx * case.ny
    faces = (nb > 0).sum(axis=1) + ((nb[:, 0::2] == 0) & (nb[:, 1::2] == 0)).sum(axis=1)
    refined = lev > 0
    nbr = np.where(nb > 0, nb - 1, 0)
    near = refined | (refined[nbr] & (nb > 0)).any(axis=1)
    top = lev == lev.max() if lev.max() > 0 else np.zeros(n, bool)
    squares = 0
    if lev.max() >= 2:
        g = np.asarray(t.geom)[lev >= 1]
        h = case.dx / 2
        squares = len(np.unique(np.floor(g[:, 1] / h) * (2 * case.nx) + np.floor(g[:, 0] / h)))
    return {
        "spmv_pp": int((36 + 12 * faces).sum()),
        "mg_down0": 29 * nbase, "mg_up0": 33 * nbase,
        "mg_coarse": 62 * sum(mg_level_cells(case.nx, case.ny)[:-1]),
        "tel_restrict": int(28 * n + 4 * nbase + (12 + 16 * faces[near]).sum()),
        "tel_fjac": int((36 + 12 * faces[top]).sum()),
        "tel_cs_restrict": int((36 + 16 * faces[lev >= 1]).sum() + 16 * squares),
    }


# Whole-step byte model (FP64, per active cell): once per step 358 (static
# sources, wet/dry, velface_eta, level rotation, mass balance); per outer
# iteration 1156 (momentum assembly 193, three solve prologues 80+80+64, three
# ILU0 factorisations 3*48, pp assembly 185, pvfc 168, gradients/interpolation
# 105, friction 64, eddy viscosity 73); per BiCGStab iteration of any of the
# three systems 376 (two SpMVs 7*8 each, two preconditioner applications 9*8
# each, vector updates and dot products 15*8).
STEP_MODEL = {"per_step": 358, "per_outer": 1156, "per_krylov_iter": 376}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=2000)
    ap.add_argument("--ny", type=int, default=1500)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-refinement form")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched ensemble")
    ap.add_argument("--no-cartesian", action="store_true",
                    help="skip the uniform-grid comparison (and the FP32 form, built there)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the discrete adjoint")
    ap.add_argument("--adj-n", type=int, default=1000, help="adjoint workload: cells a side")
    ap.add_argument("--adj-steps", type=int, default=60,
                    help="adjoint workload: time steps (the full workload is 720)")
    ap.add_argument("--adj-every", type=int, default=20, help="adjoint: checkpoint interval (steps)")
    ap.add_argument("--adj-fd-cells", type=int, default=2,
                    help="adjoint: cells checked by central differences (four to eight forward runs each per field)")
    ap.add_argument("--adj-rtol", type=float, default=1e-10, help="adjoint workload: Krylov rtol")
    ap.add_argument("--adj-check-rtol", type=float, default=1e-12,
                    help="adjoint: Krylov rtol of the runs checked by central differences")
    ap.add_argument("--ens-members", type=int, default=256)
    ap.add_argument("--ens-steps", type=int, default=3)
    ap.add_argument("--ens-threads", type=int, default=8)
    ap.add_argument("--cpu-sample", default="500x375x2",
                    help="oracle sample: NXxNYxSTEPS of the same workload definition")
    ap.add_argument("--ref-sample", default="500x375x3",
                    help="compiled-reference sample: NXxNYxSTEPS of the same workload definition")
    ap.add_argument("--ref-threads", type=int, default=min(16, os.cpu_count() or 1))
    return ap.parse_args()


def cpu_baseline(sample):
    """Times the FP64 oracle (single thread) on a bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle  # checker library, used here only as the CPU baseline
    from cms2d_amd import cases
    nx, ny, steps = (int(v) for v in sample.split("x"))
    c = cases.tidal_tel(nx=nx, ny=ny, dx=20.0, seed=1)
    o = Oracle(c, serial=1)
    o.step(1)                     # cold-start step, untimed
    t0 = time.perf_counter()
    rc, _ = o.step(steps)
    dt = time.perf_counter() - t0
    if rc != 0:
        return None
    return {"value": c.ncells * steps / dt, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle (ILU0-BiCGStab, rtol 1e-10, one thread) on the telescoping tidal workload "
                      f"generated at {nx}x{ny} base cells ({c.ncells} cells, dt 30 s), {steps} steps after "
                      f"one untimed cold-start step, {dt:.1f} s"}


def cpu_reference(sample, threads):
    """Times the reference's own implicit step (oracle/_ref/ref_run: the
    reference's Fortran sources compiled unmodified, single precision and its
    own solver settings as shipped, OpenMP) on a bounded sample of the same
    workload; the first step (cold start) is not counted."""
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import refcase  # writes the reference's ASCII inputs for a case
    from cms2d_amd import cases
    if not refcase.available():
        return None
    nx, ny, steps = (int(v) for v in sample.split("x"))
    c = cases.tidal_tel(nx=nx, ny=ny, dx=20.0, seed=1)
    with tempfile.TemporaryDirectory() as d:
        d = d + "/"
        refcase.write_case(c, d, "tidal")
        force = os.path.join(d, "forcing.bin")
        np.concatenate([np.asarray(c.wavestrx, np.float64), np.asarray(c.wavestry, np.float64)]).tofile(force)
        # the reference keeps work arrays on the stack (automatic arrays sized
        # by the cell count): unlimited main stack, large OpenMP thread stacks
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_STACKSIZE="512M")

        def stack():
            import resource
            resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

        r = subprocess.run([refcase.REF_RUN, d, "tidal", str(steps + 1), "-", force], cwd=d, env=env,
                           capture_output=True, text=True, timeout=600, preexec_fn=stack)
    t = [float(l.split()[2]) for l in r.stdout.splitlines() if l.startswith("step_seconds")]
    if r.returncode != 0 or len(t) != steps + 1:
        return None
    el = sum(t[1:])
    return {"value": c.ncells * steps / el, "unit": "cell-steps/s", "cores": threads, "kind": "reference",
            "sample": f"the reference's flow_imp/flow_update (its Fortran compiled unmodified with flang -O2 "
                      f"-fopenmp, single precision and BiCGStab-ILUT settings as shipped) on the telescoping "
                      f"tidal workload generated at {nx}x{ny} base cells ({c.ncells} cells, read from a .tel "
                      f"file), {steps} steps after one untimed cold-start step, {threads} OpenMP threads, "
                      f"{el:.1f} s"}


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
    tel = getattr(case, "tel", None)
    area = tel.areas() if tel is not None else np.full(case.ncells, case.dx * case.dy)
    edge = tel.edge_faces() if tel is not None else None
    dvol = float((area * (g.get("h") - g64.get("h"))).sum())
    nstep = a.warmup + a.steps

    def balance(m):
        # one more step: volume change minus dt times the boundary inflow (m^3)
        v0 = (m.get("h") * area).sum()
        m.step(1)
        if tel is not None:
            q = -m.get_face("flux")[edge].sum()
        else:
            f = m.get_face("flux").reshape(case.ny, case.nx, 4)
            q = -(f[-1, :, 0].sum() + f[:, -1, 1].sum() + f[0, :, 2].sum() + f[:, 0, 3].sum())
        return float((m.get("h") * area).sum() - v0 - case.dt * q), float(case.dt * q)

    d32, q32 = balance(g)
    d64, q64 = balance(g64)
    g.close()
    return {"value": v32, "unit": "cell-steps/s", "ms_per_step": el * 1e3 / a.steps,
            "speedup_vs_fp64": v32 / value64, "pp_inner_krylov_iters_per_step": kry / a.steps,
            "error_vs_fp64_form": err, "wet_cells_differ": wdiff,
            "volume_difference_m3_after_steps": dvol, "steps_compared": nstep,
            "volume_drift_m_per_step": abs(dvol) / area.sum() / nstep,
            "mass_balance_one_step": {"fp32_form": {"dV_minus_dtQ_m3": d32, "dtQ_m3": q32},
                                      "fp64_form": {"dV_minus_dtQ_m3": d64, "dtQ_m3": q64}}}


def ensemble_form(a):
    """The batched ensemble: independent basins (sides uniform in 64-512 cells,
    FP32 form, dt 60 s) advanced side by side; aggregate cell-steps/s."""
    from cms2d_amd import _lib, cases
    from cms2d_amd.ensemble import Ensemble
    ms = cases.ensemble_members(n=a.ens_members, seed=11)
    e = Ensemble(ms, nthreads=a.ens_threads, precision=_lib.PREC_FP32_REFINE)
    e.step(1)
    t0 = time.perf_counter()
    infos = e.step(a.ens_steps)
    el = time.perf_counter() - t0
    ok = all(i["iconv"] in (1, 2) for i in infos)
    tiles = {}
    for m in e.members:
        t, gr = m.tuning()
        key = f"{t}{'+graphs' if gr else ''}"
        tiles[key] = tiles.get(key, 0) + 1
    out = {"value": e.ncells * a.ens_steps / el, "unit": "cell-steps/s", "members": len(ms),
           "cells": e.ncells, "steps": a.ens_steps, "seconds": el, "host_threads": a.ens_threads,
           "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "runtime default"),
           "form": "FP32 Krylov + FP64 refinement, dt 60 s", "all_converged": ok,
           "members_per_tuning": tiles,
           "note": "members run on host threads with one stream each (not one persistent launch); "
                   "per member, the level-0 multigrid tile width timed at creation and graph replay "
                   "of BiCGStab iteration pairs (members up to 1M cells)"}
    e.close()
    return out


def adjoint_form(a):
    """The discrete adjoint: a 1000 x 1000 grid of 20 m cells (the tidal
    physics without radiation stress, dt 60 s), 16 gauges at random wet
    cells, observations from a hidden seeded Manning field (lognormal,
    sigma_log 0.2, about the case's), loss = sum over steps and gauges of the
    squared misfit; dL/dn and dL/dzb of every cell from cms2d_flow_adjoint
    with checkpoints every adj_every steps, timed against the forward run of
    the same steps, and checked by central differences (h = 1e-6 and 5e-7,
    else 1e-5 and 5e-6) at random wet cells where the gradient is resolved
    (|g| >= 1e-3 max|g|).  The
    derivative of the exponential scheme below its 1e-6 guard is that of the
    unguarded form (its limit 1 - P/2), so the gradient is the forward run's
    except on that 1e-6-wide band of each face's Peclet number."""
    import copy
    from cms2d_amd import cases
    from cms2d_amd.flow import FlowStep
    n, N, K = a.adj_n, a.adj_steps, a.adj_every
    c = cases.square(nx=n, ny=n, dx=20.0, seed=3, dt=60.0)
    rng = np.random.default_rng(17)
    wet = np.nonzero(-c.zb > 1.0)[0]
    gauges = np.sort(rng.choice(wet, 16, replace=False))
    kw = dict(rtol=a.adj_rtol, atol=0.0)

    def gauge_run(case):
        f = FlowStep(case, **kw)
        eta = np.zeros((N, len(gauges)))
        for k in range(N):
            f.step(1)
            eta[k] = f.get("eta")[gauges]
        f.close()
        return eta

    hidden = copy.copy(c)
    hidden.mann = (c.mann * np.exp(0.2 * rng.standard_normal(c.mann.shape))).astype(np.float32).astype(np.float64)
    obs = gauge_run(hidden)
    progress("adjoint: observations from the hidden Manning field done")
    f = FlowStep(c, **kw)
    t0 = time.perf_counter()
    f.step(N)
    fwd = time.perf_counter() - t0
    f.close()
    f = FlowStep(c, **kw)
    t0 = time.perf_counter()
    _, loss, gm, gz = f.adjoint(N, gauges, obs=obs, every=K)
    adj = time.perf_counter() - t0
    f.close()
    progress(f"adjoint: {N} steps forward {fwd:.2f} s, adjoint {adj:.2f} s")
    # the check: forward and adjoint with the solves at adj_check_rtol, whose
    # stopping error (~rtol relative, changing when a perturbation changes a
    # solve's iteration count) stays below what a difference at h = 1e-6 resolves
    kw = dict(rtol=a.adj_check_rtol, atol=0.0)
    f = FlowStep(c, **kw)
    _, _, gm, gz = f.adjoint(N, gauges, obs=obs, every=K)
    f.close()
    progress(f"adjoint: gradient at rtol {a.adj_check_rtol:g} for the check done")
    fd = {"h": "1e-6 and 5e-7, else 1e-5 and 5e-6", "n": [], "zb": []}
    for field, grad, base in (("n", gm, c.mann), ("zb", gz, c.zb)):
        resolved = np.nonzero(np.abs(grad) >= 1e-3 * np.abs(grad).max())[0]
        cand = np.intersect1d(resolved, wet)
        for k in rng.choice(cand, min(a.adj_fd_cells, len(cand)), replace=False):
            # central differences at h and h/2 for h = 1e-6, then 1e-5: the
            # first pair that agrees to 1e-3 (a step across a discrete switch
            # of the forward run, or one too small for the loss's rounding,
            # makes the two disagree); the loss differences are formed as
            # sum (e+ - e-)(e+ + e- - 2 obs), free of the cancellation of two
            # nearly equal sums
            best = None
            for h in (1e-6, 1e-5):
                cen = []
                for hh in (h, h / 2):
                    e = []
                    for sgn in (1.0, -1.0):
                        c2 = copy.copy(c)
                        arr = base.copy()
                        arr[k] += sgn * hh
                        setattr(c2, "mann" if field == "n" else "zb", arr)
                        e.append(gauge_run(c2))
                        progress(f"adjoint: run with {field}[{int(k)}] {sgn * hh:+g} done")
                    cen.append(((e[0] - e[1]) * (e[0] + e[1] - 2.0 * obs)).sum() / (2.0 * hh))
                agree = abs(cen[0] - cen[1]) / abs(cen[1])
                if best is None or agree < best[0]:
                    best = (agree, cen[1], h / 2)
                if agree <= 1e-3:
                    break
            spread, d, h = best
            progress(f"adjoint: difference at cell {int(k)} of {field} done")
            fd[field].append({"cell": int(k), "adjoint": float(grad[k]), "difference": float(d), "h": h,
                              "h_vs_2h_agreement": float(spread), "rel_err": float(abs(grad[k] - d) / abs(d))})
        fd[f"resolved_cells_{field}"] = int(len(resolved))
    progress("adjoint: central differences done")
    errs = [e["rel_err"] for f_ in ("n", "zb") for e in fd[f_]]
    return {"workload": f"{n}x{n} cells of 20 m (tidal physics without radiation stress, wet/dry), dt 60 s, "
                        f"16 gauges, loss = sum of squared gauge misfit against a hidden lognormal Manning field",
            "cells": n * n, "steps": N, "checkpoint_every": K, "rtol": a.adj_rtol,
            "check_rtol": a.adj_check_rtol, "loss": loss,
            "forward_s": fwd, "adjoint_s": adj, "adjoint_over_forward": adj / fwd,
            "adjoint_cell_steps_per_s": n * n * N / adj,
            "grad_norm": {"n": float(np.linalg.norm(gm)), "zb": float(np.linalg.norm(gz))},
            "central_differences": fd, "max_rel_err": max(errs) if errs else None}


def progress(msg):
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def measure(a, case, label, groups):
    """Times a.steps steps of the FP64 form on `case` after a.warmup; the live
    per-launch times of the kernel groups; the whole-step byte model."""
    from cms2d_amd.flow import FlowStep
    ncells = case.ncells
    g = FlowStep(case)
    mg_tile, graphs = g.tuning()
    for _ in range(a.warmup):
        g.step(1)
    st0 = {k: g.kernel_stats(k) for k in groups}
    kry = np.zeros(3)
    outer = 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        info = g.step(1)
        kry += np.array(info["kry"])
        outer += info["niter"]
    elapsed = time.perf_counter() - t0
    progress(f"{label}: {a.steps} steps of {ncells} cells in {elapsed:.2f} s")
    kern = {}
    for k, nbytes in groups.items():
        ms1, n1 = g.kernel_stats(k)
        ms0, n0 = st0[k]
        nl = n1 - n0
        tot = ms1 * n1 - ms0 * n0
        avg = tot / nl if nl else None
        kern[k] = {"kernel": KERNEL_DESC[k], "launches": nl, "total_ms": tot, "avg_launch_ms": avg,
                   "share_of_step": tot / (elapsed * 1e3),
                   "bytes_per_launch": nbytes,
                   "achieved_GBs": nbytes / (avg * 1e-3) / 1e9 if avg else None}
    # the roofline kernel: the single kernel with the largest total time
    # (mg_coarse is a span of nine launches, reported beside it)
    dom = max((k for k in kern if k != "mg_coarse"), key=lambda k: kern[k]["total_ms"])
    # flow_balance and flow_step_stat as the device keeps them (cms2d_flow_balance):
    # balance = storage + efflux = dV - dt Q_in of a step [m^3]
    b = g.balance()
    vbal = {"last_step": {k: b[k] for k in ("storage", "efflux", "balance", "error_pct")},
            "cumulative": {k: b["cum_" + k] for k in ("storage", "efflux", "balance", "error_pct")},
            "steps": b["steps"], "mean_balance_m3_per_step": b["cum_balance"] / max(b["steps"], 1),
            "extremes_last_step": {k: (b[k], b[k + "_cell"]) for k in ("wse", "u", "v", "courant")}}
    value = ncells * a.steps / elapsed
    model_b = (STEP_MODEL["per_step"] + STEP_MODEL["per_outer"] * outer / a.steps
               + STEP_MODEL["per_krylov_iter"] * kry.sum() / a.steps)
    model_gbs = model_b * value / 1e9
    res = {"value": value, "ms_per_step": elapsed * 1e3 / a.steps, "cells": ncells, "label": label,
           "mg_tile_x": mg_tile, "graph_replay": graphs, "outer_iters_per_step": outer / a.steps,
           "krylov_iters_per_step": {"pp": kry[0] / a.steps, "u": kry[1] / a.steps, "v": kry[2] / a.steps},
           "kern": kern, "dominant": dom, "volume_balance": vbal,
           "step_model_roofline": {
               "bytes_per_cell_step": model_b, "achieved": model_gbs, "peak": HBM_PEAK_GBS,
               "unit": "GB/s", "frac": model_gbs / HBM_PEAK_GBS,
               "model": "whole step, FP64 bytes per active cell: 358 per step + 1156 per outer "
                        "iteration + 376 per BiCGStab iteration (u, v and pp), with this run's "
                        "iteration counts, at the measured cell-steps/s"}}
    return g, res


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: --gpus {a.gpus} not supported: this build runs on one GPU "
              f"(multi-GPU sharding is designed, not built)", file=sys.stderr)
        sys.exit(2)
    from cms2d_amd import cases
    # CPU baselines first, before this process touches the GPU (the compiled
    # reference runs as a child process)
    cpu = port = None
    if not a.no_cpu_baseline:
        cpu = cpu_reference(a.ref_sample, a.ref_threads)
        progress("compiled-reference CPU sample done")
        port = cpu_baseline(a.cpu_sample)
        progress("oracle CPU sample done")
    tel_case = cases.tidal_tel(nx=a.nx, ny=a.ny, dx=20.0, seed=1)
    progress(f"telescoping grid built: {tel_case.ncells} cells")
    g, T = measure(a, tel_case, "telescoping", tel_groups(tel_case))
    fp32_tel = None
    if not a.no_fp32:
        fp32_tel = fp32_form(a, tel_case, g, T["value"])
        fp32_tel["workload"] = "the telescoping workload above (3.61 M cells)"
        progress("telescoping FP32 form done")
    g.close()
    D = T["kern"][T["dominant"]]
    achieved = D["achieved_GBs"]
    lev = tel_case.tel.level
    out = {
        "metric": "implicit cell-steps/s",
        "value": T["value"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": T["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"tidal case on a telescoping grid: {a.nx}x{a.ny} base cells of 20 m with a "
                        f"shore-parallel band of 10 m cells around a 5 m core, {T['cells']} cells; "
                        f"wet/dry, radiation stress, M2 tide, dt=30s, FP64 BiCGStab rtol 1e-10",
            "cells": T["cells"], "cells_per_level": {"20m": int((lev == 0).sum()), "10m": int((lev == 1).sum()),
                                                     "5m": int((lev == 2).sum())},
            "dt_s": tel_case.dt, "outer_iters_per_step": T["outer_iters_per_step"],
            "krylov_iters_per_step": T["krylov_iters_per_step"],
        },
        "roofline": {
            "kernel": D["kernel"],
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": TEL_TRAFFIC.get(T["dominant"]) if (a.nx, a.ny) == (2000, 1500) else None,
            "avg_launch_ms": D["avg_launch_ms"],
            "launches": D["launches"],
            "bytes_per_launch": D["bytes_per_launch"],
            "timed_groups": T["kern"],
        },
        "step_model_roofline": T["step_model_roofline"],
        "volume_balance": T["volume_balance"],
        "cpu_baseline": None,
    }
    if fp32_tel is not None:
        out["fp32_refine"] = fp32_tel
    if not a.no_cartesian:
        case = cases.tidal(nx=a.nx, ny=a.ny, dx=20.0, seed=1)
        g, C = measure(a, case, "cartesian", cartesian_groups(case))
        Dc = C["kern"][C["dominant"]]
        out["cartesian_uniform"] = {
            "workload": f"the same physics on the uniform {a.nx}x{a.ny} grid of 20 m cells",
            "value": C["value"], "unit": "cell-steps/s", "ms_per_step": C["ms_per_step"], "cells": C["cells"],
            "mg_tile_x_autotuned": C["mg_tile_x"], "outer_iters_per_step": C["outer_iters_per_step"],
            "krylov_iters_per_step": C["krylov_iters_per_step"],
            "roofline": {"kernel": Dc["kernel"], "achieved": Dc["achieved_GBs"], "peak": HBM_PEAK_GBS,
                         "frac": Dc["achieved_GBs"] / HBM_PEAK_GBS if Dc["achieved_GBs"] else None,
                         "traffic": TRAFFIC_PER_CELL[C["dominant"]] * C["cells"] if C["dominant"] in TRAFFIC_PER_CELL
                         else None, "timed_groups": C["kern"]},
            "step_model_roofline": C["step_model_roofline"], "volume_balance": C["volume_balance"]}
        if not a.no_fp32:
            out["cartesian_uniform"]["fp32_refine"] = fp32_form(a, case, g, C["value"])
        g.close()
    if not a.no_ensemble:
        out["ensemble"] = ensemble_form(a)
    if not a.no_adjoint:
        out["adjoint"] = adjoint_form(a)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu if cpu is not None else port
        out["cpu_baseline_port"] = port
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
