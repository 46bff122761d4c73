"""Builds the native parts of cms2d_amd in-tree.

* cms2d_amd/libcms2d_hip.so -- the HIP implementation behind include/cms2d_amd.h,
  compiled for gfx950 by hipcc.
* cms2d_amd/fortran/ -- the Fortran host layer (ISO_C_BINDING module and the
  drop-in solve/flow_imp wrappers), compiled by flang when it is present.

Run ``python -m cms2d_amd.build`` (or __graft_entry__.build()).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcms2d_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CMS2D_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-shared", "-std=c++17",
    # IEEE expression-by-expression arithmetic: no contraction into fma, so the
    # non-iterative kernels match the FP64 oracle operation for operation
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip():
    src = os.path.join(CSRC, "cms2d.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, src])
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    fdir = os.path.join(HERE, "fortran")
    mk = os.path.join(fdir, "Makefile")
    if not os.path.exists(mk):
        return None
    flang = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    if not os.path.exists(flang):
        print("flang not found: Fortran host layer not built", flush=True)
        return None
    _run(["make", "-s", "-C", fdir, f"FC={flang}"])
    return fdir


def build_all():
    build_hip()
    build_fortran()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what or "hip" in what:
        build_hip()
    if "all" in what or "fortran" in what:
        build_fortran()
