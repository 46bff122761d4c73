#!/usr/bin/env python3
"""Compiles the reference's own Fortran sources, as shipped, into oracle/_ref/.

TEST INFRASTRUCTURE ONLY (the reference build is a checker for the oracle).

* Sources: the file list of the reference's source/CMakeLists.txt, compiled
  where they lie under the reference tree (read-only), with AMD flang, the
  reference's preprocessor header (source/main/CMS_cpp.h) and no edits.  The
  shipped precision is kept (prec_def ikind=4: single precision).
* Fortran module dependencies are resolved by repeated passes: each pass
  compiles the files whose modules are now available.
* With --double (python oracle/build_ref.py --double) the same sources are
  built in double precision: the tree is copied to a temporary directory
  outside this repository, prec_def's ikind is set to 8
  (source/main/definitions.F90:7), and the two calls that pass single-precision
  latitude/longitude arrays to the real(ikind) argument of writescalTxt
  (source/boundaries/boundaries.F90:3140,3145, an ASCII converter of the HDF5
  inputs that the driver never reaches) pass real(lat, ikind) instead, and so
  does one more call of the same kind in the tools utility
  (source/utilities/tools.F90:675, whose module also holds string helpers the
  card readers use).  Two files pass real(ikind) arrays to the
  single-precision XMDF interface and do not compile at ikind=8: the dredging
  setup (source/dredge/dredge_setup.F90:749,816,855) and CMS-Wave's HDF5
  output; neither is on the flow step's path, they stay out of this build, and
  the driver then sets dredging off itself instead of calling dredge_default.  Outputs:
  oracle/_ref/libcmsref8.so and oracle/_ref/ref_run8.
* Outputs (objects, .mod files, the shared library) go to oracle/_ref/ only.
  The library is linked with the driver oracle/ref_driver.f90 (ours).  The
  XMDF/HDF5 C library the reference's I/O wraps is absent from this image; no
  stand-in is written for it -- its entry points stay unresolved in the shared
  library and the driver never reaches them (it builds grids and boundary
  strings in memory and calls only the flow, solver and geometry routines).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import subprocess
import sys

import shutil
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("CMS2D_REF", "/root/reference")
SRC = os.path.join(REF, "source")
OUT = os.path.join(HERE, "_ref")
DOUBLE = "--double" in sys.argv[1:]
SUFFIX = "8" if DOUBLE else ""
# the edits of the double-precision copy: (file, text, replacement, count)
DOUBLE_EDITS = [
    ("main/definitions.F90", "integer, parameter :: ikind=4", "integer, parameter :: ikind=8", 1),
    ("boundaries/boundaries.F90", "trim(dsetname), lat, 1)", "trim(dsetname), real(lat, ikind), 1)", 1),
    ("boundaries/boundaries.F90", "trim(dsetname), lon, 1)", "trim(dsetname), real(lon, ikind), 1)", 1),
    ("utilities/tools.F90", "'Maximum_WSE','maxeta',max_array)", "'Maximum_WSE','maxeta',real(max_array, 8))", 1),
]
OBJ = os.path.join(OUT, "obj" + SUFFIX)
MOD = os.path.join(OUT, "mod" + SUFFIX)
FC = os.environ.get("FC_REF", "/opt/rocm/lib/llvm/bin/flang")
FFLAGS = []


def set_flags():
    global FFLAGS
    FFLAGS = ["-cpp", "-fPIC", "-O2", "-fopenmp", "-I", os.path.join(SRC, "main"),
              "-module-dir", MOD, "-I", MOD, "-w"]


def double_copy():
    """Copies the reference's source tree to a temporary directory and applies
    DOUBLE_EDITS there; returns the copy's source directory."""
    tmp = tempfile.mkdtemp(prefix="cmsref8_")
    dst = os.path.join(tmp, "source")
    shutil.copytree(SRC, dst)
    for rel, old, new, count in DOUBLE_EDITS:
        path = os.path.join(dst, rel)
        txt = open(path).read()
        if txt.count(old) != count:
            raise SystemExit(f"{rel}: expected {count} occurrence(s) of {old!r}")
        open(path, "w").write(txt.replace(old, new))
    return dst
# The program unit is not needed in a library.
SKIP = {"main/CMS2D_main.f90"}


def source_list():
    txt = open(os.path.join(SRC, "CMakeLists.txt")).read()
    files = re.findall(r"^\s*\./(\S+\.[fF]90)\s*$", txt, flags=re.M)
    return [f for f in files if f not in SKIP]


def obj_of(rel):
    return os.path.join(OBJ, rel.replace("/", "__") + ".o")


def compile_one(rel):
    src = os.path.join(SRC, rel)
    o = obj_of(rel)
    if os.path.exists(o) and os.path.getmtime(o) >= os.path.getmtime(src):
        return rel, True, ""
    r = subprocess.run([FC, *FFLAGS, "-c", src, "-o", o], capture_output=True, text=True)
    return rel, r.returncode == 0, r.stderr[-2000:]


def main():
    global SRC
    if not os.path.isdir(SRC):
        print(f"reference sources not found under {SRC}: nothing to build", file=sys.stderr)
        return 0
    if DOUBLE:
        SRC = double_copy()
    set_flags()
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(MOD, exist_ok=True)
    todo = source_list()
    errs = {}
    npass = 0
    while todo:
        npass += 1
        with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            res = list(ex.map(compile_one, todo))
        left = [r for r, ok, _ in res if not ok]
        errs = {r: e for r, ok, e in res if not ok}
        print(f"pass {npass}: {len(todo) - len(left)} compiled, {len(left)} pending", flush=True)
        if len(left) == len(todo):
            break
        todo = left
    with open(os.path.join(OUT, f"build_ref{SUFFIX}.log"), "w") as f:
        for r, e in errs.items():
            f.write(f"=== {r}\n{e}\n")
    drv = os.path.join(HERE, "ref_driver.f90")
    lib = os.path.join(OUT, f"libcmsref{SUFFIX}.so")
    objs = [obj_of(r) for r in source_list() if os.path.exists(obj_of(r))]
    subprocess.run([FC, "-shared", "-fopenmp", "-o", lib + ".tmp", *objs], check=True)
    os.replace(lib + ".tmp", lib)
    print(f"linked {lib} from {len(objs)} reference objects", flush=True)
    if os.path.exists(drv):
        dobj = os.path.join(OBJ, "ref_driver.o")
        exe = os.path.join(OUT, f"ref_run{SUFFIX}")
        dflags = ["-DREF_NO_DREDGE"] if DOUBLE else []
        subprocess.run([FC, *FFLAGS, *dflags, "-c", drv, "-o", dobj], check=True)
        subprocess.run([FC, "-fopenmp", "-o", exe, dobj, "-L", OUT, f"-lcmsref{SUFFIX}",
                        "-Wl,-rpath,$ORIGIN", "-Wl,--allow-shlib-undefined"], check=True)
        print(f"linked {exe}", flush=True)
    if DOUBLE:
        shutil.rmtree(os.path.dirname(SRC), ignore_errors=True)
    if errs:
        print(f"{len(errs)} reference files did not compile (see {OUT}/build_ref.log): "
              + ", ".join(sorted(errs)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
