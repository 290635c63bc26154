"""Build the HIP step library (and the Fortran driver) in-tree for gfx950.

    python taxila-lbm_amd/build.py          # library + Fortran driver

Outputs go to taxila-lbm_amd/build/ (git-ignored, shipped with the tree):
  libtaxila_lbm.so        the C ABI of include/taxila_lbm.h
  txg_bindings.mod etc.   ISO_C_BINDING module of the Fortran driver
  taxila_gpu              the Fortran-90 driver executable
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
OUT = os.path.join(HERE, "build")
INC = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FLANG = os.path.join(ROCM, "bin", "amdflang")

# Packed FP32 instructions are disabled on every HIP translation unit.
COMMON = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", INC, "-I", CSRC]

UNITS = [
    # (source, extra flags): the FP64 unit must not fuse multiply-adds; the
    # FP32 units fuse a*b+c only within one source expression, so a kernel's
    # rounding does not depend on where it is inlined (tile shape, batched
    # or single launch, one-pass or two-pass)
    ("txg_fp64.hip", ["-ffp-contract=off"]),
    ("txg_fp32.hip", ["-ffp-contract=on"]),
    ("txg_fp16.hip", ["-ffp-contract=on"]),
    ("txg_api.hip", ["-ffp-contract=off"]),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_library(jobs: int = 4) -> str:
    os.makedirs(OUT, exist_ok=True)
    objs = []
    cmds = []
    for src, extra in UNITS:
        obj = os.path.join(OUT, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if out:
                print(out, file=sys.stderr)
    lib = os.path.join(OUT, "libtaxila_lbm.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, lib)
    return lib


def build_fortran() -> str | None:
    """Compile the ISO_C_BINDING module and the Fortran-90 driver with flang."""
    if not os.path.exists(FLANG) or not os.path.isdir(FSRC):
        return None
    srcs = [os.path.join(FSRC, f) for f in ("txg_bindings.F90", "txg_options.F90", "txg_flow.F90",
                                             "default_callbacks.F90", "taxila_gpu.F90")]
    srcs = [s for s in srcs if os.path.exists(s)]
    if not srcs:
        return None
    exe = os.path.join(OUT, "taxila_gpu")
    cmd = [FLANG, "-O2", "-module-dir", OUT, "-I", OUT, "-o", exe] + srcs + [
        "-L", OUT, "-ltaxila_lbm", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    return exe


def build(jobs: int = 4) -> None:
    build_library(jobs)
    build_fortran()


if __name__ == "__main__":
    build()
    print("built", os.path.join(OUT, "libtaxila_lbm.so"))
