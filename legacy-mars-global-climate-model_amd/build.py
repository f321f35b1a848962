"""Builds the MI355X radiation library in-tree: mgcm_rt/libmgcm_rt.so (gfx950).

    python build.py            # compile every HIP source into the shared library, then the
                               # Fortran host module and driver (fortran/) with amdflang

Flags: -ffp-contract=off keeps the FP64 path's evaluation order (no fused
multiply-adds the reference does not make); packed FP32 ops are disabled.
"""
from __future__ import annotations

import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "mgcm_rt", "libmgcm_rt.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")
ARCH = os.environ.get("MGCM_RT_ARCH", "gfx950")

COMMON = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics", "-pthread"]


def build(verbose: bool = False) -> str:
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for s in srcs:
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        flags = list(COMMON)
        # the FP64 plain path and the table builders keep every rounding step
        flags += ["-ffp-contract=off"]
        cmds.append([HIPCC] + flags + ["-c", s, "-o", o])
        objs.append(o)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
        for r in ex.map(lambda cmd: subprocess.run(cmd, capture_output=True, text=True), cmds):
            if r.returncode != 0:
                sys.stderr.write(r.stderr)
                raise subprocess.CalledProcessError(r.returncode, r.args)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


def build_fortran(verbose: bool = False) -> str | None:
    """The ISO_C_BINDING module and the COMP3-style driver program (amdflang)."""
    if not os.path.exists(FC):
        print(f"build.py: {FC} not found; Fortran host not built", file=sys.stderr)
        return None
    objdir = os.path.join(HERE, "_build", "fortran")
    os.makedirs(objdir, exist_ok=True)
    src = os.path.join(HERE, "fortran")
    exe = os.path.join(src, "mgcm_rt_driver")
    steps = [
        [FC, "-O2", "-c", os.path.join(src, "mgcm_rt_mod.f90"), "-J", objdir, "-o",
         os.path.join(objdir, "mgcm_rt_mod.o")],
        [FC, "-O2", "-c", os.path.join(src, "mgcm_rt_driver.f90"), "-J", objdir, "-o",
         os.path.join(objdir, "mgcm_rt_driver.o")],
        [FC, "-o", exe + ".tmp", os.path.join(objdir, "mgcm_rt_driver.o"), os.path.join(objdir, "mgcm_rt_mod.o"),
         "-L", os.path.dirname(LIB), "-lmgcm_rt", "-Wl,-rpath,$ORIGIN/../mgcm_rt"],
    ]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    os.replace(exe + ".tmp", exe)
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
