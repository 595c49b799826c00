"""Build libprandtl_amd.so in-tree with hipcc for gfx950.

The library is compiled from several translation units in parallel: the C ABI
(``csrc/prandtl_capi.hip``) and the stage launcher ``csrc/stage_kernels.hip`` once per
(dim, Np) pair, each holding that pair's RHS / fused-stage kernel variants.  Objects go
to ``prandtl_amd/_build/``; the shared library is linked from them.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")
MAIN = os.path.join(CSRC, "prandtl_capi.hip")
STAGE = os.path.join(CSRC, "stage_kernels.hip")
STAGE_PAIRS = [(d, n) for d in (2, 3) for n in (2, 3, 4, 5)]
LIB = os.path.join(_HERE, "libprandtl_amd.so")
OBJDIR = os.path.join(_HERE, "_build")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _jobs() -> int:
    try:
        n = int(os.environ.get("MAX_JOBS", "0"))
    except ValueError:
        n = 0
    return max(1, min(n or (os.cpu_count() or 4), 16))


def build_library(extra_flags=None, verbose=False, lib_path=None, objdir=None) -> str:
    """Compile and link the library (to `lib_path`, default prandtl_amd/libprandtl_amd.so; objects
    in `objdir`, default prandtl_amd/_build/)."""
    objdir = objdir or OBJDIR
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    flags = HIPCC_FLAGS + list(extra_flags or [])
    units = [(MAIN, os.path.join(objdir, "prandtl_capi.o"), [])]
    for d, n in STAGE_PAIRS:
        units.append((STAGE, os.path.join(objdir, f"stage_d{d}_np{n}.o"),
                      [f"-DPRANDTL_STAGE_DIM={d}", f"-DPRANDTL_STAGE_NP={n}"]))
    # largest units first (3D, high order), the C ABI last
    pending = units[1:][::-1] + units[:1]
    running = []
    failed = []
    jobs = _jobs()
    while pending or running:
        while pending and len(running) < jobs:
            src, obj, defs = pending.pop(0)
            cmd = [cc] + flags + defs + ["-c", "-o", obj + ".tmp", src]
            if verbose:
                print(" ".join(cmd), file=sys.stderr)
            running.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True), obj))
        proc, obj = running.pop(0)
        out, err = proc.communicate()
        if proc.returncode != 0:
            failed.append((obj, proc.returncode, err))
        else:
            os.replace(obj + ".tmp", obj)
    if failed:
        obj, rc, err = failed[0]
        raise RuntimeError(f"hipcc failed on {os.path.basename(obj)} ({rc}):\n{err[-4000:]}")
    lib = lib_path or LIB
    tmp = lib + ".tmp"
    cmd = [cc] + flags + ["-shared", "-o", tmp] + [u[1] for u in units] + ["-lrocprofiler-sdk-roctx"]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc link failed ({r.returncode}):\n{r.stderr[-4000:]}")
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build_library(verbose=True))
