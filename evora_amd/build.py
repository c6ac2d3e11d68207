"""Build libevora_hip.so (gfx950) in-tree with hipcc.

    python -m evora_amd.build          # builds evora_amd/libevora_hip.so

Every translation unit is compiled for gfx950 only.  The staged kernels, the
exact builds of the fused and conduction kernels, and the gravity kernels are
compiled without floating-point contraction so that their FP64 results follow
the reference's operation order.  The FP32-mixed builds of the march are
compiled without it too: the compiler fuses a product into a neighbouring add
differently at different inlined call sites of the same FP32 code, and the
march forms one cell's values at several (a ring lane or an interior lane, a
z chunk's first planes or the plane ahead), so that a contracted FP32 result
would depend on the launch shape.  Packed FP32 instructions are disabled for
all device code.  The library links hipFFT (ROCm) for the Poisson solve on
grids whose sides are not all powers of two.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libevora_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]

# (object, source, extra flags)
UNITS = [
    ("api.o", "api.hip", []),
    ("staged.o", "staged.hip", ["-ffp-contract=off"]),
    ("fused64.o", "fused.hip", ["-DEVH_FUSED_PREC=64"]),
    ("fused32.o", "fused.hip", ["-DEVH_FUSED_PREC=32", "-ffp-contract=off"]),
    ("fused64x.o", "fused.hip", ["-DEVH_FUSED_PREC=64", "-DEVH_FUSED_EXACT=1", "-ffp-contract=off"]),
    ("conduction.o", "conduction.hip", []),
    ("conductionx.o", "conduction.hip", ["-DEVH_COND_EXACT=1", "-ffp-contract=off"]),
    ("gravity.o", "gravity.hip", ["-ffp-contract=off"]),
    ("mathcheck.o", "mathcheck.hip", []),
    ("batch64.o", "batch.hip", ["-DEVH_FUSED_PREC=64"]),
    ("batch32.o", "batch.hip", ["-DEVH_FUSED_PREC=32", "-ffp-contract=off"]),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def build_library(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for obj, src, extra in UNITS:
        obj = os.path.join(objdir, obj)
        cmds.append([HIPCC, *COMMON, *extra, "-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd), flush=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as ex:
        list(ex.map(_run, cmds))
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-L/opt/rocm/lib", "-lhipfft"])
    os.replace(tmp, OUT)
    shutil.rmtree(objdir, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
