"""Build the gfx950 library (libmph_hip.so) and the C++ host driver in-tree with hipcc.

Run ``python -m mphexplicit_amd.build`` (or ``__graft_entry__.build()``).  No CPU fallback
exists: the Python front end refuses to run when the library is missing.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libmph_hip.so")
DRIVER = os.path.join(PKG, "mph_explicit")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPH_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the exact parts of the step (periodic minimum image, list cutoff, wrap)
# must round like the reference; fused multiply-adds are written explicitly where wanted.
# Packed FP32 instructions are disabled for all device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I" + INCLUDE, "-I" + CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    objdir = os.path.join(PKG, "_objs")
    os.makedirs(objdir, exist_ok=True)
    dev_obj = os.path.join(objdir, "mph_device.o")
    io_obj = os.path.join(objdir, "mph_io.o")
    _run([HIPCC, "-c", "-x", "hip"] + COMMON + DEVICE +
         [os.path.join(CSRC, "mph_device.hip"), "-o", dev_obj])
    _run(["g++", "-c"] + COMMON + [os.path.join(CSRC, "mph_io.cpp"), "-o", io_obj])
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=" + ARCH, dev_obj, io_obj, "-o", tmp])
    os.replace(tmp, LIB)
    drv = os.path.join(CSRC, "mph_main.cpp")
    if os.path.exists(drv):
        _run(["g++"] + COMMON + [drv, "-o", DRIVER, "-L" + PKG, "-lmph_hip",
                                 "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
