"""Build step for the gfx950 shared library ``watercube_amd/libwatercube.so``.

Compiles every HIP source under ``watercube_amd/csrc`` with hipcc for gfx950 into one in-tree
shared library.  Device code is compiled without FMA contraction and with correctly rounded
binary32 division and square root, which the step's bit-exact contract depends on.

    python -m watercube_amd.build          # build (always recompiles)
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libwatercube.so"
HARNESS = PKG / "wc_harness"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["wc_api.hip", "wc_kernels.hip", "wc_mlp.hip", "wc_scenes.cpp"]

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-fno-fast-math",
    # packed FP32 forms are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> Path:
    srcs = [str(CSRC / s) for s in SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *HIP_FLAGS, "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed ({proc.returncode}) building {LIB.name}")
    os.replace(tmp, LIB)
    # headless C++ driver over the C ABI (host code only)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-o", str(HARNESS), str(CSRC / "wc_harness.cpp"),
           "-L", str(PKG), "-lwatercube", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("g++ failed building wc_harness")
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
