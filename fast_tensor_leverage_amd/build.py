"""Compiles libftl_hip.so (gfx950) in-tree with hipcc.

Usage: ``python -m fast_tensor_leverage_amd.build`` or ``build_library()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "libftl_hip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

SOURCES = ["sampler.hip", "als.hip"]
HEADERS = ["ftl_common.hpp", "sampler_kernels.hpp", "als_kernels.hpp"]

ARCH = os.environ.get("FTL_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
    # packed FP32 instructions are disabled for every kernel in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build_library(verbose: bool = False) -> str:
    """Compile every HIP source of the package into one shared library."""
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc()] + FLAGS + ["-I", os.path.join(os.path.dirname(HERE), "include")] + sources() + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with exit code {proc.returncode}")
    errs = [ln for ln in proc.stderr.splitlines()
            if "warning" in ln and "packed-fp32-ops" not in ln]
    if verbose and errs:
        sys.stderr.write("\n".join(errs) + "\n")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
