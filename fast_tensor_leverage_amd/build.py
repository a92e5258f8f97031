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

SOURCES = ["sampler.hip", "als.hip", "serve.hip"]
HEADERS = ["ftl_common.hpp", "sampler_kernels.hpp", "als_kernels.hpp", "jacobi.hpp", "wave_ops.hpp",
           "sampler_internal.hpp"]

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


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with exit code {proc.returncode}")
    errs = [ln for ln in proc.stderr.splitlines() if "warning" in ln and "packed-fp32-ops" not in ln]
    if verbose and errs:
        sys.stderr.write("\n".join(errs) + "\n")


def build_library(verbose: bool = False) -> str:
    """Compile every HIP source of the package (one hipcc per source, in
    parallel) and link them into one shared library."""
    from concurrent.futures import ThreadPoolExecutor
    inc = ["-I", os.path.join(os.path.dirname(HERE), "include")]
    compile_flags = [f for f in FLAGS if f != "-shared"]
    objs = []
    jobs = []
    for src in sources():
        obj = os.path.join(CSRC, os.path.basename(src) + ".o")
        objs.append(obj)
        jobs.append([hipcc()] + compile_flags + inc + ["-c", src, "-o", obj])
    with ThreadPoolExecutor(max_workers=min(len(jobs), 4)) as ex:
        for f in [ex.submit(_run, j, verbose) for j in jobs]:
            f.result()
    tmp = LIB_PATH + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC"] + objs + ["-o", tmp], verbose)
    for o in objs:
        os.remove(o)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
