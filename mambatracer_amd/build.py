"""Builds libmambart.so (HIP, gfx950) in-tree with hipcc, and beside it
libmt_calib.so, the vector-ALU issue-rate calibration kernels that the
benchmark's roofline uses (not linked into libmambart).

Every HIP translation unit is compiled with:
  -ffp-contract=off        no FMA contraction: kernels round like the CPU
                           restatement used by the tests;
  -fhip-fp32-correctly-rounded-divide-sqrt
                           IEEE division and square root;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                           no packed-FP32 instructions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmambart.so")
CALIB_LIB = os.path.join(HERE, "libmt_calib.so")
CALIB_SOURCES = ["mt_calib.hip"]
ARCH = os.environ.get("MAMBART_ARCH", "gfx950")

SOURCES = [
    "mt_context.hip",
    "mt_host.cpp",
    "mt_trace.hip",
    "mt_scan.hip",
    "mt_sat.hip",
    "mt_bsdf.hip",
    "mt_render.hip",
    "mt_batch.hip",
    "mt_bvh.hip",
    "mt_image.hip",
    "mt_io.cpp",
    "mt_error.cpp",
]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; libmambart needs ROCm's hipcc to build")
    return exe


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [_hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [_hipcc(), *HIPCC_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def _link(objs, lib: str, verbose: bool) -> None:
    tmp = lib + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link of {os.path.basename(lib)} failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, lib)


def build(verbose: bool = False) -> str:
    """Compiles every source, links libmambart.so and libmt_calib.so; returns
    libmambart.so's path."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = SOURCES + CALIB_SOURCES
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    _link(objs[:len(SOURCES)], LIB, verbose)
    _link(objs[len(SOURCES):], CALIB_LIB, verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
