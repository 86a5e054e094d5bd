"""Compile the HIP sources into the in-tree C-ABI library ``liblithobat_amd.so``.

    python -m lithobat_amd.build            # gfx950, all sources, in parallel

hipcc cross-compiles for gfx950 without a GPU.  Packed FP32 instructions are
disabled for every device compile (some of their forms have produced wrong
results on the target boxes).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblithobat_amd.so")
SOURCES = ["lb_api.hip", "lb_band.hip", "lb_band_any.hip", "lb_blur.hip", "lb_field.hip",
           "lb_objective.hip", "lb_pupil.hip", "lb_socs.hip", "lb_util.hip"]
ARCH = os.environ.get("LITHOBAT_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the library)")


def device_flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-I", os.path.join(HERE, "..", "include")]


# Per-source extra flags.  The coherent-system kernels feed compile-time zeros
# (box rows and columns past the pupil support) into their first radix stage;
# without signed zeros x + 0 folds to x, so those additions drop out (the
# sign of a zero is never observed there: only |E|^2 and products leave).
EXTRA_FLAGS = {"lb_field.hip": ["-fno-signed-zeros"]}


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [hipcc(), *device_flags(), *EXTRA_FLAGS.get(src, []), "-c", os.path.join(CSRC, src),
           "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
