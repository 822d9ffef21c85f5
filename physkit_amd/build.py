"""Builds physkit_amd/_lib/libphyskit_amd.so for gfx950 with hipcc.

All device code is compiled with FMA contraction off (-ffp-contract=off) so the
operation order written in the sources is the one executed, and with packed FP32
instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libphyskit_amd.so")
COPY_LIB = os.path.join(OUT_DIR, "libpk_copy_check.so")

SOURCES = ["api.hip", "k_body.hip", "k_sort.hip", "k_broad.hip", "k_narrow.hip", "k_gjk32.hip", "k_solver.hip", "k_joint.hip",
           "k_query.hip", "scenes.cpp"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    # the copy-calibration kernel for HBM byte counters (bench.py --copy-check)
    cmd = [hipcc(), *FLAGS, "-o", COPY_LIB + ".tmp", os.path.join(CSRC, "copy_check.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed (copy_check.hip)")
    os.replace(COPY_LIB + ".tmp", COPY_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
