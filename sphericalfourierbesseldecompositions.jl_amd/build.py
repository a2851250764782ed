"""Builds libsfb_mi355x.so for gfx950 in-tree (next to the Python loader).

    python sphericalfourierbesseldecompositions.jl_amd/build.py

Host setup code (setup.cpp) is compiled with g++, the HIP library with hipcc;
both with -ffp-contract=off so every a*b+c that is not an explicit fma() keeps
the two roundings the reference's Julia code performs.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "sfb_amd")
LIB_NAME = "libsfb_mi355x.so"
BUILD_DIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wextra", "-pthread"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    setup_o = os.path.join(BUILD_DIR, "setup.o")
    copy_o = os.path.join(BUILD_DIR, "hostcopy.o")
    sfb_o = os.path.join(BUILD_DIR, "sfb.o")
    _run([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, "setup.cpp"), "-o", setup_o])
    _run([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, "hostcopy.cpp"), "-o", copy_o])
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "sfb.hip"), "-o", sfb_o])
    tmp = os.path.join(OUT_DIR, LIB_NAME + ".tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, sfb_o, setup_o, copy_o, "-pthread"])
    final = os.path.join(OUT_DIR, LIB_NAME)
    os.replace(tmp, final)
    return final


if __name__ == "__main__":
    path = build()
    print("built", path)
    sys.exit(0)
