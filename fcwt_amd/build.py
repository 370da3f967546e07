"""Builds libfcwt_hip.so (gfx950) in-tree: fcwt_amd/_lib/libfcwt_hip.so.

Host-only sources (the reference's float32 host arithmetic and the C++ API)
go through the host compiler with -ffp-contract=off; the HIP sources through
hipcc for gfx950.  Run as ``python -m fcwt_amd.build``.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libfcwt_hip.so"
INCLUDE = ROOT / "include"

HOST_SOURCES = ["host_math.cpp", "fcwt_api.cpp"]
HIP_SOURCES = ["fcwt_hip.hip"]

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Packed FP32 instructions stay off: the compiler emits them by itself at -O3.
HIP_FLAGS = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-ffp-contract=off", "-fno-fast-math"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(map(str, cmd)) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ({r.returncode})")
    # hipcc's host pass does not know the packed-fp32 feature; that warning is noise
    noise = "'-packed-fp32-ops' is not a recognized feature"
    msg = "\n".join(l for l in (r.stderr or "").splitlines() if noise not in l and l.strip())
    if msg:
        sys.stderr.write(msg + "\n")


def build(verbose: bool = False) -> pathlib.Path:
    """Compile every source and link the shared library; returns its path."""
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    objs = []
    for s in HOST_SOURCES:
        o = objdir / (s + ".o")
        _run([CXX, *HOST_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(CSRC / s), "-o", str(o)])
        objs.append(o)
    for s in HIP_SOURCES:
        o = objdir / (s + ".o")
        _run([HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(CSRC / s), "-o", str(o)])
        objs.append(o)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
