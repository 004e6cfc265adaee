"""Build the native library and the drop-in executable in-tree.

    python -m meteorscattering_amd.build

produces meteorscattering_amd/libmsc.so (HIP kernels for gfx950 + C ABI)
and meteorscattering_amd/meteor (the reference's executable contract).
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
             # packed FP32 instructions are kept out of all generated code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall"]

HIP_SOURCES = ["msc_kernels.hip", "msc_solver.hip", "msc_server.hip"]
CXX_SOURCES = ["msc_io.cpp", "msc_run.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(SRC, s), "-o", o], verbose)
        objs.append(o)
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s.replace(".cpp", ".o"))
        _run(["g++", *CXX_FLAGS, "-c", os.path.join(SRC, s), "-o", o], verbose)
        objs.append(o)
    lib = os.path.join(PKG, "libmsc.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, lib)
    exe = os.path.join(PKG, "meteor")
    # $ORIGIN first, then the build directory itself: setup/meteorrun.m:162
    # copies the executable alone into each run directory
    _run(["g++", *CXX_FLAGS, os.path.join(SRC, "meteor_main.cpp"), "-o", exe, "-L" + PKG, "-lmsc",
          "-Wl,-rpath,$ORIGIN:" + os.path.realpath(PKG)], verbose)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
