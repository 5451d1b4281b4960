"""Builds the in-tree native library dnds_amd/libdndsmi.so for gfx950.

Host C++ (mesh topology, VR setup) is compiled with g++; the HIP sources
(device context and kernels) with hipcc for gfx950; hipcc links the shared
library.  All translation units use -ffp-contract=off so every multiply and
add rounds separately, as the reference's FP64 expressions do.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdndsmi.so")

HOST_SOURCES = ["mesh.cpp", "vr_setup.cpp", "io.cpp"]
HIP_SOURCES = ["ctx.hip", "lusgs.hip", "warmstart.hip"]

HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-I" + os.path.join(HERE, "..", "include")]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(HERE, "..", "include")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose: bool = False) -> str:
    """Compile every source and link libdndsmi.so (always rebuilds)."""
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for s in HOST_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, s), "-o", o], verbose)
        objs.append(o)
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        _run(["hipcc", *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o], verbose)
        objs.append(o)
    tmp = LIB + ".tmp"
    _run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp, "-lpthread"], verbose)
    os.replace(tmp, LIB)
    build_facade_run(verbose)
    return LIB


FACADE_SRC = os.path.join(HERE, "..", "tests", "cpp", "facade_run.cpp")
FACADE_BIN = os.path.join(HERE, "..", "tests", "cpp", "facade_run")


def build_facade_run(verbose: bool = False) -> str:
    """The C++ host-class example (tests/cpp/facade_run.cpp), linked to the
    in-tree library with an rpath to it."""
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(HERE, "..", "include"), FACADE_SRC,
          "-o", FACADE_BIN, "-L" + HERE, "-ldndsmi", "-Wl,-rpath,$ORIGIN/../../dnds_amd"], verbose)
    return FACADE_BIN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
