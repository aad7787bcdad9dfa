"""Build the gfx950 HIP library (libscihpc_hip.so) in-tree with hipcc.

Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
selection disabled, into one shared library exporting the C ABI of
include/scihpc_hip.h.  The C++ drivers under drivers/ (written against
include/scihpc_facade.hpp) are then compiled with g++ and linked to it.
``build()`` always recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "scihpc_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libscihpc_hip.so")

SOURCES = ["mg.hip", "advect.hip", "unet.hip", "snapshot.hip"]
DRIVERS = ["facade_demo", "dambreak_demo"]

HIPCC_FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-fPIC",
    "-shared",
    "-std=c++17",
    "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libscihpc_hip.so")
    os.replace(tmp, LIB)
    for d in DRIVERS:
        build_driver(d, verbose)
    return LIB


def build_driver(name: str, verbose: bool = False) -> str:
    src = os.path.join(PKG, "drivers", name + ".cpp")
    out = os.path.join(PKG, "drivers", name)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), src, "-L", PKG,
           "-lscihpc_hip", "-Wl,-rpath,$ORIGIN/..", "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"g++ failed building {name}")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
