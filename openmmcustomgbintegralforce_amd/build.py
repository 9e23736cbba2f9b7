"""Ahead-of-time build of the gfx950 library (libgbi.so) with hipcc.

The library is built in-tree (``openmmcustomgbintegralforce_amd/lib/libgbi.so``) so it
travels with the repository snapshot.  Usage: ``python -m openmmcustomgbintegralforce_amd.build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libgbi.so")
ARCH = os.environ.get("GBI_OFFLOAD_ARCH", "gfx950")

SOURCES = ["gbi_kernels.hip", "gbi_context.cpp", "quadrature.cpp"]

# Packed-FP32 instructions stay disabled: the compiler must not form v_pk_*_f32.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
               "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c", path, "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
            cmd += DEVICE_FLAGS
        else:
            cmd[1:1] = ["-x", "c++"]
            cmd = [c for c in cmd if not c.startswith("--offload-arch")]
            cmd += ["-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
