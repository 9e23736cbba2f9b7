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

SOURCES = ["gbi_kernels.hip", "gbi_noalloc.hip", "gbi_context.cpp", "gbi_custom.cpp", "quadrature.cpp"]

# Packed-FP32 instructions stay disabled: the compiler must not form v_pk_*_f32.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


PLUGIN = os.path.join(PKG, "openmm_plugin")
PLUGIN_LIB = os.path.join(LIBDIR, "libgbi_openmm.so")
DRIVER = os.path.join(LIBDIR, "gbi_openmm_driver")


def build_openmm_plugin(verbose: bool = False) -> str:
    """The OpenMM kernel classes (HipCalcCharmmGB{SW,MV}ForceKernel, factory,
    registerKernelFactories) as lib/libgbi_openmm.so, plus the test driver.  Built against
    OpenMM when OPENMM_DIR and GBI_CHARMM_API_DIR (the plugin's openmmapi/include) are set,
    otherwise against the OpenMM-shaped stand-in headers."""
    cxx = shutil.which("g++") or "g++"
    omm, api = os.environ.get("OPENMM_DIR"), os.environ.get("GBI_CHARMM_API_DIR")
    if omm and api:
        incs = ["-I", os.path.join(omm, "include"), "-I", api]
        libs = ["-L", os.path.join(omm, "lib"), "-lOpenMM"]
        defs = []
    else:
        incs = ["-I", os.path.join(PLUGIN, "stand_in")]
        libs = []
        defs = ["-DGBI_OPENMM_STAND_IN"]
    common = [cxx, "-std=c++17", "-O2", "-fPIC", "-Wall", *defs, "-I", os.path.join(ROOT, "include"),
              "-I", os.path.join(PLUGIN, "src"), *incs]
    srcs = [os.path.join(PLUGIN, "src", f) for f in ("HipCharmmKernels.cpp", "HipCharmmKernelFactory.cpp")]
    tmp = PLUGIN_LIB + ".tmp"
    cmd = common + ["-shared", "-o", tmp, *srcs, "-L", LIBDIR, "-lgbi", *libs, "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, PLUGIN_LIB)
    if not (omm and api):
        # the driver also plays a GPU platform's device buffers, so it uses the HIP runtime
        cmd = common + ["-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-o", DRIVER + ".tmp",
                        os.path.join(PLUGIN, "driver", "gbi_openmm_driver.cpp"), "-L", LIBDIR, "-lgbi_openmm", "-lgbi",
                        "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(DRIVER + ".tmp", DRIVER)
    return PLUGIN_LIB


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
    # hiprtc compiles the kernels of general expressions when a context is created
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_openmm_plugin(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
