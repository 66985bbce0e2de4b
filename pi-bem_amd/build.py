"""Build libpbfmm.so (HIP kernels for gfx950 + C++ host setup + C ABI).

    python pi-bem_amd/build.py

Host-only C++ is compiled by g++ with -ffp-contract=off so that the octree
construction (box membership tests on coordinates) is reproducible; device code
is compiled by hipcc for gfx950 with packed-FP32 instructions disabled and
hardware FP64 atomics enabled.  Objects go to pi-bem_amd/_build, the library to
pi-bem_amd/lib.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libpbfmm.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HOST_SRC = ["setup.cpp", "capi.cpp", "forest.cpp"]
DEV_SRC = ["engine.hip", "postsolve.hip"]
HDRS = ["pbfmm_impl.h", "engine.h"]

HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-pthread", "-Wall", "-Wno-unused-function",
              f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]
DEV_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build(verbose: bool = False) -> str:
    """Compile every source for gfx950 and link lib/libpbfmm.so.  Always
    recompiles (the build check must exercise the compiler)."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for s in HOST_SRC:
        o = os.path.join(BUILD, s + ".o")
        _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in DEV_SRC:
        o = os.path.join(BUILD, s + ".o")
        out = _run([HIPCC, *DEV_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, s), "-o", o])
        if verbose and out:
            print(out)
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
