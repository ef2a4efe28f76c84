"""Build libsrx.so (HIP kernels + host driver) and the `srx` CLI for gfx950, in-tree."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsrx.so")
CLI = os.path.join(HERE, "srx")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed-FP32 instructions are disabled for all device code.
DEV_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose_resources: bool = False) -> None:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    kobj = os.path.join(objdir, "kernels.o")
    hobj = os.path.join(objdir, "host.o")
    extra = ["-Rpass-analysis=kernel-resource-usage"] if verbose_resources else []
    mobj = os.path.join(objdir, "mlp.o")
    nobj = os.path.join(objdir, "noalloc.o")
    run([HIPCC, *DEV_FLAGS, *extra, "-c", os.path.join(CSRC, "kernels.hip"), "-o", kobj])
    run([HIPCC, *DEV_FLAGS, *extra, "-c", os.path.join(CSRC, "mlp.hip"), "-o", mobj])
    run([HIPCC, *DEV_FLAGS, *extra, "-c", os.path.join(CSRC, "noalloc.hip"), "-o", nobj])
    run([HIPCC, "-x", "c++", *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c",
         os.path.join(CSRC, "host.cpp"), "-o", hobj])
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fopenmp", "-o", LIB, kobj, mobj, nobj, hobj])
    run([HIPCC, "-x", "c++", *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
         os.path.join(CSRC, "cli.cpp"), "-o", CLI, "-L" + HERE, "-lsrx", "-Wl,-rpath,$ORIGIN",
         "-L/opt/rocm/lib", "-lamdhip64"])


if __name__ == "__main__":
    build(verbose_resources="-v" in sys.argv)
