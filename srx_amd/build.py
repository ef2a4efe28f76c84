"""Build libsrx.so (HIP kernels + host driver) and the `srx` CLI for gfx950, in-tree."""
from __future__ import annotations

import concurrent.futures
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
DEV_SOURCES = ["newton.hip", "increment.hip", "fft.hip", "mlp.hip", "noalloc.hip"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose_resources: bool = False) -> None:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    hobj = os.path.join(objdir, "host.o")
    extra = ["-Rpass-analysis=kernel-resource-usage"] if verbose_resources else []
    # the device sources compile independently; the Newton file takes longest
    cmds, objs = [], []
    for src in DEV_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmds.append([HIPCC, *DEV_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    cmds.append([HIPCC, "-x", "c++", *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c",
                 os.path.join(CSRC, "host.cpp"), "-o", hobj])
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(run, cmds))
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fopenmp", "-o", LIB, *objs, hobj])
    run([HIPCC, "-x", "c++", *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
         os.path.join(CSRC, "cli.cpp"), "-o", CLI, "-L" + HERE, "-lsrx", "-Wl,-rpath,$ORIGIN",
         "-L/opt/rocm/lib", "-lamdhip64"])


if __name__ == "__main__":
    build(verbose_resources="-v" in sys.argv)
