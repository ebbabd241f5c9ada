"""Builds libphysim_gpu.so in-tree with hipcc for gfx950.

    python -m physim_amd.build          # or physim_amd.build.build()

Every source is compiled on each call; the product is a single shared
library holding the HIP kernels, the engine C-ABI (include/physim_gpu.h) and
the physim plugin symbols (include/physim_abi.h).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libphysim_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

SOURCES = ["engine.hip", "plugin.cpp", "generators.cpp", "server.cpp", "integrator.cpp", "runner.cpp"]

# Host-only sources are compiled by g++ with the oracle's flags, so host FP64
# (initial conditions) rounds like the reference's baseline x86-64 build.
HOST = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-I" + INCLUDE, "-I" + CSRC]
COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Wall",
    "-I" + INCLUDE,
    "-I" + CSRC,
]
DEVICE = [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    objs = []
    os.makedirs(os.path.join(HERE, "_obj"), exist_ok=True)
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(HERE, "_obj", src + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC] + COMMON + DEVICE + ["-c", path, "-o", obj]
        else:
            cmd = [CXX] + HOST + ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + objs + ["-ldl", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    # the C harness of the pipeline loop (a test program over the C-ABI)
    harness = os.path.join(ROOT, "tests", "harness")
    cmd = [CXX.replace("g++", "gcc") if CXX.endswith("g++") else "gcc", "-O2", "-std=c11",
           "-Wall", "-I" + INCLUDE, os.path.join(harness, "pipeline_harness.c"),
           "-L" + HERE, "-lphysim_gpu", "-Wl,-rpath,$ORIGIN/../../physim_amd",
           "-o", os.path.join(harness, "pipeline_harness")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
