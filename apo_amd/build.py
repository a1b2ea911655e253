"""Build the gfx950 HIP library apo_amd/lib/libapo_hip.so with hipcc.

Every translation unit is compiled for gfx950 only, without FMA contraction
(the kernels reproduce IEEE host arithmetic bit for bit) and without packed
FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libapo_hip.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["lbvh.hip", "knn.hip", "cells.hip", "validation.hip", "complete.hip", "check.hip", "topology.hip",
           "learned.hip", "cavity.hip", "api.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    if err.strip():
        print(err, file=sys.stderr)
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + res.stderr)
    os.replace(tmp, LIB)
    build_samples()
    _build_example()
    if verbose:
        print(f"built {LIB}")
    return LIB


EXAMPLE = os.path.join(LIBDIR, "apo_diagram")
SHIM_CALLER = os.path.join(LIBDIR, "apo_reference_caller")
SAMPLES = os.path.join(LIBDIR, "libapo_samples.so")


def build_samples() -> str:
    """The host library of the seeded molecule generator (csrc/samples.cpp)."""
    os.makedirs(LIBDIR, exist_ok=True)
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared",
           os.path.join(CSRC, "samples.cpp"), "-o", SAMPLES + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("samples build failed:\n" + res.stderr)
    os.replace(SAMPLES + ".tmp", SAMPLES)
    return SAMPLES


def _build_example() -> None:
    """The C++ callers of include/apo/gpu/algorithm.hpp, linked to the library:
    examples/diagram.cpp (FP32 sites), and examples/reference_caller.cpp with
    the reference-side shim examples/apo_gpu_shim.cpp (Real sites, the
    header in its declarations-only form)."""
    include = os.path.join(os.path.dirname(HERE), "include")
    ex = os.path.join(HERE, "examples")
    cxx = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-Werror", "-I", include]
    link = ["-L", LIBDIR, "-lapo_hip", "-Wl,-rpath,$ORIGIN"]
    for out, srcs, extra in ((EXAMPLE, ["diagram.cpp"], []),
                             (SHIM_CALLER, ["reference_caller.cpp", "apo_gpu_shim.cpp"],
                              ["-DAPO_GPU_DECLARATIONS_ONLY"])):
        cmd = [*cxx, *extra, *[os.path.join(ex, s) for s in srcs], "-o", out + ".tmp", *link]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"example build failed ({os.path.basename(out)}):\n" + res.stderr)
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build()
