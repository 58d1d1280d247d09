"""Build libckmpm.so (HIP, gfx950) in-tree with hipcc, and the ckmpm_scene driver (host
C++ over include/ckmpm.hpp, linked to the library) with the host compiler."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libckmpm.so")
DRIVER = os.path.join(HERE, "ckmpm_scene")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
CXX = os.environ.get("CXX", "g++")
SOURCES = ["kernels.hip", "engine.hip", "hash.hip", "adjoint.hip", "conservation.hip", "setup_wide.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
    # no packed FP32 instructions (v_pk_*_f32) in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose: bool = True) -> str:
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *[f for f in FLAGS if f != "-shared"], "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    tmp = DRIVER + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", INCLUDE, os.path.join(CSRC, "ckmpm_scene.cpp"),
           "-o", tmp, "-L", HERE, "-lckmpm", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, DRIVER)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
