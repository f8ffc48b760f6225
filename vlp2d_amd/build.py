"""Builds the solver library and its host layer in-tree.

    python -m vlp2d_amd.build

lib/libvlp2d_mg.so: the gfx950 solver, one hipcc translation unit
(csrc/mg_lib.hip and the headers it includes).  -ffp-contract=off keeps every
product and sum separately rounded, as written, so device results are
reproducible operation by operation; packed-FP32 instructions are disabled.

lib/libvlp2d_host.so and lib/vlp2d_solve: the C++ host layer (host/: input
files, scheme object, model loop) over the solver's C ABI, built with g++.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "mg_lib.hip")
OUT_DIR = os.path.join(HERE, "lib")
OUT = os.path.join(OUT_DIR, "libvlp2d_mg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


HOST_DIR = os.path.join(HERE, "host")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HOST_OUT = os.path.join(OUT_DIR, "libvlp2d_host.so")
SOLVE_OUT = os.path.join(OUT_DIR, "vlp2d_solve")
CXX = os.environ.get("CXX", "g++")
ROCM_LIB = "/opt/rocm/lib"


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_host(verbose=False):
    """Host layer (needs lib/libvlp2d_mg.so)."""
    common = ["-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-I", HOST_DIR, "-L", OUT_DIR,
              "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link," + ROCM_LIB]
    tmp = HOST_OUT + ".tmp"
    _run([CXX, "-shared", "-fPIC"] + common + ["-o", tmp, os.path.join(HOST_DIR, "vlp2d_host.cpp"), "-lvlp2d_mg"],
         verbose)
    os.replace(tmp, HOST_OUT)
    tmp = SOLVE_OUT + ".tmp"
    _run([CXX] + common + ["-o", tmp, os.path.join(HOST_DIR, "vlp2d_solve.cpp"), "-lvlp2d_host", "-lvlp2d_mg"],
         verbose)
    os.replace(tmp, SOLVE_OUT)
    return HOST_OUT


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = OUT + ".tmp"
    _run([HIPCC] + FLAGS + ["-o", tmp, SRC], verbose)
    os.replace(tmp, OUT)
    build_host(verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-q" not in sys.argv))
