"""Builds the solver library and its host layer in-tree.

    python -m vlp2d_amd.build

lib/libvlp2d_mg.so: the gfx950 solver.  csrc/mg_lib.hip holds the C ABI and
the small kernels; the streaming kernels are instantiated in one translation
unit per precision and sweep count (csrc/mg_k_down_*.hip, csrc/mg_k_up_*.hip),
all compiled in parallel with hipcc -c and linked into one shared library.
-ffp-contract=off keeps every product and sum separately rounded, as written,
so device results are reproducible operation by operation; packed-FP32
instructions are disabled.

lib/libvlp2d_host.so and lib/vlp2d_solve: the C++ host layer (host/: input
files, scheme object, model loop) over the solver's C ABI, built with g++.
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(OUT_DIR, "obj")
OUT = os.path.join(OUT_DIR, "libvlp2d_mg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
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


def sources():
    """Translation units of the solver library (the ABI unit first)."""
    tus = sorted(glob.glob(os.path.join(CSRC, "mg_k_*.hip")))
    return [os.path.join(CSRC, "mg_lib.hip")] + tus


def _jobs():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 4
    return max(1, min(n, 16))


def build(verbose=False):
    """Compiles every translation unit (in parallel) and links the library."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(OBJ_DIR, os.path.basename(p)[:-4] + ".o") for p in srcs]

    def one(pair):
        src, obj = pair
        _run([HIPCC] + FLAGS + ["-c", "-o", obj + ".tmp", src], verbose)
        os.replace(obj + ".tmp", obj)

    with ThreadPoolExecutor(max_workers=_jobs()) as ex:
        list(ex.map(one, zip(srcs, objs)))
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, OUT)
    build_host(verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-q" not in sys.argv))
