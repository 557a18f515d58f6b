"""Build recipe of the engine library (librockable_amd.so) for gfx950.

Every source is compiled by hipcc with -ffp-contract=off (no implicit fused
multiply-add, so FP64 results match the reference's x86-64 build bit for
bit) and without packed FP32 operations.  Run as a module:
    python -m rockable_amd.build [--jobs N]
"""
import argparse
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librockable_amd.so")
EXAMPLE = os.path.join(HERE, "rk_run")
SOURCES = ["rk_io.cpp", "rk_batch.cpp", "rk_step.hip", "rk_nl.hip", "rk_mlp.hip", "rk_engine.cpp"]
ARCH = os.environ.get("RK_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "--offload-arch=" + ARCH, "-Wall", "-Wno-unused-function", "-Wno-unused-result"]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), "-x", "hip"] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(jobs=4, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_example(verbose)
    return LIB


def build_example(verbose=False):
    """rk_run (examples/rk_run.cpp): the C++ caller of the C ABI, compiled by
    the host compiler against include/rockable_amd.h alone and linked to the
    library next to it."""
    src = os.path.join(HERE, "examples", "rk_run.cpp")
    inc = os.path.join(os.path.dirname(HERE), "include")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", inc, src, "-o", EXAMPLE + ".tmp", "-L", HERE,
           "-lrockable_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("example build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(EXAMPLE + ".tmp", EXAMPLE)
    if verbose:
        print("built", EXAMPLE)
    return EXAMPLE


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    build(a.jobs, verbose=True)
    sys.exit(0)
