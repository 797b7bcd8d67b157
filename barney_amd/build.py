"""Compile the gfx950 HIP library (barney_amd/libbnx.so) with hipcc.

Every call compiles all sources (the objects go to barney_amd/_build/);
``python -m barney_amd.build`` does the same from the command line.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbnx.so")
SOURCES = ["bnx_core.hip", "bnx_bvh.hip", "bnx_trace.hip", "bnx_render.hip", "bnx_volume.hip", "bnx_engine.hip",
           "bnx_env.hip", "bnx_tune.hip", "bnx_sort.hip"]
ARCH = os.environ.get("BNX_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off and correctly rounded division / square root keep every
# float operation identical to its scalar CPU evaluation; packed fp32 ops are
# disabled (the compiler otherwise forms v_pk_* instructions on its own).
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wno-unused-result",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-8000:]}")
    return obj


# host-only C++ over the C-ABI, linked into the same library: barney's bn*
# API (include/bnx_barney.h)
HOST_SOURCES = ["bnx_barney.cpp"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include")]


def _compile_host(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".host.o")
    cmd = ["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{r.stderr[-8000:]}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES) + len(HOST_SOURCES))) as ex:
        hip = ex.map(_compile, SOURCES)
        host = ex.map(_compile_host, HOST_SOURCES)
        objs = list(hip) + list(host)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-8000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_host_programs(verbose)
    return LIB


# host programs over the C-ABI (g++, no HIP headers): the C++ mirror of
# barney's rtc interface driven as barney's native layer would
HOST_PROGRAMS = [("tests/cpp/rtc_trace.cpp", "tests/cpp/rtc_trace"), ("tests/cpp/bn_render.cpp", "tests/cpp/bn_render")]


def build_host_programs(verbose: bool = False):
    for src, out in HOST_PROGRAMS:
        dst = os.path.join(ROOT, out)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
               os.path.join(ROOT, src), "-L", HERE, "-lbnx", "-Wl,-rpath,$ORIGIN/../../barney_amd",
               "-Wl,-rpath-link,/opt/rocm/lib", "-o", dst]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {src}:\n{r.stderr[-8000:]}")
        if verbose:
            print(f"built {dst}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
