"""Compile the gfx950 kernels and the C ABI into turbosort_amd/libtsort.so.

Usage: python -m turbosort_amd.build   (or turbosort_amd.build.build()).
Objects go to turbosort_amd/_build/, the library next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtsort.so")
ARCH = os.environ.get("TSORT_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # Keep the compiler from emitting packed-FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(os.path.dirname(HERE), "include"),
]

SOURCES = ["radix.hip", "segmented.hip", "learned.hip", "small.hip", "util.hip", "api.hip"]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


CPP_TEST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "hpp_test.cpp")
CPP_TEST_BIN = os.path.join(os.path.dirname(HERE), "tests", "cpp", "hpp_test.bin")


def build_cpp_test() -> str:
    """The C++ API test program (include/turbosort.hpp over libtsort.so)."""
    cmd = [HIPCC, *FLAGS, CPP_TEST_SRC, "-o", CPP_TEST_BIN, "-L", HERE, "-ltsort",
           f"-Wl,-rpath,{os.path.relpath(HERE, os.path.dirname(CPP_TEST_BIN))}".replace("-rpath,", "-rpath,$ORIGIN/")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on hpp_test.cpp:\n{r.stdout}\n{r.stderr}")
    return CPP_TEST_BIN


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_test()
    sys.exit(0)
