"""CPU model of the device sort's index arithmetic (tests/cpp/sweep_model.cpp).

Checks, without a GPU, that the segmented one-sweep passes enumerate every
tile once in an order where each tile's predecessor comes first, write every
position exactly once and in bounds, and produce the stable order.
"""
import os
import subprocess

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "sweep_model.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "sweep_model.bin")


def test_sweep_index_model():
    if not os.path.exists(BIN) or os.path.getmtime(BIN) < max(
            os.path.getmtime(SRC), os.path.getmtime(os.path.join(ROOT, "turbosort_amd", "csrc", "keys.h"))):
        subprocess.run(["g++", "-O2", "-std=c++17", "-o", BIN, SRC], check=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
