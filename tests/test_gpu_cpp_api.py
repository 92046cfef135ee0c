"""The C++ mirror of the reference API (include/turbosort.hpp) on the GPU.

Runs tests/cpp/hpp_test.bin, built by __graft_entry__.build() against
libtsort.so; it checks the reference's doctest answers and panic messages.
"""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "hpp_test.bin")


def test_cpp_api_program(gpu):
    assert os.path.exists(BIN), "tests/cpp/hpp_test.bin is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
