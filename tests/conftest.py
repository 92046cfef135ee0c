import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import closed_form as cf  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU and the built libtsort.so")


_KATS = None


def kats():
    global _KATS
    if _KATS is None:
        with open(os.path.join(ROOT, "tests", "golden", "kats.json")) as f:
            _KATS = json.load(f)
    return _KATS


def kat_array(values, dtype_name: str) -> np.ndarray:
    """Decode a fixture list: floats are hex bit patterns, integers are ints."""
    npdt = cf.NP_DTYPES[dtype_name]
    if dtype_name[0] == "f":
        w = np.dtype(npdt).itemsize
        bits = np.array([int(v, 16) for v in values], dtype=cf.UNSIGNED_OF[w])
        return bits.view(npdt)
    return np.array(values, dtype=npdt)


def bits_of(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view(cf.UNSIGNED_OF[a.dtype.itemsize])


@pytest.fixture(scope="session")
def gpu():
    """The GPU test environment: fails (never skips) if the HIP path is absent."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a GPU (run them on the MI355X)")
    import turbosort_amd as ts

    ts._native.load()
    return torch.device("cuda", 0)
