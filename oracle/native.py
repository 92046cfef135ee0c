"""ctypes access to the C++ restatement (libtsoracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from . import closed_form as cf

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtsoracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.tso_sort.argtypes = [ctypes.c_int, vp, sz]
        L.tso_sort_with_buffer.argtypes = [ctypes.c_int, vp, sz, vp, sz]
        L.tso_sort_parallel.argtypes = [ctypes.c_int, vp, sz, ctypes.c_int]
        L.tso_sort_by_key_order.argtypes = [ctypes.c_int, vp, sz, vp]
        L.tso_apply_order.argtypes = [vp, sz, vp, sz, sz]
        L.tso_log2_eighths.argtypes = [ctypes.c_uint64]
        L.tso_log2_eighths.restype = ctypes.c_uint32
        L.tso_skippable_digits.argtypes = [sz, sz, vp]
        L.tso_skippable_digits.restype = sz
        L.tso_varying_bits.argtypes = [vp, sz, ctypes.c_int, vp]
        L.tso_finish_runs_u16.argtypes = [vp, sz, ctypes.c_uint32]
        L.tso_splitmix64_fill.argtypes = [ctypes.c_uint64, sz, vp]
        L.tso_segmented_sort_pairs.argtypes = [ctypes.c_int, vp, vp, vp, sz, ctypes.c_int]
        _lib = L
    return _lib


def _code(arr: np.ndarray) -> int:
    return cf.DTYPES.index(cf.dtype_name(arr))


def sort(values: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(values).copy()
    rc = lib().tso_sort(_code(out), out.ctypes.data, out.size)
    assert rc == 0, rc
    return out


def sort_with_buffer(values: np.ndarray, buffer_len: int | None = None) -> np.ndarray:
    out = np.ascontiguousarray(values).copy()
    buf = np.empty(out.size if buffer_len is None else buffer_len, dtype=out.dtype)
    rc = lib().tso_sort_with_buffer(_code(out), out.ctypes.data, out.size, buf.ctypes.data, buf.size)
    if rc == 1:
        raise ValueError(f"buffer too small: need {out.size}, got {buf.size}")
    assert rc == 0, rc
    return out


def sort_parallel_inplace(values: np.ndarray, threads: int) -> None:
    assert values.flags.c_contiguous
    rc = lib().tso_sort_parallel(_code(values), values.ctypes.data, values.size, threads)
    assert rc == 0, rc


def sort_parallel(values: np.ndarray, threads: int) -> np.ndarray:
    out = np.ascontiguousarray(values).copy()
    sort_parallel_inplace(out, threads)
    return out


def order(keys: np.ndarray) -> np.ndarray:
    keys = np.ascontiguousarray(keys)
    out = np.empty(keys.size, dtype=np.uint32)
    rc = lib().tso_sort_by_key_order(_code(keys), keys.ctypes.data, keys.size, out.ctypes.data)
    assert rc == 0, rc
    return out


def apply_order(order_arr: np.ndarray, values: np.ndarray) -> np.ndarray:
    order_arr = np.ascontiguousarray(order_arr, dtype=np.uint32)
    out = np.ascontiguousarray(values).copy()
    elem = out.dtype.itemsize * (out[0].size if out.ndim > 1 else 1)
    rc = lib().tso_apply_order(order_arr.ctypes.data, order_arr.size, out.ctypes.data, out.shape[0], elem)
    if rc == 2:
        raise ValueError("apply_order: the slice's length differs from the last sorted slice's")
    assert rc == 0, rc
    return out


def log2_eighths(x: int) -> int:
    return int(lib().tso_log2_eighths(x))


def skippable_digits(n_live: int, length: int, max_bucket) -> int:
    arr = np.array([max_bucket(i) for i in range(n_live)], dtype=np.uint64) if n_live else np.zeros(1, np.uint64)
    return int(lib().tso_skippable_digits(n_live, length, arr.ctypes.data))


def varying_bits(keys: np.ndarray):
    w = keys.dtype.itemsize
    k64 = np.ascontiguousarray(cf.to_radix(keys).astype(np.uint64))
    out = np.zeros(1, np.uint64)
    some = lib().tso_varying_bits(k64.ctypes.data, k64.size, w, out.ctypes.data)
    return int(out[0]) if some else None


def finish_runs_u16(keys: np.ndarray, shift: int) -> np.ndarray:
    out = np.ascontiguousarray(keys, dtype=np.uint16).copy()
    lib().tso_finish_runs_u16(out.ctypes.data, out.size, shift)
    return out


def splitmix64(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    lib().tso_splitmix64_fill(seed, n, out.ctypes.data)
    return out


def segmented_sort_pairs_inplace(keys: np.ndarray, offsets: np.ndarray, values=None, threads: int = 1) -> None:
    """Every segment sorted stably by key with its u32 values, in place (the
    restated per-segment sort_by_key_with_buffer + move)."""
    assert keys.flags.c_contiguous
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    vp = None if values is None else values.ctypes.data
    rc = lib().tso_segmented_sort_pairs(_code(keys), keys.ctypes.data, vp, off.ctypes.data, off.size - 1, threads)
    assert rc == 0, rc
