"""Exhaustive checks of the LDS sorting networks on the GPU.

The reference pins its small-slice networks by the 0-1 principle (every
0/1 vector of the network's length, arch/aarch64/neon.rs:227-253) and by
every permutation of small lengths (tests/correctness.rs:50-72).  Here the
same inputs go through the device's networks:

* segmented_sort: all 2^16 0-1 vectors of length 16 as 65536 segments, and
  all 8! permutations of 0..7 as 40320 segments, with values (the global
  index), so the wave bitonic network (k_seg_bitonic) must also keep equal
  keys in input order;
* ts_sort: every permutation of lengths 2..8 and every 0-1 vector of lengths
  2..12, each its own sort (the one-workgroup network, k_small);
* ts_sort_by_key: every 0-1 vector of length 10, whose stable order pins the
  (key, position) network.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest
import torch

import turbosort_amd as ts
from gpu_util import to_dev, to_host
from oracle import closed_form as cf

pytestmark = pytest.mark.gpu


def zero_one_vectors(length: int) -> np.ndarray:
    v = np.arange(1 << length, dtype=np.uint32)
    return ((v[:, None] >> np.arange(length, dtype=np.uint32)) & 1).astype(np.uint32)


def permutations(length: int) -> np.ndarray:
    return np.array(list(itertools.permutations(range(length))), dtype=np.uint32)


@pytest.mark.parametrize("what", ["zero-one 16", "permutations 8"])
def test_segment_networks_exhaustive(gpu, what):
    rows = zero_one_vectors(16) if what == "zero-one 16" else permutations(8)
    m, length = rows.shape
    keys = rows.reshape(-1)
    vals = np.arange(keys.size, dtype=np.uint32)
    offsets = np.arange(m + 1, dtype=np.uint64) * np.uint64(length)
    kt, vt = to_dev(keys, gpu), to_dev(vals, gpu)
    ts.segmented_sort(kt, to_dev(offsets, gpu), vt, max_segment_len=length)
    want_k, want_v = cf.segmented_sort(keys, offsets, vals)
    assert np.array_equal(to_host(kt, np.uint32), want_k)
    assert np.array_equal(to_host(vt, np.uint32), want_v)


def sort_each_row(rows: np.ndarray, dev) -> np.ndarray:
    """Sort every row as its own slice (ts_sort on views of one tensor)."""
    t = to_dev(rows.reshape(-1), dev)
    length = rows.shape[1]
    for i in range(rows.shape[0]):
        ts.sort(t[i * length:(i + 1) * length])
    return to_host(t, rows.dtype).reshape(rows.shape)


@pytest.mark.parametrize("length", range(2, 9))
def test_small_sort_every_permutation(gpu, length):
    rows = permutations(length)
    got = sort_each_row(rows, gpu)
    assert np.array_equal(got, np.broadcast_to(np.arange(length, dtype=np.uint32), rows.shape))


@pytest.mark.parametrize("length", range(2, 13))
def test_small_sort_every_zero_one_vector(gpu, length):
    rows = zero_one_vectors(length)
    got = sort_each_row(rows, gpu)
    assert np.array_equal(got, np.sort(rows, axis=1))


@pytest.mark.parametrize("dtype", ["u64", "f32"])
def test_small_sort_zero_one_other_widths(gpu, dtype):
    """The same network on 8-byte and float keys (two values: 0 and 1 as the type)."""
    rows = zero_one_vectors(12)
    x = rows.astype(cf.NP_DTYPES[dtype])
    t = to_dev(x.reshape(-1), gpu)
    for i in range(rows.shape[0]):
        ts.sort(t[i * 12:(i + 1) * 12])
    got = to_host(t, x.dtype).reshape(rows.shape)
    assert np.array_equal(got, np.sort(x, axis=1))


def test_by_key_network_stable_order_on_zero_one_vectors(gpu):
    rows = zero_one_vectors(10)
    buf = ts.SortByKeyBuffer(gpu)
    orders = []
    for r in rows:
        t = to_dev(r, gpu)
        ts.sort_by_key_with_buffer(t, buf)
        orders.append(to_host(buf.order(), np.uint32))
    got = np.stack(orders)
    want = np.argsort(rows, axis=1, kind="stable").astype(np.uint32)
    assert np.array_equal(got, want)
