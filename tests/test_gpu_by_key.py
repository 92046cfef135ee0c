"""Stable sort_by_key, order() and apply_order on the GPU vs the oracle.

order() must equal the stable argsort by radix key exactly (by_key.rs:92-110),
the records must end up in that order, and apply_order must move a parallel
slice the same way (by_key.rs:136-144).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import turbosort_amd as ts
from conftest import kat_array, kats
from gpu_util import bits, to_dev, to_host
from oracle import closed_form as cf
from oracle import native as nv
from shapes import float_specials_f32, random_of

pytestmark = pytest.mark.gpu

REC = np.dtype([("id", "<u4"), ("score", "<f4"), ("payload", "<u8")])


def records(n: int, seed: int) -> np.ndarray:
    r = np.zeros(n, dtype=REC)
    r["id"] = np.arange(n, dtype=np.uint32)
    r["score"] = float_specials_f32(n, seed)
    r["payload"] = cf.splitmix64(seed, n)
    return r


def rec_to_dev(r: np.ndarray, dev) -> torch.Tensor:
    return to_dev(r.view(np.uint8).reshape(len(r), REC.itemsize), dev)


def rec_to_host(t: torch.Tensor) -> np.ndarray:
    return t.cpu().numpy().reshape(-1).view(REC)


@pytest.mark.parametrize("i", range(len(kats()["order"])))
def test_reference_order_kats(gpu, i):
    k = kats()["order"][i]
    keys = kat_array(k["keys"], k["dtype"])
    t = to_dev(keys, gpu)
    buf = ts.SortByKeyBuffer()
    ts.sort_by_key_with_buffer(t, buf)
    assert to_host(buf.order(), np.uint32).tolist() == k["order"], k["src"]
    assert to_host(t, keys.dtype).tolist() == sorted(k["keys"])
    if "apply_to" in k:
        other = to_dev(np.array(k["apply_to"], dtype=np.int64), gpu)
        buf.apply_order(other, [])
        assert other.cpu().tolist() == k["applied"]


@pytest.mark.parametrize("n", [0, 1, 2, 3, 63, 64, 100, 4096, 4097, 100_000, 1_000_000])
def test_records_keyed_on_f32_score(gpu, n):
    r = records(n, seed=n + 5)
    t = rec_to_dev(r, gpu)
    buf = ts.SortByKeyBuffer()
    ts.sort_by_key_with_buffer(t, buf, ts.KeyField(4, torch.float32))
    want = cf.order(r["score"])
    got = to_host(buf.order(), np.uint32)
    assert got.shape == (n,)
    assert np.array_equal(got, want)
    if n <= 100_000:
        assert np.array_equal(got, nv.order(r["score"]))
    out = rec_to_host(t)
    assert np.array_equal(out.view(np.uint8), r[want.astype(np.int64)].view(np.uint8))
    # apply_order on a second u64 slice
    other = cf.splitmix64(n + 99, n)
    o = to_dev(other, gpu)
    buf.apply_order(o)
    assert np.array_equal(to_host(o, np.uint64), other[want.astype(np.int64)])


@pytest.mark.parametrize("name", ["u8", "u16", "u32", "u64", "i8", "i16", "i32", "i64", "f32", "f64"])
def test_key_arrays_every_dtype(gpu, name):
    g = np.random.default_rng(17)
    for n in (2, 50, 4096, 4097, 70_000, 300_000):
        for few in (False, True):
            x = random_of(name, n, seed=n + 3)
            if few:
                x = x[g.integers(0, min(7, n), size=n)]  # duplicates, so stability shows
            t = to_dev(x, gpu)
            buf = ts.SortByKeyBuffer()
            ts.sort_by_key_with_buffer(t, buf)
            assert np.array_equal(to_host(buf.order(), np.uint32), cf.order(x)), (name, n, few)
            assert np.array_equal(bits(to_host(t, x.dtype)), bits(cf.sort(x))), (name, n, few)


def test_callable_key_and_unbuffered_sort_by_key(gpu):
    # lib.rs:145-147: people sorted by age, stable
    ages = np.array([31, 25, 31, 25], dtype=np.uint32)
    names = to_dev(np.array([0, 1, 2, 3], dtype=np.int64), gpu)  # bob, alice, carol, dave
    people = torch.stack([names, to_dev(ages.astype(np.int64), gpu)], dim=1).contiguous()
    ts.sort_by_key(people, lambda p: p[:, 1].to(torch.uint32).contiguous())
    assert people.cpu().tolist() == [[1, 25], [3, 25], [0, 31], [2, 31]]


def test_buffer_reuse_and_capacity(gpu):
    buf = ts.SortByKeyBuffer.new()
    assert buf.order().numel() == 0 and buf.capacity() == 0
    for n in (1000, 200_000, 5000):
        x = random_of("u32", n, seed=n) % np.uint32(100)
        t = to_dev(x, gpu)
        ts.sort_by_key_with_buffer(t, buf)
        assert np.array_equal(to_host(buf.order(), np.uint32), cf.order(x))
    assert buf.capacity() >= 200_000
    assert "capacity" in repr(buf)


def test_apply_order_length_mismatch_panics(gpu):
    buf = ts.SortByKeyBuffer()
    ts.sort_by_key_with_buffer(to_dev(np.array([3, 1, 2], dtype=np.uint32), gpu), buf)
    with pytest.raises(ts.TurbosortPanic, match="apply_order: the slice's length differs from the last sorted"):
        buf.apply_order(torch.zeros(4, dtype=torch.int64, device=gpu))


def test_generic_element_sizes_for_apply_order(gpu):
    n = 10_000
    x = random_of("u32", n, seed=1)
    buf = ts.SortByKeyBuffer()
    ts.sort_by_key_with_buffer(to_dev(x, gpu), buf)
    order = cf.order(x).astype(np.int64)
    for width in (1, 2, 3, 4, 6, 8, 12, 16, 20, 24, 40):
        other = np.random.default_rng(width).integers(0, 255, size=(n, width), dtype=np.uint8)
        t = torch.from_numpy(other.copy()).to(gpu)
        buf.apply_order(t)
        assert np.array_equal(t.cpu().numpy(), other[order]), width


def test_presorted_and_descending_keys_keep_the_stable_order(gpu):
    """Ascending keys give the identity order; non-increasing keys with ties
    are not reversed by key-value sorts (ties keep their input order)."""
    n = 300_000
    asc = np.sort(random_of("u32", n, seed=5) % 5000).astype(np.uint32)
    for keys in (asc, asc[::-1].copy()):
        t = to_dev(keys, gpu)
        buf = ts.SortByKeyBuffer()
        ts.sort_by_key_with_buffer(t, buf)
        want = np.argsort(keys, kind="stable").astype(np.uint32)
        assert np.array_equal(to_host(buf.order(), np.uint32), want)
        assert np.array_equal(to_host(t, np.uint32), keys[want])
