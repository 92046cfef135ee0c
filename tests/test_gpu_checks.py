"""The device-side checks, and configurations no other GPU test covers.

* ts_checksum, the fingerprint behind every full-size "sorted, same keys"
  claim: equal to a numpy restatement (sum and xor of a one-multiply
  xor-shift mixer of each key's salted bits, and the count of adjacent pairs
  out of order in the type's total order); one swapped pair or one changed
  key changes it.
* The sticky device error word (ts_ctx_check): a segment longer than the
  max_segment_len given to an asynchronous segmented sort, and invalid
  offsets in such a call, are reported by the next check and then cleared.
* A sort of more than 2^30 keys, whose look-back words are 64-bit.
* The part of a long keys-only segment that falls back to the radix sort at
  2^21 keys (small scatter tiles) inside a sort sized for more keys.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import turbosort_amd as ts
from gpu_util import bits, to_dev, to_host
from oracle import closed_form as cf
from shapes import random_of, rng

pytestmark = pytest.mark.gpu

SALT = np.uint64(0x632BE59BD9B4E019)
MUL = np.uint64(0xD6E8FEB86659FD93)


def mix(z: np.ndarray) -> np.ndarray:
    """The fingerprint's mixer: xor-shift 32, multiply, xor-shift 32 (mod 2^64)."""
    with np.errstate(over="ignore"):
        z = z ^ (z >> np.uint64(32))
        z = z * MUL
        return z ^ (z >> np.uint64(32))


def fingerprint(x: np.ndarray):
    """Sum and xor (mod 2^64) of mix(bits + salt), and the unsorted pairs."""
    with np.errstate(over="ignore"):
        h = mix(bits(x).astype(np.uint64) + SALT)
    s = int(np.sum(h, dtype=np.uint64))
    xr = int(np.bitwise_xor.reduce(h)) if h.size else 0
    k = cf.to_radix(x)
    unsorted = int(np.count_nonzero(k[:-1] > k[1:])) if x.size > 1 else 0
    return s, xr, unsorted


@pytest.mark.parametrize("name", ["u8", "i16", "u32", "f32", "u64", "f64"])
def test_checksum_equals_numpy_restatement(gpu, name):
    x = random_of(name, 1_000_003, seed=31)
    t = to_dev(x, gpu)
    assert ts.checksum(t) == fingerprint(x)
    s = cf.sort(x)
    t = to_dev(s, gpu)
    got = ts.checksum(t)
    assert got == fingerprint(s) and got[2] == 0
    assert got[:2] == fingerprint(x)[:2]  # a permutation keeps the multiset fingerprint


def test_checksum_catches_a_swapped_pair_and_a_changed_key(gpu):
    x = cf.sort(random_of("u64", 1_000_000, seed=5))
    base = ts.checksum(to_dev(x, gpu))
    assert base[2] == 0
    y = x.copy()
    y[[1000, 1001]] = y[[1001, 1000]]
    swapped = ts.checksum(to_dev(y, gpu))
    assert swapped[:2] == base[:2] and swapped[2] >= 1
    z = x.copy()
    z[777] ^= np.uint64(1)
    changed = ts.checksum(to_dev(z, gpu))
    assert changed[0] != base[0] and changed[1] != base[1]
    assert changed == fingerprint(z)


def test_ctx_check_is_clean_after_ordinary_work(gpu):
    c = ts.Context(gpu)
    t = to_dev(random_of("u32", 1 << 20, seed=2), gpu)
    ts.sort(t, ctx=c)
    c.check_device()


def test_async_segment_longer_than_the_bound_is_reported(gpu):
    """An asynchronous segmented sort trusts max_segment_len; a segment
    longer than it is recorded on the device and reported by the next check,
    which then clears the record."""
    c = ts.Context(gpu)
    keys = to_dev(random_of("u32", 5000, seed=3), gpu)
    offsets = torch.tensor([0, 100, 3000, 5000], dtype=torch.uint64, device=gpu)
    ts.segmented_sort(keys, offsets, max_segment_len=256, ctx=c)  # returns without waiting
    with pytest.raises(RuntimeError, match="longer than max_segment_len"):
        c.check_device()
    c.check_device()  # cleared
    # Within the bound nothing is recorded, and the segments are sorted.
    x = random_of("u32", 5000, seed=4)
    keys = to_dev(x, gpu)
    ts.segmented_sort(keys, offsets, max_segment_len=2900, ctx=c)
    c.check_device()
    got = to_host(keys, np.uint32)
    off = [0, 100, 3000, 5000]
    for a, b in zip(off[:-1], off[1:]):
        assert np.array_equal(got[a:b], np.sort(x[a:b]))


def test_async_invalid_offsets_are_reported(gpu):
    c = ts.Context(gpu)
    keys = to_dev(random_of("u32", 1000, seed=5), gpu)
    offsets = torch.tensor([0, 600, 400, 1000], dtype=torch.uint64, device=gpu)  # decreasing pair
    ts.segmented_sort(keys, offsets, max_segment_len=1000, ctx=c)
    with pytest.raises(RuntimeError, match="offsets"):
        c.check_device()
    c.check_device()


def test_split_part_of_2p21_keys_in_a_longer_sort(gpu):
    """Keys only, no split rounds: a 2^21-key segment of a 2^21 + 1 key array
    goes straight to the radix sort, whose small tiles need more look-back
    rows than the array's size alone would reserve."""
    c = ts.Context(gpu)
    c.set_split_rounds(0)
    n = (1 << 21) + 1
    for name in ("u32", "u64"):
        x = random_of(name, n, seed=9)
        keys = to_dev(x, gpu)
        offsets = torch.tensor([0, 1 << 21, n], dtype=torch.uint64, device=gpu)
        ts.segmented_sort(keys, offsets, ctx=c)
        got = to_host(keys, x.dtype)
        want = np.concatenate([cf.sort(x[: 1 << 21]), x[1 << 21:]])
        assert np.array_equal(bits(got), bits(want)), name
    c.check_device()


def test_more_than_2p30_keys_use_wide_look_back_words(gpu):
    """2^30 + 17 u32 keys (SplitMix64 state 2), sorted on the device: the
    look-back status words are 64-bit from 2^30 keys on.  The output keeps
    the multiset fingerprint and has no pair out of order."""
    c = ts.Context(gpu)
    n = (1 << 30) + 17
    t = torch.empty(n, dtype=torch.uint32, device=gpu)
    ts.fill_splitmix64(t, 2, ctx=c)
    head = to_host(t[:1000], np.uint32)
    assert np.array_equal(head, cf.u32_keys(2, 1000))
    s0, x0, _ = ts.checksum(t, ctx=c)
    ts.sort(t, ctx=c)
    info = c.last_sort_info()
    s1, x1, unsorted = ts.checksum(t, ctx=c)
    assert (s0, x0) == (s1, x1) and unsorted == 0
    assert info["executed_passes"] == 4, info
    c.check_device()
    del t
    torch.cuda.empty_cache()
