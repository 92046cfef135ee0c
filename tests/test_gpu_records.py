"""By-key sorts that carry 8- and 16-byte records through the radix passes,
and the by-key sort beyond u32::MAX elements, against the oracle.

* Records of 8 or 16 bytes aligned to their size move through the passes
  (each pass reads the key out of the record) instead of being gathered by
  the order afterwards.  For every key width and every aligned key offset,
  for signed and float keys, for odd and even numbers of live passes, for
  presorted, descending and all-equal keys, order() must equal the oracle's
  stable argsort and the records must end in that order, byte for byte.
  A misaligned slice of such records takes the gather path and must agree.
* ts_sort_by_key_unordered -- the path of sort_by_key above u32::MAX
  elements (by_key.rs:190-194), which leaves no order -- is checked at
  ordinary sizes through the C ABI against the oracle's stable order, for
  keys inside the records (carried, or sorted as 16-byte (key, position)
  records and gathered) and keys in an array of their own.
* Above u32::MAX elements: 2^32 + 4099 8-byte records, each its own index
  above a 16-bit key, and 2^32 + 5000 u32 ids keyed by a separate u8 array;
  the outputs are checked on the device by properties (each record keeps its
  key, keys non-decreasing, indices increasing within equal keys, the
  records a permutation of the input: distinct indices in range, or equal
  per-key counts and sums), and order() stays empty.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import turbosort_amd as ts
from gpu_util import bits, to_dev, to_host
from oracle import closed_form as cf
from shapes import random_of

pytestmark = pytest.mark.gpu

NAMES = {1: ("u8", "i8"), 2: ("u16", "i16"), 4: ("u32", "i32", "f32"), 8: ("u64", "i64", "f64")}
TORCH = {"u8": torch.uint8, "i8": torch.int8, "u16": torch.uint16, "i16": torch.int16, "u32": torch.uint32,
         "i32": torch.int32, "f32": torch.float32, "u64": torch.uint64, "i64": torch.int64, "f64": torch.float64}


def make_records(n: int, stride: int, name: str, offset: int, seed: int, few: bool) -> tuple[np.ndarray, np.ndarray]:
    """n records of random bytes with a key of dtype `name` planted at `offset`."""
    g = np.random.default_rng(seed)
    r = g.integers(0, 256, size=(n, stride), dtype=np.uint8)
    k = random_of(name, n, seed=seed + 1)
    if few:
        k = k[g.integers(0, min(9, n), size=n)]  # ties, so that stability shows
    w = k.dtype.itemsize
    r[:, offset:offset + w] = k.view(np.uint8).reshape(n, w)
    return r, k


def by_key(r: np.ndarray, name: str, offset: int, gpu, misalign: bool = False):
    n, stride = r.shape
    if misalign:
        base = torch.empty(n * stride + 8, dtype=torch.uint8, device=gpu)
        t = base[8:].view(n, stride)
        t.copy_(torch.from_numpy(r))
    else:
        t = torch.from_numpy(r.copy()).to(gpu)
    buf = ts.SortByKeyBuffer()
    ts.sort_by_key_with_buffer(t, buf, ts.KeyField(offset, TORCH[name]))
    return to_host(buf.order(), np.uint32), t.cpu().numpy()


CASES = [(stride, w, off, NAMES[w][(off // w) % len(NAMES[w])])
         for stride in (8, 16) for w in (1, 2, 4, 8) for off in range(0, stride, w)]


@pytest.mark.parametrize("stride,w,off,name", CASES)
def test_carried_records_every_key_width_and_offset(gpu, stride, w, off, name):
    for n, few in ((70_001, True), (5_000, False)):
        r, k = make_records(n, stride, name, off, seed=stride * 100 + off * 10 + w, few=few)
        want = cf.order(k)
        order, out = by_key(r, name, off, gpu)
        assert np.array_equal(order, want), (n, few)
        assert np.array_equal(out, r[want.astype(np.int64)]), (n, few)


@pytest.mark.parametrize("stride", [8, 16])
def test_carried_pass_counts_and_presorted_inputs(gpu, stride):
    """One live digit (an odd pass count ends in the scratch half and is
    copied back), three live digits, dead digits only, keys already in
    order, non-increasing keys with ties, and 10^6 f32 scores."""
    n = 300_000
    g = np.random.default_rng(stride)
    shapes = {
        "one digit": (g.integers(0, 256, n) << 8).astype(np.uint32),
        "three digits": (g.integers(0, 1 << 24, n) | (7 << 24)).astype(np.uint32),
        "all equal": np.full(n, 12345, dtype=np.uint32),
        "ascending": np.sort(g.integers(0, 5000, n)).astype(np.uint32),
        "descending ties": np.sort(g.integers(0, 5000, n))[::-1].astype(np.uint32).copy(),
    }
    for label, k in shapes.items():
        r = g.integers(0, 256, size=(n, stride), dtype=np.uint8)
        r[:, 4:8] = k.view(np.uint8).reshape(n, 4)
        want = np.argsort(k, kind="stable").astype(np.uint32)
        order, out = by_key(r, "u32", 4, gpu)
        assert np.array_equal(order, want), label
        assert np.array_equal(out, r[want.astype(np.int64)]), label
    n = 1_000_003
    r, k = make_records(n, stride, "f32", 4, seed=9, few=False)
    want = cf.order(k)
    order, out = by_key(r, "f32", 4, gpu)
    assert np.array_equal(order, want)
    assert np.array_equal(out, r[want.astype(np.int64)])


@pytest.mark.parametrize("stride,name,off", [(16, "f32", 4), (16, "u64", 8), (8, "i32", 4)])
def test_misaligned_records_take_the_gather_path(gpu, stride, name, off):
    n = 100_000
    r, k = make_records(n, stride, name, off, seed=stride + off, few=True)
    want = cf.order(k)
    order, out = by_key(r, name, off, gpu, misalign=True)
    assert np.array_equal(order, want)
    assert np.array_equal(out, r[want.astype(np.int64)])


def unordered(ctx, name, keys: torch.Tensor, key_stride: int, key_off: int, records: torch.Tensor, n: int,
              record_bytes: int):
    code = ts._dtype_code(TORCH[name])
    rc = ctx._lib.ts_sort_by_key_unordered(ctx.handle, code, ctypes.c_void_p(keys.data_ptr()), key_stride, key_off,
                                           ctypes.c_void_p(records.data_ptr()), n, record_bytes,
                                           ts._stream(records.device))
    ctx.check(rc)


@pytest.mark.parametrize("stride,name,off", [(16, "f64", 8), (8, "u16", 2), (12, "i32", 8), (24, "u64", 16),
                                             (4, "u8", 3), (3, "u8", 1), (8, "u64", 0)])
def test_unordered_sort_with_keys_inside_the_records(gpu, stride, name, off):
    ctx = ts.Context(gpu)
    for n in (2, 3000, 250_001):
        r, k = make_records(n, stride, name, off, seed=n + stride, few=True)
        t = torch.from_numpy(r.copy()).to(gpu)
        unordered(ctx, name, t, stride, off, t, n, stride)
        want = cf.order(k).astype(np.int64)
        assert np.array_equal(t.cpu().numpy(), r[want]), (n, stride, name)
    ctx.check_device()


@pytest.mark.parametrize("name", ["u8", "i16", "f32", "u64"])
def test_unordered_sort_with_keys_of_their_own(gpu, name):
    """The keys a closure gave, in an array of their own; the records (12
    bytes here) are put in the keys' stable order, the keys stay as they are."""
    ctx = ts.Context(gpu)
    n = 400_000
    g = np.random.default_rng(3)
    k = random_of(name, n, seed=11)[g.integers(0, 50, n)]
    r = g.integers(0, 256, size=(n, 12), dtype=np.uint8)
    kt = to_dev(k, gpu)
    t = torch.from_numpy(r.copy()).to(gpu)
    unordered(ctx, name, kt, k.dtype.itemsize, 0, t, n, 12)
    want = cf.order(k).astype(np.int64)
    assert np.array_equal(t.cpu().numpy(), r[want])
    assert np.array_equal(bits(to_host(kt, k.dtype)), bits(k))


def test_ts_sort_by_key_without_gather_above_u32_max_is_too_long(gpu):
    """An order cannot index more than u32::MAX elements: asking for one is
    TS_ERR_TOO_LONG, before anything is read."""
    ctx = ts.Context(gpu)
    t = torch.empty(16, dtype=torch.uint8, device=gpu)
    rc = ctx._lib.ts_sort_by_key(ctx.handle, ts._dtype_code(torch.uint32), ctypes.c_void_p(t.data_ptr()),
                                 (1 << 32) + 1, 4, 0, None, 0, ts._stream(t.device))
    assert rc == 3


def mixed_key(idx: torch.Tensor, bits_: int) -> torch.Tensor:
    """A key of `bits_` bits from an index (int64 in, int64 out), a fixed multiplicative hash."""
    return ((idx * 0x5851F42D4C957F2D) >> 40) & ((1 << bits_) - 1)


def free_gpu():
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_above_u32_max_records_carried(gpu):
    """2^32 + 4099 8-byte records r_i = i << 16 | key(i), key 16 bits,
    sorted by the u16 at offset 0: carried through the passes, no order."""
    n = (1 << 32) + 4099
    chunk = 1 << 28
    rec = torch.empty(n, dtype=torch.int64, device=gpu)
    for a in range(0, n, chunk):
        i = torch.arange(a, min(n, a + chunk), dtype=torch.int64, device=gpu)
        rec[a:a + i.numel()] = (i << 16) | mixed_key(i, 16)
    del i
    buf = ts.SortByKeyBuffer()
    ts.sort_by_key_with_buffer(rec, buf, ts.KeyField(0, torch.uint16))
    assert buf.order().numel() == 0
    with pytest.raises(ts.TurbosortPanic):
        buf.apply_order(torch.zeros(3, dtype=torch.int64, device=gpu))
    del buf
    prev = None
    for a in range(0, n, chunk):
        x = rec[a:min(n, a + chunk)]
        key, idx = x & 0xFFFF, x >> 16
        assert bool((key == mixed_key(idx, 16)).all()), a  # each record is whole
        comb = (key << 40) | idx  # (key, index) order: sorted by key, stable
        assert bool((comb[1:] > comb[:-1]).all()), a
        if prev is not None:
            assert int(comb[0]) > prev
        prev = int(comb[-1])
        assert bool(((idx >= 0) & (idx < n)).all())
    del rec, x, key, idx, comb
    free_gpu()


def test_above_u32_max_keys_of_their_own(gpu):
    """2^32 + 5000 u32 records (the low 32 bits of each index) keyed by a
    separate u8 array key(i mod 2^32): the (key, position) records are sorted
    and the ids gathered.  Within each key the ids increase, except once,
    where the indices pass 2^32 and the 5000 ids below 5000 follow."""
    extra = 5000
    n = (1 << 32) + extra
    chunk = 1 << 28
    keys = torch.empty(n, dtype=torch.uint8, device=gpu)
    ids = torch.empty(n, dtype=torch.int32, device=gpu)
    for a in range(0, n, chunk):
        i = torch.arange(a, min(n, a + chunk), dtype=torch.int64, device=gpu) & 0xFFFFFFFF
        keys[a:a + i.numel()] = mixed_key(i, 8).to(torch.uint8)
        ids[a:a + i.numel()] = i.to(torch.int32)
    del i

    def tally(ids_):
        """Per key: how many ids, and the sum of the ids (mod 2^64)."""
        cnt = torch.zeros(256, dtype=torch.int64, device=gpu)
        tot = torch.zeros(256, dtype=torch.int64, device=gpu)
        for a in range(0, n, chunk):
            x = ids_[a:min(n, a + chunk)].to(torch.int64) & 0xFFFFFFFF
            k = mixed_key(x, 8)
            cnt.index_add_(0, k, torch.ones_like(k))
            tot.index_add_(0, k, x)
        return cnt, tot

    before = tally(ids)
    ts.sort_by_key(ids, lambda _: keys)
    after = tally(ids)
    assert torch.equal(before[0], after[0]) and torch.equal(before[1], after[1])
    prev_key = -1
    descents = torch.zeros(256, dtype=torch.int64, device=gpu)
    last = None
    for a in range(0, n, chunk):
        x = ids[a:min(n, a + chunk)].to(torch.int64) & 0xFFFFFFFF
        k = mixed_key(x, 8)
        if last is not None:
            x = torch.cat([last[0], x])
            k = torch.cat([last[1], k])
        assert bool((k[1:] >= k[:-1]).all()), a  # keys non-decreasing
        assert int(k[0]) >= prev_key
        same = k[1:] == k[:-1]
        down = same & (x[1:] < x[:-1])
        descents.index_add_(0, k[1:][down], torch.ones(int(down.sum()), dtype=torch.int64, device=gpu))
        # a descent restarts at an id below `extra`
        assert bool((x[1:][down] < extra).all())
        prev_key = int(k[-1])
        last = (x[-1:], k[-1:])
    assert int(descents.max()) <= 1  # each key's run: ascending ids, then the wrapped tail
    assert int(descents.sum()) >= 1
    del keys, ids, x, k
    free_gpu()
