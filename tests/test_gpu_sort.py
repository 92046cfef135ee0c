"""Parity of the gfx950 sort with the oracle, bit for bit.

Every test calls libtsort.so through its C ABI (via turbosort_amd's ctypes
mirror of the reference API) and compares with the closed-form oracle on the
same seeded input; the reference's own known answers (tests/golden) run too.
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
from shapes import adversarial_u32, float_specials_f32, random_of, rng, structured_u64

pytestmark = pytest.mark.gpu


def gpu_sort(x: np.ndarray, dev, buffer_len=None) -> np.ndarray:
    t = to_dev(x, dev)
    if buffer_len is None:
        ts.sort(t)
    else:
        buf = torch.empty(buffer_len, dtype=t.dtype, device=dev)
        ts.sort_with_buffer(t, buf)
    return to_host(t, x.dtype)


def check(x: np.ndarray, dev, what="", buffer_len=None):
    got = gpu_sort(x, dev, buffer_len)
    want = cf.sort(x)
    if not np.array_equal(bits(got), bits(want)):
        bad = np.nonzero(bits(got) != bits(want))[0]
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {bad[:5]} "
                             f"(n={x.size}, dtype={x.dtype})")


def test_c1_one_million_u32_splitmix_seed1(gpu):
    """The 1M-key parity workload: device generator and sort vs both oracles."""
    n = 1_000_000
    t = torch.empty(n, dtype=torch.uint32, device=gpu)
    ts.fill_splitmix64(t, 1)
    x = cf.u32_keys(1, n)
    assert np.array_equal(to_host(t, np.uint32), x), "device generator differs from the oracle's"
    ts.sort(t)
    got = to_host(t, np.uint32)
    assert np.array_equal(got, cf.sort(x))
    assert np.array_equal(got, nv.sort(x))


@pytest.mark.parametrize("i", range(len(kats()["sort"])))
def test_reference_kats(gpu, i):
    k = kats()["sort"][i]
    x = kat_array(k["input"], k["dtype"])
    want = bits(kat_array(k["expected"], k["dtype"]))
    assert np.array_equal(bits(gpu_sort(x, gpu)), want), k["src"]
    assert np.array_equal(bits(gpu_sort(x, gpu, buffer_len=x.size)), want), k["src"]


SIZES = [0, 1, 2, 3, 15, 16, 17, 48, 128, 511, 512, 513, 1024, 4095, 4096, 4097, 8191, 8192, 65_537, 300_000]


@pytest.mark.parametrize("name", cf.DTYPES)
def test_random_every_dtype_and_tier(gpu, name):
    for n in SIZES:
        check(random_of(name, n, seed=1000 + n), gpu, f"{name} n={n}")


@pytest.mark.parametrize("name", cf.DTYPES)
def test_both_tile_shapes_every_dtype(gpu, name):
    """Keys-only sorts up to 2^21 keys run small scatter tiles, larger ones
    the large tiles: both sides of the switch and a ragged larger size."""
    for n in (1 << 21, (1 << 21) + 1, 3_000_017):
        check(random_of(name, n, seed=n + 7), gpu, f"{name} n={n}")
    check(random_of(name, 3_000_017, seed=11), gpu, f"{name} buffer", buffer_len=3_000_017)


@pytest.mark.parametrize("name", cf.DTYPES)
def test_sort_with_buffer_every_dtype(gpu, name):
    for n in (17, 513, 4097, 70_000):
        check(random_of(name, n, seed=n), gpu, f"{name} n={n}", buffer_len=n + 5)


@pytest.mark.parametrize("name", cf.DTYPES)
def test_one_workgroup_tier_every_dtype(gpu, name):
    """Keys-only sorts of 4097 keys up to one workgroup's LDS capacity (16384
    keys of 4 bytes and narrower, 17408 of 8) run the segment kernel for their
    length on one segment: both sides of every bound, random keys, few
    distinct values (crowded bins: the kernels' LSD fallback) and descending
    keys, through sort and sort_with_buffer, whose buffer holds the segment
    list when it is 8-byte aligned (an unaligned view falls back to the
    workspace)."""
    width = np.dtype(cf.NP_DTYPES[name]).itemsize
    cap = 17408 if width == 8 else 16384
    g = rng(width * 31 + len(name))
    for n in (4097, 5000, 8191, 8192, 8193, 12_345, cap - 1, cap, cap + 1):
        x = random_of(name, n, seed=n + 11)
        check(x, gpu, f"{name} n={n}")
        check(x, gpu, f"{name} n={n} buffer", buffer_len=n)
        check(x[g.integers(0, 3, n)], gpu, f"{name} n={n} few distinct")
        check(cf.sort(x)[::-1].copy(), gpu, f"{name} n={n} descending")
    n = 9000
    x = random_of(name, n, seed=5)
    t = to_dev(x, gpu)
    spare = torch.empty(n + 1, dtype=t.dtype, device=gpu)
    ts.sort_with_buffer(t, spare[1:])
    assert np.array_equal(bits(to_host(t, x.dtype)), bits(cf.sort(x))), f"{name} unaligned buffer"


def test_sort_with_buffer_too_small_panics(gpu):
    t = torch.zeros(10, dtype=torch.uint32, device=gpu)
    buf = torch.zeros(9, dtype=torch.uint32, device=gpu)
    with pytest.raises(ts.TurbosortPanic, match=r"^buffer too small: need 10, got 9$"):
        ts.sort_with_buffer(t, buf)
    # The reference checks the length before looking at the slice (dispatch.rs:78-83).
    with pytest.raises(ts.TurbosortPanic, match=r"^buffer too small: need 1, got 0$"):
        ts.sort_with_buffer(t[:1], buf[:0])


def test_structured_u64(gpu):
    # tests/correctness.rs:129-190, through both entry points
    for n in (70_000, 200_000):
        for shape, x in structured_u64(n, seed=77).items():
            check(x, gpu, shape)
            check(x, gpu, shape, buffer_len=n)


def test_adversarial_patterns(gpu):
    # tests/correctness.rs:577-635, 784-792
    for n in (8, 16, 17, 100, 512, 513, 1000, 200_000):
        for shape, x in adversarial_u32(n).items():
            check(x, gpu, f"{shape} n={n}")


def test_presorted_and_duplicates(gpu):
    # tests/correctness.rs:327-400
    for n in (17, 100, 128, 129, 512, 513, 5000, 200_000):
        check((np.arange(n, dtype=np.uint32)[::-1] // 3).copy(), gpu, "reverse with duplicates")
        check(((np.arange(n)[::-1] // 3) - 50).astype(np.float64), gpu, "f64 reverse")
        check((np.arange(n)[::-1] % 60_000).astype(np.int16), gpu, "i16 reverse")
        x = np.arange(n, dtype=np.uint64)
        x[[-2, -1]] = x[[-1, -2]]
        check(x, gpu, "sorted except the tail")
        y = np.arange(n, dtype=np.uint64)[::-1].copy()
        y[-1] = np.iinfo(np.uint64).max
        check(y, gpu, "descending with max tail", buffer_len=n)
    check(np.full(1000, 42, dtype=np.uint32), gpu, "all equal")
    check(np.full(100_000, 42, dtype=np.uint32), gpu, "all equal, radix tier")


def test_float_specials(gpu):
    for n in (100, 4096, 5000, 250_000):
        x = float_specials_f32(n, seed=n)
        check(x, gpu, f"f32 specials n={n}")
        check(x.astype(np.float64), gpu, f"f64 specials n={n}")


def test_equal_bucket_inputs(gpu):
    # tests/correctness.rs:196-245: permutations and pipe organs of every width
    g = np.random.default_rng(1234)
    for n in (16_384, 65_536 + 17, 200_000):
        perm = g.permutation(n).astype(np.uint32)
        organ = np.where(np.arange(n) < n // 2, 2 * np.arange(n), 2 * (n - 1 - np.arange(n))).astype(np.uint32)
        for base in (perm, organ):
            check(base, gpu, "u32")
            check(base.astype(np.uint64) * 3, gpu, "u64", buffer_len=n)
            check(base.astype(np.uint16), gpu, "u16")
            check(base.astype(np.float32) - 1000.0, gpu, "f32")


def test_context_reuse_across_sizes_and_types(gpu):
    ctx = ts.Context(gpu)
    for i, (name, n) in enumerate([("u32", 300_000), ("u64", 5000), ("f32", 1_000_000), ("i8", 70_000),
                                   ("u32", 10_000), ("f64", 200_000)]):
        x = random_of(name, n, seed=i)
        t = to_dev(x, gpu)
        ts.sort(t, ctx=ctx)
        assert np.array_equal(bits(to_host(t, x.dtype)), bits(cf.sort(x))), (name, n)


def test_sorts_on_a_side_stream(gpu):
    x = random_of("u32", 500_000, seed=3)
    t = to_dev(x, gpu)
    s = torch.cuda.Stream(gpu)
    s.wait_stream(torch.cuda.current_stream(gpu))
    with torch.cuda.stream(s):
        ts.sort(t)
    torch.cuda.current_stream(gpu).wait_stream(s)
    assert np.array_equal(to_host(t, np.uint32), np.sort(x))


def test_both_ranking_methods(gpu):
    """The scatter passes rank a wave's keys by LDS atomic adds where the
    device passed the context's lane-order check (MI355X does), else by
    ballots.  Both must give the oracle's output: keys, and the stable order of
    equal keys (a long segment of the segmented sort runs the by-key passes
    with positions as values)."""
    ctx = ts.Context(gpu)
    assert ctx.ranking == "atomic"
    g = rng(91)
    skew = np.where(g.random(600_000) < 0.9, np.uint32(12345), random_of("u32", 600_000, seed=92))
    dup = (random_of("u32", 400_000, seed=93) % 1000).astype(np.uint32)
    for method in ("ballot", "atomic"):
        ctx.set_ranking(method)
        assert ctx.ranking == method
        for name, n in (("u32", 1_000_003), ("u64", 700_001), ("f32", 65_537), ("i16", 300_000), ("f64", 250_000)):
            x = random_of(name, n, seed=n)
            t = to_dev(x, gpu)
            ts.sort(t, ctx=ctx)
            assert np.array_equal(bits(to_host(t, x.dtype)), bits(cf.sort(x))), (method, name, n)
        t = to_dev(skew, gpu)
        ts.sort(t, ctx=ctx)
        assert np.array_equal(to_host(t, np.uint32), np.sort(skew)), method
        k = to_dev(dup, gpu)
        v = torch.arange(dup.size, dtype=torch.int32, device=gpu)
        off = torch.tensor([0, dup.size], dtype=torch.int64, device=gpu)
        ts.segmented_sort(k, off, v, ctx=ctx)
        want = np.argsort(dup, kind="stable").astype(np.uint32)
        assert np.array_equal(v.cpu().numpy().astype(np.uint32), want), method
        assert np.array_equal(to_host(k, np.uint32), dup[want]), method


@pytest.mark.parametrize("pinned", [False, True])
def test_sort_host(gpu, pinned):
    """sort_host: a host slice is copied in, sorted on the device and copied
    back (lib.rs:99-101 on host memory), pageable or pinned, reusing the
    context's device buffer across sizes."""
    ctx = ts.Context(gpu)
    for name, n in (("u32", 0), ("u32", 1), ("f32", 2), ("u64", 4097), ("i16", 70_000), ("f64", 300_000),
                    ("u32", 5_000_000), ("u8", 1000)):
        x = random_of(name, n, seed=n + 3)
        t = torch.from_numpy(x.copy())
        if pinned:
            t = t.pin_memory()
        ts.sort_host(t, ctx=ctx)
        assert np.array_equal(bits(t.numpy()), bits(cf.sort(x))), (name, n, pinned)
    with pytest.raises(ValueError):
        ts.sort_host(torch.zeros(4, dtype=torch.uint32, device=gpu), ctx=ctx)


@pytest.mark.parametrize("n", [4097, 100_000, 1_000_003, 2_500_001])
def test_presorted_shortcuts(gpu, n):
    """Non-decreasing input skips every pass and non-increasing input is
    reversed (presorted.rs:15-65); any single out-of-place pair, including at
    the ends and at workgroup-chunk boundaries, must fall back to the sort."""
    g = rng(n)
    for name in ("u32", "u64", "f32", "i16", "f64"):
        x = cf.sort(random_of(name, n, seed=n + 1))
        check(x, gpu, f"{name} ascending")
        check(x[::-1].copy(), gpu, f"{name} descending")
        ties = np.repeat(x[: n // 3], 3)[:n]
        check(ties[::-1].copy(), gpu, f"{name} descending with ties")
        for pos in {1, n // 2, n - 1, int(g.integers(1, n - 1)), 65_536 if n > 65_536 else n // 3}:
            y = x.copy()
            y[pos - 1], y[pos] = y[pos], y[pos - 1]
            check(y, gpu, f"{name} ascending, pair {pos} swapped")
            z = x[::-1].copy()
            z[pos - 1], z[pos] = z[pos], z[pos - 1]
            check(z, gpu, f"{name} descending, pair {pos} swapped")


@pytest.mark.parametrize("name,n", [("u32", 1_000_003), ("u32", 3_000_017), ("f64", 200_000), ("i16", 70_000),
                                    ("u64", 3_000), ("f32", 10_000)])
def test_sort_replays_from_a_graph(gpu, name, n):
    """sort makes no host call once the workspace has grown, so it can be
    captured in a graph; every replay sorts freshly restored input exactly,
    with eager work and synchronisations between the replays."""
    x = random_of(name, n, seed=n + 3)
    want = bits(cf.sort(x))
    t = to_dev(x, gpu)
    src = t.clone()
    ctx = ts.Context(gpu)
    ts.sort(t, ctx=ctx)  # grows the workspace before capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        t.copy_(src)
        with torch.cuda.graph(g, stream=s):
            ts.sort(t, ctx=ctx)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    for _ in range(3):
        t.copy_(src)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(bits(to_host(t, x.dtype)), want)

