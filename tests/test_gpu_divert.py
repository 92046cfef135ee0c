"""The divert of 8-byte keys (radix/mod.rs:240-242,268-270; divert.rs:38-90)
on the GPU, bit for bit against the oracle.

Sorts of u64 / i64 / f64 without values sort only the top live digits that
carry about log2(n) + 2 bits (decided on the device from the histograms,
divert.rs:38-54) and then finish every run of keys sharing those digits
(k_finish).  A run longer than the finish sorts that is out of order makes
the device sort every digit after all.  Each case also checks what the sort
did (ts_ctx_last_sort_info), so that both the finished and the fallback
branches are known to have run.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import turbosort_amd as ts
from gpu_util import bits, to_dev, to_host
from oracle import closed_form as cf
from oracle import native as nv
from shapes import random_of, rng, structured_u64

pytestmark = pytest.mark.gpu

SMALL_N = 1 << 21


@pytest.fixture(scope="module")
def dctx(gpu):
    c = ts.Context(gpu)
    c.set_divert_min(SMALL_N + 1)
    return c


def sort_on(x: np.ndarray, dev, ctx, buffer=False):
    t = to_dev(x, dev)
    if buffer:
        ts.sort_with_buffer(t, torch.empty_like(t), ctx=ctx)
    else:
        ts.sort(t, ctx=ctx)
    info = ctx.last_sort_info()
    return to_host(t, x.dtype), info


FINISH_RUN = 256  # the longest run the finish sorts itself


def expected_divert(x: np.ndarray):
    """The reference's decision on these keys (skippable_digits over the
    largest bucket of each live digit, divert.rs:38-54, in the oracle's
    restatement) and whether a run longer than the finish's limit comes out
    of the top passes out of order: (live digits, skipped, second sort)."""
    k = cf.to_radix(x).astype(np.uint64)
    n = k.size
    live, maxb = [], []
    for p in range(8):
        h = np.bincount(((k >> np.uint64(8 * p)) & np.uint64(0xFF)).astype(np.int64), minlength=256)
        if h.max() < n:
            live.append(p)
            maxb.append(int(h.max()))
    skip = nv.skippable_digits(len(live), n, lambda i: maxb[i])
    if skip == 0:
        return len(live), 0, 0
    prefix = k >> np.uint64(8 * live[skip])
    arranged = k[np.argsort(prefix, kind="stable")]
    pre = prefix[np.argsort(prefix, kind="stable")]
    starts = np.flatnonzero(np.concatenate([[True], pre[1:] != pre[:-1], [True]]))
    second = 0
    for a, b in zip(starts[:-1], starts[1:]):
        if b - a > FINISH_RUN and np.any(arranged[a + 1:b] < arranged[a:b - 1]):
            second = 1
            break
    return len(live), skip, second


def check(x: np.ndarray, dev, ctx, what: str, buffer=False) -> dict:
    got, info = sort_on(x, dev, ctx, buffer)
    want = cf.sort(x)
    if not np.array_equal(bits(got), bits(want)):
        bad = np.nonzero(bits(got) != bits(want))[0]
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {bad[:5]}; {info}")
    ctx.check_device()
    if x.size > SMALL_N and x.size >= ctx.divert_min:
        live, skip, second = expected_divert(x)
        assert (info["live_digits"], info["skipped_digits"]) == (live, skip), (what, info, (live, skip, second))
        # The second sort runs only when a run too long for the finish is out
        # of order (the finish's first rounds may still put a nearly sorted
        # one in order).
        assert info["fallback"] <= second, (what, info, (live, skip, second))
        assert info["executed_passes"] == (live if info["fallback"] else live - skip), info
    return info


@pytest.mark.parametrize("name", ["u64", "i64", "f64"])
@pytest.mark.parametrize("n", [3_000_017, 9_000_001])
def test_uniform_8byte_keys_divert(gpu, dctx, name, n):
    """Uniform bit patterns: 3M keys sort three top digits (an odd count:
    the finish moves every key from the scratch half), 9M keys four (even:
    the finish works in place)."""
    x = random_of(name, n, seed=n + len(name))
    info = check(x, gpu, dctx, f"{name} n={n}")
    assert info["skipped_digits"] == (5 if n < (1 << 22) else 4), info
    assert info["executed_passes"] == 8 - info["skipped_digits"] and info["fallback"] == 0, info
    info = check(x, gpu, dctx, f"{name} n={n} buffer", buffer=True)
    assert info["skipped_digits"] > 0


def test_structured_u64_shapes(gpu, dctx):
    """The reference's structured u64 shapes (tests/correctness.rs:129-190):
    whatever the device decides for each, the output is the oracle's."""
    decided = {}
    for shape, x in structured_u64(3_000_017, seed=77).items():
        info = check(x, gpu, dctx, shape)  # the decision equals the reference's
        decided[shape] = info["skipped_digits"]
    assert decided["random"] == 5


@pytest.mark.parametrize("runs", [100, 1000, 30_000])
def test_runs_of_shared_top_bits(gpu, dctx, runs):
    """Top halves from a few random values (every byte of them looks uniform,
    so the low half is left to the finish), low halves random: runs of about
    `runs` keys.  Runs up to the finish's limit are sorted by it; longer
    out-of-order runs send the sort to the second, full sort."""
    n = 3_000_017
    g = rng(runs)
    tops = g.integers(0, 2**32, size=max(1, n // runs), dtype=np.uint64) << np.uint64(32)
    x = tops[g.integers(0, tops.size, size=n)] | g.integers(0, 2**32, size=n, dtype=np.uint64)
    info = check(x, gpu, dctx, f"runs of {runs}")
    assert info["skipped_digits"] > 0, info
    if runs == 1000:
        assert info["fallback"] == 1, info  # the second sort ran


@pytest.mark.parametrize("runs", [500, 2000])
def test_long_runs_already_in_order_need_no_second_sort(gpu, dctx, runs):
    """Low halves that grow with the position: every run of a shared top half
    comes out of the top passes in order, however long, so the finish leaves
    it and no second sort runs."""
    n = 3_000_017
    g = rng(runs + 1)
    tops = g.integers(0, 2**32, size=max(1, n // runs), dtype=np.uint64) << np.uint64(32)
    x = tops[g.integers(0, tops.size, size=n)] | np.arange(n, dtype=np.uint64)
    info = check(x, gpu, dctx, f"ordered runs of {runs}")
    assert info["skipped_digits"] > 0 and info["fallback"] == 0, info


def test_f64_specials_and_negative_runs(gpu, dctx):
    """Float keys: the finish compares radix keys and writes the float bits
    back, NaN of both signs, infinities and -0.0 included."""
    n = 3_000_017
    g = rng(5)
    x = g.standard_normal(n) * 1e6
    b = x.view(np.uint64).copy()
    specials = np.array([0x7FF8000000000000, 0xFFF8000000000000, 0x7FF0000000000000, 0xFFF0000000000000,
                         0x8000000000000000, 0, 0x7FF0000000000001, 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)
    pick = g.random(n) < 0.01
    b[pick] = specials[g.integers(0, specials.size, size=int(pick.sum()))]
    check(b.view(np.float64), gpu, dctx, "f64 specials")


def test_default_threshold_and_small_sizes_do_not_divert(gpu):
    c = ts.Context(gpu)
    assert c.divert_min == 1 << 23
    x = random_of("u64", SMALL_N, seed=3)
    c.set_divert_min(1)
    info = check(x, gpu, c, "small tiles")
    assert info["skipped_digits"] == 0 and info["executed_passes"] == 8, info


def test_zero_extended_u32_in_u64_skips_the_top_digits(gpu):
    """Config C2c's shape at 3M keys: u64 holding zero-extended u32 values
    (SplitMix64 high halves).  The top four digits are dead and no divert
    applies (four live digits carry only 32 bits)."""
    c = ts.Context(gpu)
    n = 3_000_017
    x = cf.u64_zero_extended_u32(4, n)
    info = check(x, gpu, c, "zero-extended u32")
    assert (info["live_digits"], info["executed_passes"], info["skipped_digits"]) == (4, 4, 0), info


def test_zero_extended_u32_in_u64_full_size(gpu):
    """C2c at its full size, 2^28 keys, generated on the device: the sort
    keeps the key multiset (fingerprint before = after), leaves no adjacent
    pair out of order, and runs four scatter passes."""
    c = ts.Context(gpu)
    n = 1 << 28
    t = torch.empty(n, dtype=torch.uint64, device=gpu)
    ts.fill_splitmix64_zero_extended(t, 4, ctx=c)
    s0, x0, _ = ts.checksum(t, ctx=c)
    ts.sort(t, ctx=c)
    info = c.last_sort_info()
    s1, x1, unsorted = ts.checksum(t, ctx=c)
    assert (s0, x0) == (s1, x1) and unsorted == 0
    assert (info["live_digits"], info["executed_passes"]) == (4, 4), info
    c.check_device()


def test_uniform_u64_full_size_diverts(gpu):
    """C2b at its full size, 2^28 uniform u64 keys (SplitMix64 state 3):
    the default context diverts after the top four digits and the finish
    leaves every key in order with the same multiset."""
    c = ts.Context(gpu)
    n = 1 << 28
    t = torch.empty(n, dtype=torch.uint64, device=gpu)
    ts.fill_splitmix64(t, 3, ctx=c)
    s0, x0, _ = ts.checksum(t, ctx=c)
    ts.sort(t, ctx=c)
    info = c.last_sort_info()
    s1, x1, unsorted = ts.checksum(t, ctx=c)
    assert (s0, x0) == (s1, x1) and unsorted == 0
    assert (info["skipped_digits"], info["executed_passes"], info["fallback"]) == (4, 4, 0), info
    c.check_device()
