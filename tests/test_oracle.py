"""Pins the CPU oracle before anything is checked against it.

Two independent oracles exist: the numpy closed form (oracle/closed_form.py)
and the C++ restatement of the reference's algorithm (oracle/restate.cpp).
Each is checked against the reference's own known answers (tests/golden), and
against definitions taken from the mathematics rather than from either
implementation: IEEE-754 totalOrder for the float transforms, Python's integer
order for the integer ones, brute-force permutations, and published SplitMix64
outputs.  Then the two oracles are checked against each other on every tier
and shape the reference's tests exercise.
"""
from __future__ import annotations

import itertools
import math

import numpy as np
import pytest

from conftest import bits_of, kat_array, kats
from oracle import closed_form as cf
from oracle import native as nv
from shapes import adversarial_u32, float_specials_f32, random_of, structured_u64


@pytest.fixture(scope="module", autouse=True)
def built():
    nv.build()


# ------------------------------------------------------------------ KATs
@pytest.mark.parametrize("i", range(len(kats()["sort"])))
def test_sort_kats(i):
    k = kats()["sort"][i]
    x = kat_array(k["input"], k["dtype"])
    want = bits_of(kat_array(k["expected"], k["dtype"]))
    assert np.array_equal(bits_of(cf.sort(x)), want), k["src"]
    assert np.array_equal(bits_of(nv.sort(x)), want), k["src"]
    if k.get("with_buffer"):
        assert np.array_equal(bits_of(nv.sort_with_buffer(x)), want), k["src"]


def test_sort_with_buffer_too_small_message():
    # dispatch.rs:78-83
    with pytest.raises(ValueError, match=r"^buffer too small: need 5, got 4$"):
        nv.sort_with_buffer(np.arange(5, dtype=np.uint32), buffer_len=4)


@pytest.mark.parametrize("i", range(len(kats()["order"])))
def test_order_kats(i):
    k = kats()["order"][i]
    keys = kat_array(k["keys"], k["dtype"])
    want = np.array(k["order"], dtype=np.uint32)
    assert np.array_equal(cf.order(keys), want), k["src"]
    assert np.array_equal(nv.order(keys), want), k["src"]
    if "apply_to" in k:
        other = np.array(k["apply_to"], dtype=np.int64)
        assert nv.apply_order(want, other).tolist() == k["applied"]
        assert cf.apply_order(want, other).tolist() == k["applied"]


def test_apply_order_length_message():
    with pytest.raises(ValueError, match="apply_order: the slice's length differs from the last sorted slice's"):
        nv.apply_order(np.arange(3, dtype=np.uint32), np.arange(4, dtype=np.uint64))


def test_divert_kats():
    d = kats()["divert"]
    for k in d["log2_eighths"]:
        assert nv.log2_eighths(k["x"]) == k["expect"], k["src"]
    for k in d["skippable"]:
        assert nv.skippable_digits(k["n_live"], k["len"], lambda _i: k["max"]) == k["expect"], k["src"]


def test_log2_eighths_brackets_the_true_value():
    # divert.rs:279-290: never above 8*log2(x), less than two eighths below, monotone.
    prev = 0
    for x in range(1, 5000):
        est = nv.log2_eighths(x)
        exact = 8.0 * math.log2(x)
        assert est <= exact + 1e-9 and est > exact - 2.0 and est >= prev, x
        prev = est


def test_finish_runs_sorts_within_runs_only():
    # divert.rs:319-332
    data = []
    for hi in range(20):
        run = (hi * 7) % 40 + 1
        data += [(hi << 8) | ((i * 37 + hi) % 251) for i in range(run)]
    got = nv.finish_runs_u16(np.array(data, dtype=np.uint16), 8)
    assert got.tolist() == sorted(data)


def test_histogram_kats():
    hk = {k["gen"]: k for k in kats()["histogram"]}
    keys = np.array([0x1111222200004455 | ((i % 256) << 16) | ((i // 256) << 63) for i in range(300)],
                    dtype=np.uint64)
    assert nv.varying_bits(keys) == int(hk["u64_byte2_top"]["expect"], 16)
    assert nv.varying_bits(np.full(5, 7, dtype=np.uint32)) == 0
    assert nv.varying_bits((np.arange(300, dtype=np.uint32) * 0x9E37).astype(np.uint16)) is None
    keys = np.array([((i * 2654435761) % 4099) << 24 for i in range(10_000)], dtype=np.uint64)
    v = nv.varying_bits(keys)
    assert [p for p in range(8) if (v >> (8 * p)) & 0xFF] == hk["u64_live_3_4"]["live"]


def test_splitmix64_published_outputs():
    s = kats()["splitmix64"]
    want = np.array(s["outputs"], dtype=np.uint64)
    assert np.array_equal(cf.splitmix64(s["seed"], 5), want)
    assert np.array_equal(nv.splitmix64(s["seed"], 5), want)
    c1 = kats()["c1_first16"]
    assert cf.u32_keys(c1["seed"], 16).tolist() == c1["u32"]


# ------------------------------------------- transforms vs the definitions
def _xorshift_patterns():
    # key.rs:253-282: edge bit patterns plus 2000 xorshift64 (13, 7, 17) outputs.
    s = 0x9E3779B97F4A7C15
    out = []
    for _ in range(2000):
        s ^= (s << 13) & (2**64 - 1)
        s ^= s >> 7
        s ^= (s << 17) & (2**64 - 1)
        out.append(s)
    return out


def _total_order_key(bits: int, w: int):
    """IEEE 754-2008 5.10 totalOrder as a Python sort key: by sign (negative
    first), then by the magnitude bits, ascending for positive and descending
    for negative values.  Covers -0 < +0, infinities and NaN payloads."""
    sign = bits >> (8 * w - 1)
    mag = bits & ((1 << (8 * w - 1)) - 1)
    return (0, -mag) if sign else (1, mag)


@pytest.mark.parametrize("name,w", [("f32", 4), ("f64", 8)])
def test_float_transform_is_total_order(name, w):
    edges = [0, 1, 0x7FFFFFFF, 0x80000000, 0x80000001, 0xFFFFFFFF]
    pats = edges + ([1 << 63, 2**64 - 1] if w == 8 else [])
    pats += [p & ((1 << (8 * w)) - 1) for p in _xorshift_patterns()]
    u = cf.UNSIGNED_OF[w]
    bits = np.array(pats, dtype=u)
    vals = bits.view(cf.NP_DTYPES[name])
    keys = cf.to_radix(vals)
    assert np.array_equal(bits_of(cf.from_radix(keys, name)), bits)  # round trip
    by_key = sorted(range(len(pats)), key=lambda i: (int(keys[i]), i))
    by_def = sorted(range(len(pats)), key=lambda i: (_total_order_key(pats[i], w), i))
    assert by_key == by_def
    # the C++ restatement sorts the same way
    assert np.array_equal(bits_of(nv.sort(vals)), bits[by_def])


@pytest.mark.parametrize("name", ["i8", "i16", "i32", "i64", "u8", "u16", "u32", "u64"])
def test_integer_transform_matches_integer_order(name):
    x = random_of(name, 3000, seed=5)
    assert np.array_equal(np.sort(x), cf.sort(x))
    assert np.array_equal(cf.from_radix(cf.to_radix(x), name), x)


# ------------------------------------ restatement vs brute force and closed form
def test_exhaustive_permutations_u8_and_i8():
    # tests/correctness.rs:50-72
    for n in range(0, 9):
        base = list(range(n))
        for perm in itertools.permutations(base):
            got = nv.sort(np.array(perm, dtype=np.uint8)).tolist()
            assert got == base
    for n in range(0, 7):
        base = list(range(-3, -3 + n))
        for perm in itertools.permutations(base):
            assert nv.sort(np.array(perm, dtype=np.int8)).tolist() == base


SIZES = [0, 1, 2, 3, 15, 16, 17, 47, 48, 128, 511, 512, 513, 1024, 4096, 70_000]


@pytest.mark.parametrize("name", cf.DTYPES)
def test_restatement_matches_closed_form(name):
    for n in SIZES:
        x = random_of(name, n, seed=n + 11)
        assert np.array_equal(bits_of(nv.sort(x)), bits_of(cf.sort(x))), (name, n)
        assert np.array_equal(bits_of(nv.sort_with_buffer(x)), bits_of(cf.sort(x))), (name, n)


def test_restatement_structured_and_adversarial():
    for n in (70_000, 200_000):
        for shape, x in structured_u64(n, seed=77).items():
            want = np.sort(x)
            assert np.array_equal(nv.sort(x), want), shape
            assert np.array_equal(nv.sort_with_buffer(x), want), shape
    for n in (8, 16, 17, 100, 512, 513, 1000, 200_000):
        for shape, x in adversarial_u32(n).items():
            assert np.array_equal(nv.sort(x), np.sort(x)), (shape, n)
    f = float_specials_f32(50_000, seed=3)
    assert np.array_equal(bits_of(nv.sort(f)), bits_of(cf.sort(f)))


@pytest.mark.parametrize("name", ["u8", "u32", "u64", "i32", "f32", "f64", "i16"])
def test_restated_sort_parallel(name):
    # tests/correctness.rs:639-781 at N = 300,000
    x = random_of(name, 300_000, seed=7)
    assert np.array_equal(bits_of(nv.sort_parallel(x, 4)), bits_of(cf.sort(x)))


def test_restated_sort_parallel_structured():
    for shape, x in structured_u64(300_000, seed=23).items():
        assert np.array_equal(nv.sort_parallel(x, 4), np.sort(x)), shape
    few = (np.arange(300_000) % 7).astype(np.uint32)
    assert np.array_equal(nv.sort_parallel(few, 4), np.sort(few))
    rev = np.arange(300_000, dtype=np.uint64)[::-1] // 5
    assert np.array_equal(nv.sort_parallel(np.ascontiguousarray(rev), 4), np.sort(rev))


@pytest.mark.parametrize("name", ["u32", "u64", "f64", "f32", "i16", "i8"])
def test_restated_order_matches_stable_argsort(name):
    g = np.random.default_rng(9)
    for n in (0, 1, 2, 63, 64, 127, 128, 129, 1000, 5000, 70_000):
        for mod in (7, None):
            x = random_of(name, n, seed=n)
            if mod and n:
                x = x[g.integers(0, min(n, mod), size=n)]  # few distinct keys: stability shows
            assert np.array_equal(nv.order(x), cf.order(x)), (name, n, mod)
    # ascending and non-increasing inputs take the reference's shortcuts (by_key.rs:203-226)
    asc = np.arange(1000, dtype=np.uint32) // 3
    assert np.array_equal(nv.order(asc), np.arange(1000, dtype=np.uint32))
    desc = np.ascontiguousarray(asc[::-1])
    assert np.array_equal(nv.order(desc), cf.order(desc))


def test_segmented_oracle_matches_restated_order_per_segment():
    """The batched oracle equals the restated by-key order (by_key.rs:175-310)
    applied segment by segment, values moved by apply_order."""
    g = np.random.default_rng(21)
    for name in ("u32", "f32", "i64", "u8"):
        lens = [0, 1, 2, 5, 16, 17, 100, 129, 1000, 3000]
        g.shuffle(lens)
        off = np.concatenate([[3], 3 + np.cumsum(lens)]).astype(np.uint64)
        n = int(off[-1]) + 4
        keys = random_of(name, n, seed=len(lens) + n)
        keys[10:400] = keys[g.integers(10, 400, size=390)]  # ties: stability shows
        vals = np.arange(n, dtype=np.uint32)
        got_k, got_v = cf.segmented_sort(keys, off, vals)
        want_k, want_v = keys.copy(), vals.copy()
        for lo, hi in zip(off[:-1].astype(np.int64), off[1:].astype(np.int64)):
            o = nv.order(np.ascontiguousarray(keys[lo:hi]))
            want_k[lo:hi] = nv.apply_order(o, keys[lo:hi])
            want_v[lo:hi] = nv.apply_order(o, vals[lo:hi])
        assert np.array_equal(bits_of(got_k), bits_of(want_k)), name
        assert np.array_equal(got_v, want_v), name


def test_segment_lengths_are_log_uniform_in_range():
    lens = cf.segment_lengths(7, 100_000)
    assert lens.min() >= 2 and lens.max() <= 16384
    # log-uniform: each of the 13 octaves [2^k, 2^(k+1)) holds about 1/13 of the segments
    octave = np.floor(np.log2(lens)).astype(int)
    share = np.bincount(octave, minlength=15)[1:14] / lens.size
    assert np.all(np.abs(share - 1 / 13) < 0.01), share
    assert abs(lens.mean() - (16385 - 2) / np.log(16385 / 2)) < 40


def test_pwl_gradient_oracle_matches_finite_differences():
    """The learned model's float64 gradient (oracle/pwl.py) against central
    differences of its loss, on parameters of both signs and on points that
    hit the first and last intervals."""
    from oracle import pwl

    g = np.random.default_rng(5)
    theta = g.normal(0.0, 2.0, pwl.K)
    u = np.sort(np.concatenate([g.uniform(0, pwl.K, 3000), [0.0, 0.25, pwl.K - 0.5, pwl.K - 1e-3]]))
    t = np.sort(g.uniform(0, 1, u.size))
    _, grad = pwl.loss_grad(theta, u, t)
    for j in (0, 1, 17, 511, 1000, pwl.K - 1):
        e = 1e-5
        tp, tm = theta.copy(), theta.copy()
        tp[j] += e
        tm[j] -= e
        fd = (pwl.loss(tp, u, t) - pwl.loss(tm, u, t)) / (2 * e)
        assert abs(fd - grad[j]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-11, (j, fd, grad[j])
    # knots: monotone, from 0 to 1
    y = pwl.knots(theta)
    assert y[0] == 0.0 and abs(y[-1] - 1.0) < 1e-12 and np.all(np.diff(y) >= 0)


def test_lognormal_mixture_generator():
    x = cf.lognormal_mixture_keys(8, 200_000)
    assert x.dtype == np.uint64
    assert np.array_equal(x, cf.lognormal_mixture_keys(8, 200_000))
    # bulk near 2^47 (x ~ 1 -> e^-12 * 2^64), a 5 % hot set with heavy repeats
    med = np.median(np.log2(x.astype(np.float64)))
    assert 45.5 < med < 47.5, med
    _, counts = np.unique(x, return_counts=True)
    assert counts.max() > 0.005 * x.size
    f = cf.lognormal_mixture_keys(8, 1000, as_f64=True)
    assert f.dtype == np.float64 and np.all(f > 0)


def test_bench_generators_match_the_oracle_generators():
    """bench.py builds its workloads without the oracle; they must be the
    same inputs the parity tests use."""
    import bench

    off, _, _ = cf.segmented_workload(7, 5000)
    assert np.array_equal(bench.segment_offsets(7, 5000), off.astype(np.int64))
    assert np.array_equal(bench.np_splitmix64(9, 1000), cf.splitmix64(9, 1000))
    rec = bench.score_records(20_000, 5)
    assert np.array_equal(rec["id"], np.arange(20_000, dtype=np.uint32))
    scores = rec["score"].view(np.float32)
    nan = np.isnan(scores)
    assert 0.001 < nan.mean() < 0.01  # two of the five specials, 1 % of the records
    assert np.isinf(scores).any() and (rec["score"] == 0x80000000).any()


def test_restated_segmented_sort_matches_the_closed_form():
    """tso_segmented_sort_pairs (the CPU baseline of the segmented sort) gives
    the closed form's per-segment stable order, with empty, one-key and long
    segments and heavy ties, for several key types and thread counts."""
    import numpy as np

    from oracle import closed_form as cf
    from oracle import native as nv

    g = np.random.default_rng(11)
    for name, threads in (("u32", 1), ("f32", 3), ("i64", 8), ("u8", 2)):
        lens = g.integers(0, 600, size=300)
        lens[:5] = [0, 1, 2, 0, 5000]
        off = np.zeros(lens.size + 1, dtype=np.uint64)
        off[1:] = np.cumsum(lens)
        n = int(off[-1])
        keys = cf.from_radix(g.integers(0, 40, size=n).astype(cf.UNSIGNED_OF[np.dtype(cf.NP_DTYPES[name]).itemsize]),
                             name) if name != "f32" else g.standard_normal(n).astype(np.float32).round(1)
        vals = np.arange(n, dtype=np.uint32)
        k, v = keys.copy(), vals.copy()
        nv.segmented_sort_pairs_inplace(k, off, v, threads)
        wk, wv = cf.segmented_sort(keys, off, vals)
        assert np.array_equal(k.view(cf.UNSIGNED_OF[k.dtype.itemsize]), wk.view(cf.UNSIGNED_OF[k.dtype.itemsize]))
        assert np.array_equal(v, wv)
