"""Batched segmented sort on the GPU vs the oracle.

Each segment must equal a stable key-value sort of that segment: keys in
radix-key order and values moved with them, ties in input order (per segment:
sort_by_key_with_buffer keyed by the key plus apply_order, by_key.rs:136-235).
Lengths cover every size class on both sides of its bound, the device-radix
class above the LDS capacity, empty and one-key segments and elements outside
every segment.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import turbosort_amd as ts
from gpu_util import bits, to_dev, to_host
from oracle import closed_form as cf
from shapes import float_specials_f32, random_of, rng

pytestmark = pytest.mark.gpu

# Both sides of every class bound (16, 32, 64, 256, 1024, ..., 16384) and the
# device class beyond the LDS capacity.
EDGE_LENGTHS = [0, 1, 2, 3, 5, 15, 16, 17, 31, 32, 33, 63, 64, 65, 200, 255, 256, 257, 1000, 1024, 1025, 2047,
                2048, 2049, 4095, 4096, 4097, 8191, 8192, 8193, 11264, 11265, 12000, 16383, 16384, 16385, 17408, 17409,
                40000]


def offsets_of(lengths, lead: int = 0) -> np.ndarray:
    o = np.zeros(len(lengths) + 1, dtype=np.uint64)
    o[0] = lead
    o[1:] = lead + np.cumsum(np.asarray(lengths, dtype=np.int64)).astype(np.uint64)
    return o


def run(keys: np.ndarray, offsets: np.ndarray, values, dev, max_len: int = 0):
    tk = to_dev(keys, dev)
    to = to_dev(offsets, dev)
    tv = None if values is None else to_dev(values, dev)
    ts.segmented_sort(tk, to, tv, max_segment_len=max_len)
    torch.cuda.synchronize()
    return to_host(tk, keys.dtype), None if tv is None else to_host(tv, values.dtype)


def check(keys, offsets, values, dev, max_len=0):
    got_k, got_v = run(keys, offsets, values, dev, max_len)
    want_k, want_v = cf.segmented_sort(keys, offsets, values)
    assert np.array_equal(bits(got_k), bits(want_k))
    if values is not None:
        assert np.array_equal(got_v, want_v)


@pytest.mark.parametrize("dtype", cf.DTYPES)
@pytest.mark.parametrize("with_values", [True, False])
def test_every_class_every_dtype(gpu, dtype, with_values):
    lens = list(EDGE_LENGTHS)
    rng(3).shuffle(lens)
    off = offsets_of(lens, lead=7)
    n = int(off[-1]) + 5  # untouched elements before and after the segments
    keys = random_of(dtype, n, seed=len(dtype) * 7 + int(with_values))
    vals = np.arange(n, dtype=np.uint32) if with_values else None
    check(keys, off, vals, gpu)


@pytest.mark.parametrize("dtype", ["u32", "f32", "u64", "i16"])
def test_bounded_length_is_asynchronous_and_equal(gpu, dtype):
    cap = ts.segmented_lds_max_len(getattr(torch, {"u32": "uint32", "f32": "float32", "u64": "uint64",
                                                   "i16": "int16"}[dtype]), True)
    lens = [x for x in EDGE_LENGTHS if x <= cap]
    off = offsets_of(lens)
    n = int(off[-1])
    keys = random_of(dtype, n, seed=11)
    check(keys, off, np.arange(n, dtype=np.uint32), gpu, max_len=max(lens))


@pytest.mark.parametrize("shape", ["few unique", "constant", "sorted", "reversed", "one digit varies"])
def test_ties_keep_input_order(gpu, shape):
    lens = [2, 9, 30, 64, 100, 700, 3000, 9000, 16384, 20000]
    off = offsets_of(lens)
    n = int(off[-1])
    i = np.arange(n, dtype=np.uint64)
    keys = {
        "few unique": (i % 3).astype(np.uint32),
        "constant": np.full(n, 0xDEADBEEF, dtype=np.uint32),
        "sorted": i.astype(np.uint32),
        "reversed": (n - i).astype(np.uint32),
        "one digit varies": (0x12345600 | (i * 7 % 16) << np.uint64(4)).astype(np.uint32),
    }[shape]
    check(keys, off, np.arange(n, dtype=np.uint32), gpu)


def test_float_specials(gpu):
    lens = [3, 17, 64, 500, 5000, 16384, 17000]
    off = offsets_of(lens)
    n = int(off[-1])
    keys = float_specials_f32(n, seed=5)
    check(keys, off, np.arange(n, dtype=np.uint32), gpu)
    check(keys.astype(np.float64), off, np.arange(n, dtype=np.uint32), gpu)


@pytest.mark.parametrize("n_segments", [1, 100, 3000])
def test_log_uniform_workload(gpu, n_segments):
    off, keys, vals = cf.segmented_workload(7, n_segments)
    check(keys, off, vals, gpu, max_len=16384)


def test_invalid_offsets_raise(gpu):
    keys = to_dev(np.arange(100, dtype=np.uint32), gpu)
    bad = to_dev(np.array([0, 50, 40, 100], dtype=np.uint64), gpu)
    with pytest.raises(ValueError, match="non-decreasing"):
        ts.segmented_sort(keys, bad)
    past = to_dev(np.array([0, 50, 101], dtype=np.uint64), gpu)
    with pytest.raises(ValueError, match="non-decreasing"):
        ts.segmented_sort(keys, past)


def test_full_size_properties(gpu):
    """10^5 log-uniform segments (1.82e8 pairs) with values = global index.

    Output is correct iff every value stays in its segment, the values form a
    permutation, each key is the input key at its value's position, keys do
    not decrease inside a segment and equal keys keep increasing values.
    """
    off, keys, _ = cf.segmented_workload(7, 100_000)
    n = int(off[-1])
    k = to_dev(keys, gpu)
    orig = k.view(torch.int32).clone()
    v = torch.arange(n, dtype=torch.int32, device=gpu)
    to = to_dev(off, gpu)
    ts.segmented_sort(k, to, v, max_segment_len=16384)
    torch.cuda.synchronize()
    lens = torch.from_numpy(np.diff(off.astype(np.int64))).to(gpu)
    seg = torch.repeat_interleave(torch.arange(lens.numel(), device=gpu), lens)
    vl = v.long()

    def take(t, idx, chunk=1 << 24):  # torch's one-call gather has returned zeros past 2^30 output bytes
        return torch.cat([t[idx[c:c + chunk]] for c in range(0, idx.numel(), chunk)])

    assert torch.equal(take(seg, vl), seg)
    seen = torch.zeros(n, dtype=torch.bool, device=gpu)
    seen[vl] = True
    assert bool(seen.all())
    assert torch.equal(take(orig, vl), k.view(torch.int32))
    kk = k.view(torch.int32).long() & 0xFFFFFFFF
    same = seg[1:] == seg[:-1]
    assert not bool((same & (kk[1:] < kk[:-1])).any())
    assert not bool((same & (kk[1:] == kk[:-1]) & (vl[1:] <= vl[:-1])).any())


@pytest.mark.parametrize("dtype", ["u32", "u64", "f32", "i8"])
def test_long_keys_only_segments_split(gpu, dtype):
    """Keys-only segments beyond the LDS capacity: repeated values, few
    distinct values and random keys, sorted through three-way splits."""
    n_hot = 200_000
    g = rng(55)
    a = random_of(dtype, n_hot, seed=56)
    a[g.random(n_hot) < 0.6] = a[0]
    b = random_of(dtype, 5, seed=57)[g.integers(0, 5, size=150_000)]
    c = random_of(dtype, 90_000, seed=58)
    keys = np.concatenate([a, b, c, random_of(dtype, 300, seed=59)])
    off = offsets_of([n_hot, 150_000, 90_000, 300])
    check(keys, off, None, gpu)


@pytest.mark.parametrize("dtype", ["u64", "f32"])
def test_very_long_keys_only_segments(gpu, dtype):
    """Keys-only segments of over 64 workgroups' capacity (several split
    rounds, few distinct values in one of them) beside shorter ones."""
    cap = ts.segmented_lds_max_len(torch.uint64 if dtype == "u64" else torch.float32, False)
    big = 64 * cap + 1000
    g = rng(61)
    a = random_of(dtype, big, seed=62)
    b = random_of(dtype, 7, seed=63)[g.integers(0, 7, size=big + 5)]
    keys = np.concatenate([a, random_of(dtype, 3 * cap, seed=64), b, random_of(dtype, 100, seed=65)])
    off = offsets_of([big, 3 * cap, big + 5, 100], lead=0)
    check(keys, off, None, gpu)


@pytest.mark.parametrize("rounds", [0, 1, 2])
def test_split_round_limit_hands_the_rest_to_the_radix_sort(gpu, rounds):
    """Long keys-only segments left after the set number of split rounds take
    the device radix sort; the output does not depend on the limit."""
    cap = ts.segmented_lds_max_len(torch.uint32, False)
    g = rng(71)
    a = random_of("u32", 40 * cap, seed=72)
    b = random_of("u32", 9, seed=73)[g.integers(0, 9, size=5 * cap)]
    keys = np.concatenate([a, b, random_of("u32", 2 * cap + 3, seed=74), random_of("u32", 500, seed=75)])
    off = offsets_of([40 * cap, 5 * cap, 2 * cap + 3, 500])
    ctx = ts.Context(gpu)
    ctx.set_split_rounds(rounds)
    assert ctx.split_rounds == rounds
    tk, to = to_dev(keys, gpu), to_dev(off, gpu)
    ts.segmented_sort(tk, to, None, ctx=ctx)
    torch.cuda.synchronize()
    want, _ = cf.segmented_sort(keys, off, None)
    assert np.array_equal(to_host(tk, np.uint32), want)
    with pytest.raises(ValueError):
        ctx.set_split_rounds(65)


@pytest.mark.parametrize("with_values", [False, True])
def test_more_long_segments_than_one_readback(gpu, with_values):
    """More segments past the LDS capacity than the first bounds readback
    holds (8192): the rest arrive in a second copy; every segment sorted."""
    dt = torch.uint8
    cap = ts.segmented_lds_max_len(dt, with_values)
    count = 8192 + 37
    lens = [cap + 1 + (i % 3) for i in range(count)]
    off = offsets_of(lens, lead=3)
    n = int(off[-1]) + 2
    keys = random_of("u8", n, seed=91)
    vals = np.arange(n, dtype=np.uint32) if with_values else None
    got_k, got_v = run(keys, off, vals, gpu, 0)
    assert np.array_equal(got_k[:3], keys[:3]) and np.array_equal(got_k[-2:], keys[-2:])
    for i in range(count):
        lo, hi = int(off[i]), int(off[i + 1])
        seg = keys[lo:hi]
        if with_values:  # stable: equal keys keep their input order
            order = np.argsort(seg, kind="stable")
            assert np.array_equal(got_k[lo:hi], seg[order])
            assert np.array_equal(got_v[lo:hi], vals[lo:hi][order])
        else:
            assert np.array_equal(got_k[lo:hi], np.sort(seg))


def test_graph_capture_replays(gpu):
    """With a bound that fits one workgroup the sort makes no host call, so
    it can be captured in a graph; every replay sorts fresh input exactly
    (the input is restored eagerly between replays)."""
    off, keys, vals = cf.segmented_workload(7, 2000)
    tk = to_dev(keys, gpu)
    tv = to_dev(vals, gpu)
    to = to_dev(off, gpu)
    src_k, src_v = tk.clone(), tv.clone()
    want_k, want_v = cf.segmented_sort(keys, off, vals)
    ctx = ts.Context(gpu)
    ts.segmented_sort(tk, to, tv, max_segment_len=16384, ctx=ctx)  # grows the workspace before capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        tk.copy_(src_k)
        tv.copy_(src_v)
        with torch.cuda.graph(g, stream=s):
            ts.segmented_sort(tk, to, tv, max_segment_len=16384, ctx=ctx)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    for _ in range(3):
        tk.copy_(src_k)
        tv.copy_(src_v)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(to_host(tk, np.uint32), want_k)
        assert np.array_equal(to_host(tv, np.uint32), want_v)


@pytest.mark.parametrize("dtype,with_values", [("u32", True), ("u64", False), ("f32", True), ("i64", True),
                                               ("u64", True)])
def test_every_kernel_variant_is_exact(gpu, dtype, with_values):
    """Every kernel shape the autotuner may pick sorts exactly."""
    from turbosort_amd import autotune

    tdt = {"u32": torch.uint32, "u64": torch.uint64, "f32": torch.float32, "i64": torch.int64}[dtype]
    ctx = ts.Context(gpu)
    g = rng(91)
    for c, cap in enumerate(autotune.CLASS_CAPS):
        names = autotune.variants(tdt, with_values, c)
        if len(names) <= 1:
            continue
        lo = autotune.CLASS_CAPS[c - 1] + 1
        lens = g.integers(lo, cap + 1, size=7)
        off = offsets_of(list(lens))
        n = int(off[-1])
        keys = random_of(dtype, n, seed=c * 10 + len(dtype))
        keys[: n // 3] = keys[g.integers(0, max(1, n // 3), size=n // 3)]  # ties
        vals = np.arange(n, dtype=np.uint32) if with_values else None
        want_k, want_v = cf.segmented_sort(keys, off, vals)
        for v in range(len(names)):
            autotune.set_variant(ctx, tdt, with_values, c, v)
            tk, tv = to_dev(keys, gpu), None if vals is None else to_dev(vals, gpu)
            ts.segmented_sort(tk, to_dev(off, gpu), tv, max_segment_len=cap, ctx=ctx)
            torch.cuda.synchronize()
            assert np.array_equal(bits(to_host(tk, keys.dtype)), bits(want_k)), (c, names[v])
            if vals is not None:
                assert np.array_equal(to_host(tv, np.uint32), want_v), (c, names[v])


def test_autotuner_picks_a_variant_per_class(gpu):
    from turbosort_amd import autotune

    ctx = ts.Context(gpu)
    table = autotune.tune_segmented(torch.uint32, True, keys_per_class=1 << 18, repeats=1, ctx=ctx)
    assert set(table["choice"]) == {c for c in range(9) if len(autotune.variants(torch.uint32, True, c)) > 1}
    for c, v in table["choice"].items():
        assert table["times_ms"][c][v] == min(table["times_ms"][c])
        assert autotune.get_variant(ctx, torch.uint32, True, c) == v
    off, keys, vals = cf.segmented_workload(7, 500)
    tk, tv = to_dev(keys, gpu), to_dev(vals, gpu)
    ts.segmented_sort(tk, to_dev(off, gpu), tv, max_segment_len=16384, ctx=ctx)
    torch.cuda.synchronize()
    want_k, want_v = cf.segmented_sort(keys, off, vals)
    assert np.array_equal(to_host(tk, np.uint32), want_k) and np.array_equal(to_host(tv, np.uint32), want_v)


@pytest.mark.parametrize("dtype", ["u32", "u64", "f64", "i32"])
def test_msd_shapes_on_structured_segments(gpu, dtype):
    """The keys-only MSD shapes (plain and pipelined) on segments that take
    each of their paths: short runs (random keys), keys differing only in their
    low bits, every key equal, few distinct values and a tight cluster plus one
    outlier (runs beyond 64 keys: the LSD fallback)."""
    from turbosort_amd import autotune

    tdt = {"u32": torch.uint32, "u64": torch.uint64, "f64": torch.float64, "i32": torch.int32}[dtype]
    npdt = cf.NP_DTYPES[dtype]
    udt = cf.UNSIGNED_OF[np.dtype(npdt).itemsize]
    ctx = ts.Context(gpu)
    g = rng(123)
    for c in (6, 7, 8):
        names = autotune.variants(tdt, False, c)
        msd = [v for v, nm in enumerate(names) if "MSD" in nm]
        assert len(msd) == 2, names  # the plain and the pipelined form
        cap, lo = autotune.CLASS_CAPS[c], autotune.CLASS_CAPS[c - 1] + 1
        parts = []
        for kind in range(5):
            m = int(g.integers(lo, cap + 1))
            base = random_of(dtype, m, seed=c * 100 + kind)
            if kind == 1:
                base = np.sort(base)[:1].repeat(m)
                base = (base.view(udt) ^ g.integers(0, 1024, m).astype(udt)).view(npdt)
            elif kind == 2:
                base = base[:1].repeat(m)
            elif kind == 3:
                base = base[g.integers(0, 5, m)]
            elif kind == 4:
                u = base.view(udt).copy()
                u[:] = u[0] ^ g.integers(0, 1 << 16, m).astype(udt)
                u[0] = ~u[1]
                base = u.view(npdt)
            parts.append(base.astype(npdt))
        keys = np.concatenate(parts)
        off = offsets_of([len(p) for p in parts])
        want, _ = cf.segmented_sort(keys, off, None)
        for v in msd:
            autotune.set_variant(ctx, tdt, False, c, v)
            tk = to_dev(keys, gpu)
            ts.segmented_sort(tk, to_dev(off, gpu), None, max_segment_len=cap, ctx=ctx)
            torch.cuda.synchronize()
            assert np.array_equal(bits(to_host(tk, keys.dtype)), bits(want)), (dtype, c, names[v])


@pytest.mark.parametrize("dtype", ["u32", "i32", "f32"])
def test_pair_msd_shapes_on_structured_segments(gpu, dtype):
    """The key-value MSD shapes (one MSD pass, ranks inside the bins) on
    segments that take each of their paths: random keys (short bins), keys
    differing only in their low bits, every key equal (nothing moves), few
    distinct values and a tight cluster plus one outlier (bins beyond 64 keys:
    the LSD fallback), and float specials; ties must keep input order."""
    from turbosort_amd import autotune

    tdt = {"u32": torch.uint32, "i32": torch.int32, "f32": torch.float32}[dtype]
    npdt = cf.NP_DTYPES[dtype]
    udt = cf.UNSIGNED_OF[np.dtype(npdt).itemsize]
    ctx = ts.Context(gpu)
    g = rng(321)
    seen = 0
    for c in range(len(autotune.CLASS_CAPS)):
        names = autotune.variants(tdt, True, c)
        shapes = [v for v, nm in enumerate(names) if "ranks in bins" in nm]
        if not shapes:
            continue
        cap, lo = autotune.CLASS_CAPS[c], autotune.CLASS_CAPS[c - 1] + 1
        parts = []
        for kind in range(6):
            m = int(g.integers(lo, cap + 1)) if kind else cap
            base = random_of(dtype, m, seed=c * 100 + kind + 7)
            if kind == 1:
                base = np.sort(base)[:1].repeat(m)
                base = (base.view(udt) ^ g.integers(0, 1024, m).astype(udt)).view(npdt)
            elif kind == 2:
                base = base[:1].repeat(m)
            elif kind == 3:
                base = base[g.integers(0, 5, m)]
            elif kind == 4:
                u = base.view(udt).copy()
                u[:] = u[0] ^ g.integers(0, 1 << 16, m).astype(udt)
                u[0] = ~u[1]
                base = u.view(npdt)
            elif kind == 5 and dtype == "f32":
                base = float_specials_f32(m, seed=c + 50)
            parts.append(base.astype(npdt))
        keys = np.concatenate(parts)
        off = offsets_of([len(p) for p in parts])
        vals = g.integers(0, 1 << 32, keys.size, dtype=np.uint64).astype(np.uint32)
        want_k, want_v = cf.segmented_sort(keys, off, vals)
        for v in shapes:
            autotune.set_variant(ctx, tdt, True, c, v)
            tk, tv = to_dev(keys, gpu), to_dev(vals, gpu)
            ts.segmented_sort(tk, to_dev(off, gpu), tv, max_segment_len=cap, ctx=ctx)
            torch.cuda.synchronize()
            assert np.array_equal(bits(to_host(tk, keys.dtype)), bits(want_k)), (dtype, c, names[v])
            assert np.array_equal(to_host(tv, np.uint32), want_v), (dtype, c, names[v])
            seen += 1
    assert seen >= 5
