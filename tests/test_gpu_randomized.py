"""Seeded randomized parity of the one-workgroup kernels against the oracle.

Each trial draws a key type, a key distribution, a segment layout spanning the
size classes (empty and one-key segments, lengths on and around every class
bound, now and then one longer than a workgroup holds) and, per size class, a
kernel shape the autotuner could install; the batched segmented sort must
equal the oracle's per-segment stable sort (by_key.rs:175-235 per segment)
and leave the elements outside every segment alone.  A second set draws plain
sorts of sizes on both sides of the small tiers.  Seeds are fixed, so every
run checks the same cases.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import turbosort_amd as ts
from gpu_util import TORCH_OF, bits, to_dev, to_host
from oracle import closed_form as cf
from shapes import float_specials_f32, random_of, rng

pytestmark = pytest.mark.gpu

CAPS = [16, 32, 64, 256, 1024, 2048, 4096, 8192, 16384]
DISTS = ["uniform", "few distinct", "narrow range", "ascending", "descending", "equal runs", "float specials"]


def keys_of(name: str, n: int, dist: str, g: np.random.Generator) -> np.ndarray:
    x = random_of(name, n, seed=int(g.integers(1 << 30)))
    u = bits(x).copy()
    if dist == "few distinct":
        u = u[g.integers(0, 7, n)] if n else u
    elif dist == "narrow range":  # the keys differ in their low bits only
        mask = u.dtype.type((1 << min(12, 4 * u.dtype.itemsize)) - 1)
        u = (u & mask) | (u[:1] & ~mask) if n else u
    elif dist == "ascending":
        return cf.sort(x)
    elif dist == "descending":
        return cf.sort(x)[::-1].copy()
    elif dist == "equal runs":
        u = np.repeat(u[: n // 40 + 1], 40)[:n]
    elif dist == "float specials" and x.dtype == np.float32:
        return float_specials_f32(n, seed=int(g.integers(1 << 30)))
    return u.view(x.dtype)


def lengths_of(g: np.random.Generator, m: int, tdt: torch.dtype, with_values: bool) -> list:
    cap = ts.segmented_lds_max_len(tdt, with_values)
    out = []
    for _ in range(m):
        r = g.random()
        if r < 0.08:
            out.append(int(g.integers(0, 2)))  # empty and one-key segments
        elif r < 0.25:
            b = int(g.choice([c for c in CAPS if c <= cap] + [cap]))
            out.append(max(2, b + int(g.integers(-1, 2))) if b + 1 <= cap else b)  # on a class bound
        elif r < 0.27:
            out.append(int(g.integers(cap + 1, cap + 3000)))  # longer than one workgroup holds
        else:
            out.append(int(np.exp(g.uniform(np.log(2), np.log(cap)))))
    return out


@pytest.mark.parametrize("trial", range(48))
def test_segmented_random_layouts_and_shapes(gpu, trial):
    from turbosort_amd import autotune

    g = rng(7000 + trial)
    name = cf.DTYPES[trial % len(cf.DTYPES)]
    npdt = np.dtype(cf.NP_DTYPES[name])
    with_values = trial % 3 != 2
    dist = DISTS[trial % len(DISTS)]
    tdt = TORCH_OF[npdt]
    lens = lengths_of(g, int(g.integers(20, 120)), tdt, with_values)
    lead, tail = int(g.integers(0, 5)), int(g.integers(0, 5))
    off = np.zeros(len(lens) + 1, dtype=np.uint64)
    off[0] = lead
    off[1:] = lead + np.cumsum(np.asarray(lens, dtype=np.int64))
    n = int(off[-1]) + tail
    keys = keys_of(name, n, dist, g)
    vals = g.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32) if with_values else None
    ctx = ts.Context(gpu)
    shapes = {}
    for c in range(len(CAPS)):
        k = len(autotune.variants(tdt, with_values, c))
        if k > 1:
            shapes[c] = int(g.integers(0, k))
            autotune.set_variant(ctx, tdt, with_values, c, shapes[c])
    want_k, want_v = cf.segmented_sort(keys, off, vals)
    tk, tv = to_dev(keys, gpu), None if vals is None else to_dev(vals, gpu)
    ts.segmented_sort(tk, to_dev(off, gpu), tv, ctx=ctx)  # unknown bound: long segments take the device sort
    torch.cuda.synchronize()
    what = (name, dist, with_values, shapes)
    assert np.array_equal(bits(to_host(tk, npdt)), bits(want_k)), what
    if vals is not None:
        assert np.array_equal(to_host(tv, np.uint32), want_v), what


@pytest.mark.parametrize("trial", range(24))
def test_plain_sort_random_small_sizes(gpu, trial):
    """Sizes drawn on both sides of the network tier (4096 keys) and of the
    one-workgroup tier (16384 keys of 4 bytes and narrower, 17408 of 8),
    random distributions, through sort and sort_with_buffer."""
    g = rng(9100 + trial)
    name = cf.DTYPES[trial % len(cf.DTYPES)]
    cap = 17408 if np.dtype(cf.NP_DTYPES[name]).itemsize == 8 else 16384
    for _ in range(6):
        n = int(g.choice([int(g.integers(2, 4097)), int(g.integers(4097, cap + 1)), int(g.integers(cap + 1, 3 * cap))]))
        x = keys_of(name, n, DISTS[int(g.integers(len(DISTS)))], g)
        t = to_dev(x, gpu)
        if g.random() < 0.5:
            ts.sort(t)
        else:
            ts.sort_with_buffer(t, torch.empty(n, dtype=t.dtype, device=gpu))
        torch.cuda.synchronize()
        assert np.array_equal(bits(to_host(t, x.dtype)), bits(cf.sort(x))), (name, n)
