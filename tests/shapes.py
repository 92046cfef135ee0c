"""Seeded input shapes for parity tests.

The reference's tests draw from rand::StdRng (ChaCha), which is not available
here; these generators reproduce the *shapes* its tests use (file:line given
per shape) from numpy's seeded PCG64.  Expected outputs always come from the
oracle, never from the generator.
"""
from __future__ import annotations

import numpy as np

from oracle import closed_form as cf

GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def rng(seed: int) -> np.random.Generator:
    return np.random.default_rng(seed)


def random_of(dtype: str, n: int, seed: int) -> np.ndarray:
    """Uniform random bit patterns of a dtype (proptest any::<T>(), correctness.rs:404-438)."""
    npdt = cf.NP_DTYPES[dtype]
    w = np.dtype(npdt).itemsize
    bits = rng(seed).integers(0, 2**63, size=n, dtype=np.uint64, endpoint=True)
    bits = bits ^ (rng(seed + 1).integers(0, 2, size=n, dtype=np.uint64) << np.uint64(63))
    return bits.astype(cf.UNSIGNED_OF[w]).view(npdt)


def structured_u64(n: int, seed: int) -> dict:
    """The seven structured u64 shapes of tests/correctness.rs:129-190."""
    g = rng(seed)
    r = lambda: g.integers(0, 2**64 - 1, size=n, dtype=np.uint64, endpoint=True)  # noqa: E731
    with np.errstate(over="ignore"):
        b = r() & np.uint64(0xFF)
        return {
            "random": r(),
            "low 40 bits": r() >> np.uint64(24),
            "constant low bytes": g.integers(0, 5000, size=n, dtype=np.uint64) << np.uint64(32),
            "top bit split": ((r() & np.uint64(1)) << np.uint64(63)) | (r() >> np.uint64(8)),
            "correlated top bytes": (b << np.uint64(56)) | (b << np.uint64(48)) | (b << np.uint64(40)) |
                                    (r() >> np.uint64(24)),
            "1000 distinct": g.integers(0, 1000, size=n, dtype=np.uint64) * GOLDEN,
            "heavy tail": np.power(1.0 / (g.random(n) + 1e-9), 1.5).astype(np.uint64),
        }


def adversarial_u32(n: int) -> dict:
    """tests/correctness.rs:577-635, 784-792."""
    i = np.arange(n, dtype=np.uint64)
    half = n // 2
    return {
        "presorted": i.astype(np.uint32),
        "reverse": i[::-1].astype(np.uint32).copy(),
        "pipe organ": np.concatenate([np.arange(half), np.arange(half)[::-1]]).astype(np.uint32),
        "sawtooth 8": (i % 8).astype(np.uint32),
        "sawtooth 64": (i % 64).astype(np.uint32),
        "few unique": (i % 4).astype(np.uint32),
        "m3 killer": np.where(i % 2 == 0, i, n - i).astype(np.uint32),
    }


def float_specials_f32(n: int, seed: int) -> np.ndarray:
    """Scores with NaN of both signs, +-inf, +-0.0 and duplicates."""
    g = rng(seed)
    x = g.standard_normal(n).astype(np.float32)
    bits = x.view(np.uint32).copy()
    specials = np.array([0x7FC00000, 0xFFC00000, 0x7F800000, 0xFF800000, 0x80000000, 0x00000000, 0x7F800001,
                         0xFFFFFFFF], dtype=np.uint32)
    pick = g.random(n) < 0.05
    bits[pick] = specials[g.integers(0, len(specials), size=int(pick.sum()))]
    dup = g.random(n) < 0.2
    bits[dup] = bits[g.integers(0, n, size=int(dup.sum()))]
    return bits.view(np.float32)
