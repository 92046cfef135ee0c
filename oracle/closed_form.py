"""Closed-form oracle for turbosort's sort path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product path never does.

Why a closed form pins the reference's output:

* Each ``SortableKey`` transform is a bijection onto an unsigned key of the same
  width whose unsigned order is the type's total order (key.rs:115-198), so
  equal keys are equal bit patterns (presorted.rs:10-14) and ``turbosort::sort``
  has exactly one possible output: the values in ascending radix-key order.
* ``sort_by_key_with_buffer`` is stable (by_key.rs:1-13), so ``order()`` is the
  unique stable argsort by radix key (by_key.rs:92-110); ``apply_order`` and the
  element move are ``out[i] = in[order[i]]`` (by_key.rs:136-144, 315-334).

The transforms below are written from the mathematics of IEEE-754 total order
and two's complement, not from the reference's expressions; tests check them
against ``total_cmp`` semantics on edge and random bit patterns and against the
C++ restatement in restate.cpp.
"""
from __future__ import annotations

import numpy as np

# dtype numbering shared with include/tsort.h (ts_dtype)
DTYPES = ["u8", "u16", "u32", "u64", "i8", "i16", "i32", "i64", "f32", "f64"]
NP_DTYPES = {
    "u8": np.uint8, "u16": np.uint16, "u32": np.uint32, "u64": np.uint64,
    "i8": np.int8, "i16": np.int16, "i32": np.int32, "i64": np.int64,
    "f32": np.float32, "f64": np.float64,
}
UNSIGNED_OF = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def dtype_name(arr_or_dtype) -> str:
    dt = np.dtype(getattr(arr_or_dtype, "dtype", arr_or_dtype))
    for name, npdt in NP_DTYPES.items():
        if np.dtype(npdt) == dt:
            return name
    raise TypeError(f"not a sortable key type: {dt}")


def to_radix(values: np.ndarray) -> np.ndarray:
    """Map values to unsigned keys whose unsigned order is the total order.

    unsigned: identity.  signed: adding 2^(w-1) modulo 2^w moves MIN to 0 and
    MAX to 2^w-1, which is a flip of the top bit.  floats (IEEE-754 total
    order): non-negative sign -> set the top bit (places them above every
    negative); negative sign -> complement every bit (reverses their
    magnitude order and clears the top bit).
    """
    name = dtype_name(values)
    w = np.dtype(NP_DTYPES[name]).itemsize
    u = UNSIGNED_OF[w]
    bits = np.ascontiguousarray(values).view(u)
    top = u(1 << (8 * w - 1))
    if name[0] == "u":
        return bits.copy()
    if name[0] == "i":
        return bits ^ top
    neg = (bits & top) != 0
    return np.where(neg, ~bits, bits | top).astype(u)


def from_radix(keys: np.ndarray, name: str) -> np.ndarray:
    """Inverse of :func:`to_radix` for the dtype called ``name``."""
    w = np.dtype(NP_DTYPES[name]).itemsize
    u = UNSIGNED_OF[w]
    keys = np.ascontiguousarray(keys, dtype=u)
    top = u(1 << (8 * w - 1))
    if name[0] == "u":
        bits = keys.copy()
    elif name[0] == "i":
        bits = keys ^ top
    else:
        from_pos = (keys & top) != 0  # came from a non-negative-sign float
        bits = np.where(from_pos, keys & ~top, ~keys).astype(u)
    return bits.view(NP_DTYPES[name])


def sort(values: np.ndarray) -> np.ndarray:
    """Expected output of ``turbosort::sort`` on ``values`` (a new array)."""
    name = dtype_name(values)
    return from_radix(np.sort(to_radix(values), kind="stable"), name)


def order(keys: np.ndarray) -> np.ndarray:
    """Expected ``SortByKeyBuffer::order()`` after sorting by ``keys``."""
    if keys.size > 0xFFFFFFFF:
        raise ValueError("order() is empty above u32::MAX elements")
    return np.argsort(to_radix(keys), kind="stable").astype(np.uint32)


def apply_order(order_arr: np.ndarray, values: np.ndarray) -> np.ndarray:
    """``out[i] = values[order[i]]``."""
    if order_arr.shape[0] != values.shape[0]:
        raise ValueError("apply_order: the slice's length differs from the last sorted slice's")
    return values[order_arr.astype(np.int64)]


def segmented_sort(keys: np.ndarray, offsets: np.ndarray, values=None):
    """Expected batched segmented sort: (keys, values) as new arrays.

    Every segment ``[offsets[s], offsets[s+1])`` is put in the stable order
    by radix key -- per segment ``sort_by_key_with_buffer`` keyed by the key
    itself with ``apply_order`` on the values (by_key.rs:136-235); elements
    outside ``[offsets[0], offsets[-1])`` keep their place.
    """
    keys_out = keys.copy()
    vals_out = None if values is None else values.copy()
    offs = np.asarray(offsets, dtype=np.int64)
    if offs.size < 2:
        return keys_out, vals_out
    lo, hi = int(offs[0]), int(offs[-1])
    seg = np.repeat(np.arange(offs.size - 1, dtype=np.int64), np.diff(offs))
    perm = np.lexsort((to_radix(keys[lo:hi]), seg))  # stable: segment, then key, then position
    keys_out[lo:hi] = keys[lo:hi][perm]
    if values is not None:
        vals_out[lo:hi] = values[lo:hi][perm]
    return keys_out, vals_out


# --------------------------------------------------------------- generators
GAMMA = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, n: int, start: int = 0) -> np.ndarray:
    """SplitMix64 outputs ``start .. start+n`` of the stream with state ``seed``.

    out[i] = mix(seed + (i + 1) * gamma) with Stafford's variant-13 mixer, as
    in Vigna's reference splitmix64.c.
    """
    with np.errstate(over="ignore"):
        idx = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def u32_keys(seed: int, n: int) -> np.ndarray:
    """Uniform u32 keys: the high 32 bits of each SplitMix64 output."""
    return (splitmix64(seed, n) >> np.uint64(32)).astype(np.uint32)


def u64_keys(seed: int, n: int) -> np.ndarray:
    return splitmix64(seed, n)


def u64_zero_extended_u32(seed: int, n: int) -> np.ndarray:
    return (splitmix64(seed, n) >> np.uint64(32)).astype(np.uint64)


def unit_f64(bits: np.ndarray) -> np.ndarray:
    """u64 -> uniform double in [0, 1) from the top 53 bits."""
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def segment_lengths(seed: int, n_segments: int, lo: int = 2, hi: int = 16384) -> np.ndarray:
    """Log-uniform segment lengths: clamp(floor(exp(U(ln lo, ln(hi + 1)))), lo, hi)
    with U from the SplitMix64 stream with state ``seed``."""
    u = unit_f64(splitmix64(seed, n_segments))
    a, b = np.log(float(lo)), np.log(float(hi + 1))
    return np.clip(np.floor(np.exp(a + u * (b - a))), lo, hi).astype(np.int64)


# The key stream of a segmented workload: state seed + 2^32 (so that it differs
# from the length stream, state seed).
SEGMENT_KEY_STREAM = 1 << 32


def segmented_workload(seed: int, n_segments: int):
    """(offsets u64, keys u32, values u32 = global index) of the batched
    segmented workload: log-uniform lengths in [2, 16384], uniform u32 keys."""
    lens = segment_lengths(seed, n_segments)
    offsets = np.zeros(n_segments + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum(lens).astype(np.uint64)
    n = int(offsets[-1])
    keys = u32_keys(seed + SEGMENT_KEY_STREAM, n)
    return offsets, keys, np.arange(n, dtype=np.uint32)


# --------------------------------------------------------------- skewed keys
def lognormal_mixture_keys(seed: int, n: int, as_f64: bool = False, hot_share: float = 0.05,
                           n_hot: int = 10_000) -> np.ndarray:
    """Skewed keys for the learned partitioner.

    x from a mixture of lognormals (mu 0, sigma 0.5 / 1 / 2 with weights
    0.5 / 0.3 / 0.2); u64 key = floor(min(x / e^12, 1 - 2^-53) * 2^64), or x
    itself as f64.  A share ``hot_share`` of the positions is replaced by
    Zipf(s = 1.1) draws over ``n_hot`` hot values from the same mixture.
    Streams: SplitMix64 states seed + k * 2^32 (k = 0 .. 7).
    """
    st = lambda k: seed + k * (1 << 32)  # noqa: E731

    def mixture(n_, k0):
        u1 = unit_f64(splitmix64(st(k0), n_))
        u2 = unit_f64(splitmix64(st(k0 + 1), n_))
        pick = unit_f64(splitmix64(st(k0 + 2), n_))
        z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
        sigma = np.where(pick < 0.5, 0.5, np.where(pick < 0.8, 1.0, 2.0))
        return np.exp(sigma * z)

    x = mixture(n, 0)
    hot = mixture(n_hot, 3)
    w = np.arange(1, n_hot + 1, dtype=np.float64) ** -1.1
    cdf = np.cumsum(w) / w.sum()
    rank = np.minimum(np.searchsorted(cdf, unit_f64(splitmix64(st(6), n)), side="right"), n_hot - 1)
    repl = unit_f64(splitmix64(st(7), n)) < hot_share
    x = np.where(repl, hot[rank], x)
    if as_f64:
        return x
    v = np.minimum(x / np.exp(12.0), 1.0 - 2.0 ** -53)
    hi = np.floor(v * 4294967296.0)
    lo = np.floor((v * 4294967296.0 - hi) * 4294967296.0)
    return (hi.astype(np.uint64) << np.uint64(32)) | lo.astype(np.uint64)
