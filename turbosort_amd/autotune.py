"""Autotuner of the batched segmented sort's size classes.

Each size class of ``segmented_sort`` (segments of up to 16, 32, 64, 256, 1024,
2048, 4096, 8192 and 16384 keys) runs one of a few precompiled kernel shapes:
4-bit or byte digits, lanes per workgroup, keys per lane
(``ts_segmented_variants`` / ``ts_segmented_variant_name``).  For one key type,
with or without values, ``tune_segmented`` times every shape of every class
that has alternatives on a batch of segments whose lengths fill that class's
range, keeps the fastest, and returns the table; ``apply`` installs a table on a
context, ``save`` / ``load`` keep it as JSON.  Results never depend on the
choice: every shape sorts exactly.
"""
from __future__ import annotations

import json
from typing import Optional

import torch

from . import Context, _default_context, _dtype_code, _native, fill_splitmix64, segmented_sort

CLASS_CAPS = [16, 32, 64, 256, 1024, 2048, 4096, 8192, 16384]


def variants(dtype: torch.dtype, with_values: bool, size_class: int) -> list:
    """Names of the kernel shapes size class ``size_class`` can run (index = variant)."""
    L = _native.load()
    code = _dtype_code(dtype)
    n = L.ts_segmented_variants(code, int(with_values), size_class)
    return [L.ts_segmented_variant_name(code, int(with_values), size_class, v).decode() for v in range(n)]


def set_variant(ctx: Context, dtype: torch.dtype, with_values: bool, size_class: int, variant: int) -> None:
    ctx.check(ctx._lib.ts_segmented_set_variant(ctx.handle, _dtype_code(dtype), int(with_values), size_class, variant))


def get_variant(ctx: Context, dtype: torch.dtype, with_values: bool, size_class: int) -> int:
    """The variant ``ctx`` runs for size class ``size_class``."""
    v = ctx._lib.ts_segmented_get_variant(ctx.handle, _dtype_code(dtype), int(with_values), size_class)
    if v < 0:
        raise ValueError(f"no size class {size_class}")
    return v


def _random_keys(n: int, dtype: torch.dtype, seed: int, ctx: Context) -> torch.Tensor:
    width = torch.empty(0, dtype=dtype).element_size()
    bits = torch.empty(n, dtype=torch.uint64, device=ctx.device)
    fill_splitmix64(bits, seed, ctx=ctx)
    raw = bits.view(torch.uint8).view(n, 8)[:, :width].contiguous()
    return raw.view(dtype).reshape(n)


def tune_segmented(dtype: torch.dtype = torch.uint32, with_values: bool = True, *, keys_per_class: int = 1 << 24,
                   repeats: int = 3, seed: int = 11, ctx: Optional[Context] = None) -> dict:
    """Times every kernel shape of every size class with alternatives and
    installs the fastest on ``ctx``; returns the table."""
    ctx = ctx or _default_context(torch.device("cuda", torch.cuda.current_device()))
    dev = ctx.device
    gen = torch.Generator(device="cpu").manual_seed(seed)
    table = {"dtype": str(dtype).replace("torch.", ""), "with_values": bool(with_values), "choice": {}, "times_ms": {},
             "names": {}}
    for c in range(len(CLASS_CAPS)):
        names = variants(dtype, with_values, c)
        if len(names) <= 1:
            continue
        lo = (CLASS_CAPS[c - 1] + 1) if c > 0 else 2
        hi = CLASS_CAPS[c]
        count = max(1, keys_per_class // ((lo + hi) // 2))
        lens = torch.randint(lo, hi + 1, (count,), generator=gen, dtype=torch.int64)
        offsets = torch.zeros(count + 1, dtype=torch.int64)
        offsets[1:] = torch.cumsum(lens, 0)
        offsets = offsets.to(dev)
        n = int(offsets[-1])
        src = _random_keys(n, dtype, seed + c, ctx)
        vsrc = torch.arange(n, dtype=torch.int32, device=dev) if with_values else None
        keys = torch.empty_like(src)
        vals = torch.empty_like(vsrc) if with_values else None
        times = []
        for v in range(len(names)):
            set_variant(ctx, dtype, with_values, c, v)
            best = float("inf")
            for r in range(repeats + 1):
                keys.copy_(src)
                if with_values:
                    vals.copy_(vsrc)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                segmented_sort(keys, offsets, vals, max_segment_len=hi, ctx=ctx)
                e1.record()
                torch.cuda.synchronize(dev)
                if r:  # the first run warms up
                    best = min(best, e0.elapsed_time(e1))
            times.append(best)
        choice = min(range(len(names)), key=lambda v: times[v])
        set_variant(ctx, dtype, with_values, c, choice)
        table["choice"][c] = choice
        table["times_ms"][c] = times
        table["names"][c] = names
        del src, vsrc, keys, vals
    return table


def apply(table: dict, ctx: Context) -> None:
    """Installs a table from ``tune_segmented`` (or ``load``) on a context."""
    dtype = getattr(torch, table["dtype"])
    for c, v in table["choice"].items():
        set_variant(ctx, dtype, table["with_values"], int(c), int(v))


def save(table: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(table, f, indent=1)


def load(path: str) -> dict:
    with open(path) as f:
        return json.load(f)


__all__ = ["CLASS_CAPS", "variants", "set_variant", "get_variant", "tune_segmented", "apply", "save", "load"]
