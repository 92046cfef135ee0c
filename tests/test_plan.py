"""Workspace planning of the device radix sort, checked on the host.

The look-back status a context reserves for n keys must hold the status of
every radix sort of at most n keys of that width, of every kind (keys only,
keys with order values, learned bucket digits): sorts of at most 2^21 keys
run smaller tiles, so a shorter sort can need more rows than a longer one.
The figures come from libtsort.so's own tile geometry (ts_plan_status_bytes,
host arithmetic only; no GPU call is made).
"""
from __future__ import annotations

import os

import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "turbosort_amd", "libtsort.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libtsort.so not built")

SMALL_N = 1 << 21
CODES = {1: 0, 2: 1, 4: 2, 8: 3}  # a dtype of each width


def sizes():
    out = {4097, 5000, 65536, SMALL_N - 1, SMALL_N, SMALL_N + 1, 3_000_000, 1 << 24, (1 << 28) + 3,
           (1 << 30) - 1, 1 << 30, (1 << 30) + 17, 10**9, 1 << 33}
    return sorted(out)


def test_reserved_status_covers_every_shorter_sort_of_every_kind():
    from turbosort_amd import _native

    L = _native.load()
    for width, code in CODES.items():
        kinds = (0, 1, 2) if width >= 4 else (0, 1)
        for n in sizes():
            reserved = L.ts_plan_status_bytes(code, n, -1)
            assert reserved > 0
            for m in sizes() + [n // 2, n - 1]:
                if m > n or m <= 4096:
                    continue
                for kind in kinds:
                    used = L.ts_plan_status_bytes(code, m, kind)
                    assert 0 < used <= reserved, (width, n, m, kind, used, reserved)


def test_reserved_status_grows_with_n():
    from turbosort_amd import _native

    L = _native.load()
    for code in CODES.values():
        prev = 0
        for n in sizes():
            r = L.ts_plan_status_bytes(code, n, -1)
            assert r >= prev
            prev = r


def test_split_part_after_small_tile_switch():
    """A sort of 2^21 + 1 u32 keys whose part of 2^21 keys runs the small
    tiles: the part's status (256-key rows of 8192-key tiles) must fit the
    reservation for the whole sort."""
    from turbosort_amd import _native

    L = _native.load()
    part = L.ts_plan_status_bytes(2, SMALL_N, 0)
    assert part == 2 * (SMALL_N // 8192 + 8) * 256 * 4
    assert L.ts_plan_status_bytes(2, SMALL_N + 1, -1) >= part
    # 64-bit look-back words from 2^30 keys on.
    assert L.ts_plan_status_bytes(2, 1 << 30, 0) == 2 * ((1 << 30) // 32768 + 8) * 256 * 8


def test_record_kinds_use_their_own_tiles():
    """Carried 8- / 16-byte records, with or without the order, run tiles of
    their own (512 lanes x 24 / 14 keys with the order, x 36 / 16 without):
    each kind's status is two rows of 256 look-back words per tile, plus one
    partial tile per segment, as wide as n's words."""
    from turbosort_amd import _native

    L = _native.load()
    tiles = {3: 512 * 24, 4: 512 * 14, 5: 512 * 36, 6: 512 * 16}
    for code in CODES.values():
        for n in sizes():
            word = 8 if n >= 1 << 30 else 4
            for kind, tile in tiles.items():
                want = 2 * ((n + tile - 1) // tile + 8) * 256 * word
                assert L.ts_plan_status_bytes(code, n, kind) == want, (code, n, kind)
