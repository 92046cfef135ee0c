"""The C-ABI library loads and exports every symbol include/*.h declares.

No compute call is made: this runs on a machine without a GPU.
"""
from __future__ import annotations

import glob
import os
import re
import subprocess

import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "turbosort_amd", "libtsort.so")


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(ts_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_api():
    names = declared_symbols()
    for must in ("ts_sort", "ts_sort_with_buffer", "ts_sort_by_key", "ts_apply_order", "ts_ctx_create"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libtsort.so not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = declared_symbols() - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"


@pytest.mark.skipif(not os.path.exists(LIB), reason="libtsort.so not built")
def test_ctypes_binding_covers_the_header():
    from turbosort_amd import _native

    L = _native.load()
    assert set(_native.EXPORTS) == declared_symbols()
    for name in _native.EXPORTS:
        assert hasattr(L, name)
    assert L.ts_status_string(1).decode() == "buffer too small"


def test_library_has_no_oracle_dependency():
    # The product must not route through the CPU oracle.
    if os.path.exists(LIB):
        out = subprocess.run(["ldd", LIB], capture_output=True, text=True).stdout
        assert "tsoracle" not in out
    for src in glob.glob(os.path.join(ROOT, "turbosort_amd", "**", "*.*"), recursive=True):
        if src.endswith((".py", ".hip", ".h", ".cpp")):
            assert "oracle" not in open(src).read().lower().replace("oracle_", ""), src


def test_segmented_variant_tables_are_consistent():
    """Host-only queries (no GPU call): every size class has a default shape,
    4- and 8-byte keys offer alternatives, 1- and 2-byte keys do not, and
    8-byte keys stop at 11264 keys with values and at 17408 without."""
    from turbosort_amd import _native

    L = _native.load()
    for code in range(10):
        width = [1, 2, 4, 8, 1, 2, 4, 8, 4, 8][code]
        for vals in (0, 1):
            counts = [L.ts_segmented_variants(code, vals, c) for c in range(9)]
            top = 0 if (width == 8 and vals) else 1
            assert all(n >= 1 for n in counts[:8]) and counts[8] >= top
            if width <= 2:
                assert max(counts) == 1
            else:
                assert max(counts) > 1
            for c, n in enumerate(counts):
                for v in range(n):
                    assert L.ts_segmented_variant_name(code, vals, c, v)
            cap = L.ts_segmented_lds_max_len(code, vals)
            assert cap == ((11264 if vals else 17408) if width == 8 else 16384)
    assert L.ts_segmented_variants(2, 1, 9) == 0 and L.ts_segmented_variants(2, 1, -1) == 0


def test_learned_defaults_agree_across_layers():
    """ts_learned_cfg_default (host-only) and the Python LearnedConfig give
    the same defaults."""
    import ctypes

    import turbosort_amd as ts
    from turbosort_amd import _native

    L = _native.load()
    c = _native.LearnedCfg()
    L.ts_learned_cfg_default(ctypes.byref(c))
    py = ts.LearnedConfig()
    assert (c.buckets_log2, c.sample_per_mille, bool(c.train), c.iters, c.points, bool(c.verify)) == (
        py.buckets_log2, py.sample_per_mille, py.train, py.iters, py.points, py.verify)
    assert abs(c.lr - py.lr) < 1e-7
