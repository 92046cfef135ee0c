"""The host-side branch of sort_by_key above u32::MAX elements, without a GPU.

by_key.rs:190-194: a slice longer than u32::MAX is sorted stably by key
(`slice.sort_by_key`) and the buffer's order is cleared, so that a later
apply_order of a non-empty slice is a length mismatch.  The facade must send
such a slice to ts_sort_by_key_unordered -- keys inside the records for a
KeyField, keys of their own for a key closure, the slice itself for bare keys
-- and must leave order() empty.  The device call itself is stood in for by
a recording library; the tensors are shape-only (meta) tensors, so nothing
is allocated.
"""
from __future__ import annotations

import pytest
import torch

import turbosort_amd as ts

N = (1 << 32) + 7


class FakeLib:
    def __init__(self):
        self.calls = []

    def ts_sort_by_key_unordered(self, *args):
        self.calls.append(("unordered", args))
        return 0

    def ts_sort_by_key(self, *args):
        self.calls.append(("ordered", args))
        return 0

    def ts_apply_order(self, handle, order, order_len, other, n, eb, stream):
        self.calls.append(("apply", (order_len, n)))
        return 0 if order_len == n else 2


class FakeCtx:
    def __init__(self):
        self._lib = FakeLib()
        self.handle = object()
        self.device = torch.device("meta")

    def check(self, rc):
        if rc == 2:
            raise ts.TurbosortPanic("apply_order: the slice's length differs from the last sorted slice's")
        assert rc == 0


@pytest.fixture
def fake(monkeypatch):
    ctx = FakeCtx()
    monkeypatch.setattr(ts, "_slice", lambda t, what="slice": None)
    monkeypatch.setattr(ts, "_stream", lambda device: None)

    def ensure_context(self, device):
        self._ctx = ctx

    monkeypatch.setattr(ts.SortByKeyBuffer, "_ensure_context", ensure_context)
    monkeypatch.setattr(ts.SortByKeyBuffer, "_ensure", lambda self, device, n: pytest.fail("an order was allocated"))
    return ctx


def args_of(call):
    kind, a = call
    assert kind == "unordered"
    _, code, _keys, key_stride, key_offset, _records, n, record_bytes, _stream = a
    return code, key_stride, key_offset, n, record_bytes


def test_key_field_sorts_the_records_unordered(fake):
    recs = torch.empty((N, 16), dtype=torch.uint8, device="meta")
    buf = ts.SortByKeyBuffer()
    ts.sort_by_key_with_buffer(recs, buf, ts.KeyField(4, torch.float32))
    assert len(fake._lib.calls) == 1
    assert args_of(fake._lib.calls[0]) == (ts._dtype_code(torch.float32), 16, 4, N, 16)
    assert buf.order().numel() == 0
    with pytest.raises(ts.TurbosortPanic, match="length differs"):
        buf.apply_order(torch.empty(3, dtype=torch.int64, device="meta"))
    assert fake._lib.calls[-1] == ("apply", (0, 3))


def test_key_closure_passes_its_keys(fake):
    recs = torch.empty((N, 3), dtype=torch.uint8, device="meta")
    keys = torch.empty(N, dtype=torch.int16, device="meta")
    buf = ts.SortByKeyBuffer()
    ts.sort_by_key_with_buffer(recs, buf, lambda r: keys)
    assert args_of(fake._lib.calls[0]) == (ts._dtype_code(torch.int16), 2, 0, N, 3)
    assert buf.order().numel() == 0


def test_bare_keys_sort_by_themselves(fake):
    keys = torch.empty(N, dtype=torch.uint64, device="meta")
    buf = ts.SortByKeyBuffer()
    ts.sort_by_key_with_buffer(keys, buf)
    assert args_of(fake._lib.calls[0]) == (ts._dtype_code(torch.uint64), 8, 0, N, 8)
    assert buf.order().numel() == 0


def test_at_u32_max_the_ordered_path_is_kept(monkeypatch, fake):
    """Exactly u32::MAX elements still get an order (by_key.rs:190 is a strict >)."""
    n = (1 << 32) - 1
    seen = {}

    def ensure(self, device, m):
        seen["n"] = m
        self._ctx = fake
        self._order = torch.empty(m, dtype=torch.uint32, device="meta")

    monkeypatch.setattr(ts.SortByKeyBuffer, "_ensure", ensure)
    keys = torch.empty(n, dtype=torch.uint32, device="meta")
    buf = ts.SortByKeyBuffer()
    ts.sort_by_key_with_buffer(keys, buf)
    assert seen["n"] == n
    assert fake._lib.calls[0][0] == "ordered"
    assert buf.order().numel() == n
