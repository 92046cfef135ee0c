"""ts_sort_with_buffer_static, the entry point that allocates nothing.

Called through the C ABI and compared, bit for bit, with the existing
ts_sort_with_buffer (and the oracle) on two calls in a row, the second with
a smaller input, for 4- and 8-byte keys (the 8-byte case also through the
divert and its finish), and at sizes whose scatter passes run as persistent
grids against ts_sort_with_buffer on the device; the context it runs on
grows no workspace, and n beyond the stated maximum is refused from the host
before any launch.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import turbosort_amd as ts
from gpu_util import bits, to_dev, to_host
from oracle import closed_form as cf
from shapes import random_of

pytestmark = pytest.mark.gpu


def static_sort(ctx, t, buf):
    L = ts._native.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)
    return L.ts_sort_with_buffer_static(ctx.handle, ts.DTYPE_CODES[t.dtype], ctypes.c_void_p(t.data_ptr()),
                                        t.numel(), ctypes.c_void_p(buf.data_ptr()), buf.numel(), stream)


@pytest.mark.parametrize("name,sizes", [("u32", (3_000_017, 1_000_003, 3000)),
                                        ("f64", (9_000_001, 2_500_001, 77)),
                                        ("i16", (300_000, 5000, 2))])
def test_static_matches_sort_with_buffer_on_shrinking_calls(gpu, name, sizes):
    ctx = ts.Context(gpu)
    ctx.set_divert_min((1 << 21) + 1)
    ref = ts.Context(gpu)
    ref.set_divert_min((1 << 21) + 1)
    for i, n in enumerate(sizes):
        x = random_of(name, n, seed=100 + i + n)
        a, b = to_dev(x, gpu), to_dev(x, gpu)
        ctx.check(static_sort(ctx, a, torch.empty_like(a)))
        ts.sort_with_buffer(b, torch.empty_like(b), ctx=ref)
        got, want = to_host(a, x.dtype), to_host(b, x.dtype)
        assert np.array_equal(bits(got), bits(want)), (name, n)
        assert np.array_equal(bits(got), bits(cf.sort(x))), (name, n)
        if name == "f64" and n > (1 << 21):
            assert ctx.last_sort_info()["skipped_digits"] > 0  # diverted, finished through the static array
    ctx.check_device()
    assert ctx.workspace_bytes() == 0  # the context itself grew nothing


@pytest.mark.parametrize("dtype,sizes", [(torch.int32, ((1 << 27) + 5, (1 << 24) + 43)),
                                         (torch.uint64, ((1 << 26) + 15, 3_000_001))])
def test_static_matches_sort_with_buffer_at_persistent_sizes(gpu, dtype, sizes):
    """Large enough that the scatter passes run as persistent grids (the
    first call; the second, smaller one launches a workgroup per tile), on
    the device against ts_sort_with_buffer of the same keys."""
    ctx = ts.Context(gpu)
    ref = ts.Context(gpu)
    for i, n in enumerate(sizes):
        g = torch.Generator(device=gpu).manual_seed(40 + i)
        lo, hi = (-(1 << 31), 1 << 31) if dtype == torch.int32 else (-(1 << 63), (1 << 63) - 1)
        a = torch.randint(lo, hi, (n,), dtype=torch.int64 if dtype != torch.int32 else dtype, device=gpu,
                          generator=g)
        a = a.view(dtype) if dtype != torch.int32 else a
        b = a.clone()
        ctx.check(static_sort(ctx, a, torch.empty_like(a)))
        ts.sort_with_buffer(b, torch.empty_like(b), ctx=ref)
        torch.cuda.synchronize()
        assert torch.equal(a.view(torch.uint8), b.view(torch.uint8)), (dtype, n)
    ctx.check_device()
    assert ctx.workspace_bytes() == 0


def test_static_refuses_more_than_its_maximum(gpu):
    ctx = ts.Context(gpu)
    t = torch.empty(16, dtype=torch.uint32, device=gpu)
    n = (1 << 28) + 1
    L = ts._native.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    rc = L.ts_sort_with_buffer_static(ctx.handle, 2, ctypes.c_void_p(t.data_ptr()), n,
                                      ctypes.c_void_p(t.data_ptr()), n, stream)
    assert rc == ts._native.TS_ERR_INVALID_ARG
    assert "at most 2^28" in L.ts_ctx_last_error(ctx.handle).decode()
    with pytest.raises(ts.TurbosortPanic, match="buffer too small: need 16, got 8"):
        ctx.check(static_sort(ctx, t, t[:8]))
