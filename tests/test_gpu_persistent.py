"""Sorts large enough for the persistent scatter passes, key for key.

Keys-only and key-value sorts of 4- and 8-byte keys, record-carrying by-key
sorts and the learned sort's partition run their scatter passes as
persistent grids (one workgroup per CU claiming tiles, each claiming its
next tile on the segment it works on while its stores go out) from 8 tiles
per CU on: about 6.7e7 4-byte keys, 3.8e7 8-byte keys, 1.5e7 16-byte
records with their positions or 3.4e7 learned keys on 256 CUs.  These sizes
are too large for the oracle to sort in seconds, so each output is compared
with
torch.sort of the same keys on the device (a keys-only sort has one correct
output), after a check that the path under test ran (executed passes).
Cases: uniform u32 at the headline size, i32 and f32 with ragged last tiles
(the signed and float transforms of the last pass), u32 whose upper three
bits of every byte are zero (all keys in one value segment from the second
pass on, so every workgroup but those of one XCD finds its own segment empty
and takes tiles elsewhere), skewed u64 (a lognormal mixture: unequal
segments, workgroups moving between them), 16- and 8-byte records carried
with their positions (the by-key passes), checked against a stable
torch.sort,
the learned sort's bucket-digit passes on the lognormal mixture, and the
(key, position) passes of a by-key sort keyed by a separate array.
"""
from __future__ import annotations

import pytest
import torch

import turbosort_amd as ts

pytestmark = pytest.mark.gpu

PERSIST_TILES_PER_CU = 8


def persistent_size(gpu, tile: int) -> int:
    cus = torch.cuda.get_device_properties(gpu).multi_processor_count
    return PERSIST_TILES_PER_CU * cus * tile


def as_signed_order(t: torch.Tensor) -> torch.Tensor:
    """An int64 view of unsigned keys whose signed order is their unsigned order."""
    if t.dtype == torch.uint32:
        return t.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    return t.view(torch.int64) ^ (-(1 << 63))


def check_against_torch(t: torch.Tensor, ctx: ts.Context, passes: int) -> None:
    if t.dtype in (torch.uint32, torch.uint64):
        want = torch.sort(as_signed_order(t)).values
    else:
        want = torch.sort(t).values
    ts.sort(t, ctx=ctx)
    info = ctx.last_sort_info()
    got = as_signed_order(t) if t.dtype in (torch.uint32, torch.uint64) else t
    torch.cuda.synchronize()
    assert info["executed_passes"] == passes, info
    assert torch.equal(got, want)
    ctx.check_device()


def test_uniform_u32_headline_size(gpu):
    c = ts.Context(gpu)
    n = 1 << 28
    assert n >= persistent_size(gpu, 32768)
    g = torch.Generator(device=gpu).manual_seed(11)
    t = torch.randint(-(1 << 31), 1 << 31, (n,), dtype=torch.int32, device=gpu, generator=g).view(torch.uint32)
    check_against_torch(t, c, 4)


@pytest.mark.parametrize("dtype", [torch.int32, torch.float32])
def test_ragged_signed_and_float(gpu, dtype):
    c = ts.Context(gpu)
    n = persistent_size(gpu, 32768) + 12_345
    g = torch.Generator(device=gpu).manual_seed(12)
    if dtype == torch.int32:
        t = torch.randint(-(1 << 31), 1 << 31, (n,), dtype=torch.int32, device=gpu, generator=g)
    else:
        # Normal values and infinities, no NaN or zero (torch.sort orders
        # -0.0 and 0.0 as equal, the radix key does not).
        t = torch.randn(n, device=gpu, generator=g) * 1e3
        t[t == 0] = 1.0
        t[:: 1001] = float("inf")
        t[5 :: 1003] = float("-inf")
    check_against_torch(t, c, 4)


def test_u32_one_value_segment(gpu):
    """Every byte below 0x20: from the second pass on all keys are in value
    segment 0, so the workgroups whose own segment is empty take its tiles."""
    c = ts.Context(gpu)
    n = persistent_size(gpu, 32768) + 777
    g = torch.Generator(device=gpu).manual_seed(13)
    t = torch.randint(-(1 << 31), 1 << 31, (n,), dtype=torch.int32, device=gpu, generator=g)
    t = (t & 0x1F1F1F1F).view(torch.uint32)
    check_against_torch(t, c, 4)


def test_skewed_u64(gpu):
    """Lognormal mixture (the learned workload's distribution) of 2^26 + 3
    u64 keys: value segments of very different sizes in every pass."""
    import bench

    c = ts.Context(gpu)
    n = (1 << 26) + 3
    assert n >= persistent_size(gpu, 512 * 36)
    t = bench.lognormal_keys_dev(n, 21, c, as_f64=False)
    assert t.dtype == torch.uint64
    want = torch.sort(as_signed_order(t)).values
    ts.sort(t, ctx=c)
    torch.cuda.synchronize()
    assert torch.equal(as_signed_order(t), want)
    c.check_device()


@pytest.mark.parametrize("rec_bytes,tile", [(16, 512 * 14), (8, 512 * 24)])
def test_carried_records_with_order(gpu, rec_bytes, tile):
    """The by-key shape at persistent size: 16-byte records {u32 id, f32
    score, u64 payload}, or 8-byte records {u32 id, f32 score}, keyed on the
    score (duplicates, NaN, -0.0, inf), carried through the passes with their
    positions.  order() must be the stable argsort of the scores' radix keys
    (torch.sort with stable=True of the same keys), and the records must end
    in that order."""
    n = persistent_size(gpu, tile) + 1001
    g = torch.Generator(device=gpu).manual_seed(14)
    ids = torch.arange(n, dtype=torch.int32, device=gpu)
    score = torch.randn(n, generator=g, device=gpu)
    score = torch.where(torch.rand(n, generator=g, device=gpu) < 0.2, score.round(), score)
    score[::997] = float("nan")
    score[1::1009] = -0.0
    score[2::1013] = float("inf")
    rec = torch.empty(n, rec_bytes, dtype=torch.uint8, device=gpu)
    rec[:, 0:4] = ids.view(torch.uint8).reshape(n, 4)
    rec[:, 4:8] = score.view(torch.uint8).reshape(n, 4)
    if rec_bytes == 16:
        payload = torch.randint(-(1 << 62), 1 << 62, (n,), generator=g, device=gpu)
        rec[:, 8:16] = payload.view(torch.uint8).reshape(n, 8)
    src = rec.clone()
    buf = ts.SortByKeyBuffer()
    ts.sort_by_key_with_buffer(rec, buf, ts.KeyField(4, torch.float32))
    order = buf.order().view(torch.int32).long() & 0xFFFFFFFF
    sbits = src[:, 4:8].contiguous().view(torch.int32).reshape(-1).long() & 0xFFFFFFFF
    radix = torch.where(sbits >= 0x80000000, sbits ^ 0xFFFFFFFF, sbits | 0x80000000)
    want = torch.sort(radix, stable=True).indices
    torch.cuda.synchronize()
    assert torch.equal(order, want)
    assert torch.equal(rec, src[want])


@pytest.mark.parametrize("as_f64", [False, True])
def test_learned_partition_at_persistent_size(gpu, as_f64):
    """The learned sort's bucket-digit passes run as persistent grids from 8
    tiles of 16384 keys per CU (about 3.4e7 keys): 6e7 keys of the lognormal
    mixture with hot values, at the default 1 % sample, judged; the output
    must equal torch.sort of the same keys."""
    import bench

    c = ts.Context(gpu)
    n = 60_000_000
    assert n >= persistent_size(gpu, 1024 * 16)
    t = bench.lognormal_keys_dev(n, 23, c, as_f64=as_f64)
    want = torch.sort(t).values if as_f64 else torch.sort(as_signed_order(t)).values
    rep = ts.learned_sort(t, config=ts.LearnedConfig(sample_per_mille=10), ctx=c)
    torch.cuda.synchronize()
    assert rep["verified"] == 1, rep
    assert torch.equal(t if as_f64 else as_signed_order(t), want)
    c.check_device()


@pytest.mark.parametrize("key_dtype,tile", [(torch.int32, 512 * 36), (torch.float64, 512 * 24)])
def test_keys_of_their_own_with_positions(gpu, key_dtype, tile):
    """A by-key sort keyed by a separate array sorts (key, position) pairs:
    the key-value passes at persistent size.  The records (their indices)
    must end in the stable order of the keys (torch.sort with stable=True;
    duplicated keys, and no NaN or zero, whose orders differ in torch)."""
    n = persistent_size(gpu, tile) + 4321
    g = torch.Generator(device=gpu).manual_seed(15)
    if key_dtype == torch.int32:
        keys = torch.randint(-(1 << 20), 1 << 20, (n,), dtype=torch.int32, device=gpu, generator=g)
    else:
        keys = torch.randn(n, dtype=torch.float64, device=gpu, generator=g).round(decimals=3)
        keys[keys == 0] = 0.5
    ids = torch.arange(n, dtype=torch.int32, device=gpu)
    want = torch.sort(keys, stable=True).indices.to(torch.int32)
    ts.sort_by_key(ids, lambda _: keys)
    torch.cuda.synchronize()
    assert torch.equal(ids, want)
