"""turbosort's sort API on one AMD Instinct MI355X (gfx950).

Mirrors the reference crate's public functions (turbosort 0.2.2, lib.rs:99-229
and by_key.rs:51-159) over device-resident torch tensors:

* ``sort(t)``                          -- lib.rs:99 ``sort``
* ``sort_with_buffer(t, buf)``         -- lib.rs:122 ``sort_with_buffer``
* ``sort_by_key(records, key)``        -- lib.rs:152 ``sort_by_key``
* ``sort_by_key_with_buffer(records, buffer, key)`` -- lib.rs:196
* ``SortByKeyBuffer`` with ``new``, ``with_capacity``, ``capacity``,
  ``order`` and ``apply_order``   -- by_key.rs:51-159
* ``segmented_sort(keys, offsets, values)`` -- a batched form with no single
  reference function: a stable key-value sort of every segment
* ``learned_sort(t)`` -- ``sort``'s output reached through a learned
  partitioner (no reference counterpart)

A tensor stands for a slice: its first dimension indexes the elements.  The
ten sealed key types (key.rs:64-77) are torch.uint8/16/32/64, int8/16/32/64,
float32 and float64; floats sort in IEEE-754 total order, so -0.0 precedes
+0.0, a NaN with its sign bit clear sorts last and one with it set first.

Where the reference panics, these functions raise ``TurbosortPanic`` with the
reference's message.  All work runs in hand-written HIP kernels of
libtsort.so on the tensor's device and the current torch stream; there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes
import sys
from dataclasses import dataclass
from typing import Callable, Optional, Union

import torch

from . import _native

__all__ = [
    "TurbosortPanic", "Context", "KeyField", "sort", "sort_with_buffer", "sort_host", "sort_by_key",
    "sort_by_key_with_buffer", "SortByKeyBuffer", "fill_splitmix64", "checksum", "DTYPE_CODES",
]

DTYPE_CODES = {
    torch.uint8: 0, torch.uint16: 1, torch.uint32: 2, torch.uint64: 3,
    torch.int8: 4, torch.int16: 5, torch.int32: 6, torch.int64: 7,
    torch.float32: 8, torch.float64: 9,
}


class TurbosortPanic(RuntimeError):
    """Raised where the reference crate panics (same message)."""


def _dtype_code(dt: torch.dtype) -> int:
    try:
        return DTYPE_CODES[dt]
    except KeyError:
        raise TypeError(f"{dt} is not a SortableKey type (u8..u64, i8..i64, f32, f64)") from None


def _stream(device: torch.device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class Context:
    """A device workspace (``ts_ctx``): look-back state, histograms, scratch.

    Grow-only, like the reference's ``SortByKeyBuffer``: once it has grown to
    a size, sorts up to that size allocate nothing.  Not thread-safe.
    """

    def __init__(self, device: Union[int, torch.device, None] = None):
        self._lib = _native.load()
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        if dev.type != "cuda":
            raise ValueError("turbosort_amd runs on GPU tensors only")
        self.device = torch.device("cuda", dev.index if dev.index is not None else torch.cuda.current_device())
        h = ctypes.c_void_p()
        rc = self._lib.ts_ctx_create(ctypes.byref(h), self.device.index)
        if rc != 0:
            raise RuntimeError(f"ts_ctx_create failed: {self._lib.ts_status_string(rc).decode()}")
        self._h = h

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def check(self, rc: int) -> None:
        if rc == _native.TS_OK:
            return
        msg = self._lib.ts_ctx_last_error(self._h).decode()
        if rc in (_native.TS_ERR_BUFFER_TOO_SMALL, _native.TS_ERR_LENGTH_MISMATCH):
            raise TurbosortPanic(msg)
        if rc == _native.TS_ERR_INVALID_ARG:
            raise ValueError(msg)
        raise RuntimeError(f"{self._lib.ts_status_string(rc).decode()}: {msg}")

    def reserve(self, dtype: torch.dtype, n: int, record_bytes: int = 0) -> None:
        self.check(self._lib.ts_ctx_reserve(self._h, _dtype_code(dtype), n, record_bytes))

    def workspace_bytes(self) -> int:
        return int(self._lib.ts_ctx_workspace_bytes(self._h))

    def set_profiling(self, on: bool) -> None:
        self.check(self._lib.ts_ctx_set_profiling(self._h, 1 if on else 0))

    RANKINGS = ("ballot", "atomic")  # TS_RANK_BALLOT, TS_RANK_ATOMIC

    @property
    def ranking(self) -> str:
        """How the scatter passes rank keys within a wave: "atomic" (LDS
        atomic adds, chosen where this device passed the lane-order check) or
        "ballot"."""
        return self.RANKINGS[self._lib.ts_ctx_ranking(self._h)]

    def set_ranking(self, method: str) -> None:
        self.check(self._lib.ts_ctx_set_ranking(self._h, self.RANKINGS.index(method)))

    @property
    def split_rounds(self) -> int:
        """Three-way split rounds that long keys-only segments (and the
        learned sort's overflowing buckets) take before the device radix sort."""
        return int(self._lib.ts_ctx_split_rounds(self._h))

    def set_split_rounds(self, rounds: int) -> None:
        self.check(self._lib.ts_ctx_set_split_rounds(self._h, int(rounds)))

    @property
    def divert_min(self) -> int:
        """Keys from which sorts of 8-byte keys without values divert (sort
        only the top digits that carry about log2(n) + 2 bits, then finish the
        runs, as the reference's radix/mod.rs:240-242 does)."""
        return int(self._lib.ts_ctx_divert_min(self._h))

    def set_divert_min(self, n: int) -> None:
        self.check(self._lib.ts_ctx_set_divert_min(self._h, int(n)))

    def check_device(self) -> None:
        """Wait for this context's stream work and raise if a kernel recorded
        an error (a look-back wait that hit its bound, a segment longer than a
        given max_segment_len, invalid segment offsets); clears the record."""
        self.check(self._lib.ts_ctx_check(self._h, _stream(self.device)))

    def last_sort_info(self) -> dict:
        """What the last radix sort on this context did: live digits, passes
        executed, digits left to the divert's finish, whether the divert's
        second sort ran, whether the input was reversed (waits for the device)."""
        info = _native.SortInfo()
        self.check(self._lib.ts_ctx_last_sort_info(self._h, ctypes.byref(info), _stream(self.device)))
        return {name: int(getattr(info, name)) for name, _ in info._fields_}

    def kernel_times(self):
        """[(kernel name, ms, algorithmic bytes)] of the last call, from HIP events."""
        cap = 256
        names = (ctypes.c_char_p * cap)()
        ms = (ctypes.c_float * cap)()
        by = (ctypes.c_double * cap)()
        k = self._lib.ts_ctx_kernel_times(self._h, names, ms, by, cap)
        return [(names[i].decode(), float(ms[i]), float(by[i])) for i in range(k)]

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.ts_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self, _finalizing=sys.is_finalizing):
        # No HIP calls while the interpreter (and the HIP runtime) shuts down.
        # The module global ``sys`` may already be gone then, hence the
        # default argument bound at class creation.
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


_contexts: dict = {}


def _default_context(device: torch.device) -> Context:
    idx = device.index if device.index is not None else torch.cuda.current_device()
    ctx = _contexts.get(idx)
    if ctx is None:
        ctx = _contexts[idx] = Context(idx)
    return ctx


def _slice(t: torch.Tensor, what: str = "slice") -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{what} must live on a GPU")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")


def sort(t: torch.Tensor, *, ctx: Optional[Context] = None) -> None:
    """Sort a 1-D tensor of a SortableKey type in place (lib.rs:99-101)."""
    _slice(t)
    code = _dtype_code(t.dtype)
    ctx = ctx or _default_context(t.device)
    ctx.check(ctx._lib.ts_sort(ctx.handle, code, ctypes.c_void_p(t.data_ptr()), t.numel(), _stream(t.device)))


def sort_with_buffer(t: torch.Tensor, buffer: torch.Tensor, *, ctx: Optional[Context] = None) -> None:
    """Sort in place using ``buffer`` as scratch (lib.rs:122-124).

    Raises ``TurbosortPanic("buffer too small: need N, got M")`` when the buffer
    is shorter than the slice, as the reference panics (dispatch.rs:78-83).
    """
    _slice(t)
    _slice(buffer, "buffer")
    code = _dtype_code(t.dtype)
    if buffer.dtype != t.dtype:
        raise TypeError("buffer must have the slice's dtype")
    ctx = ctx or _default_context(t.device)
    ctx.check(ctx._lib.ts_sort_with_buffer(ctx.handle, code, ctypes.c_void_p(t.data_ptr()), t.numel(),
                                           ctypes.c_void_p(buffer.data_ptr()), buffer.numel(), _stream(t.device)))


def sort_host(t: torch.Tensor, *, ctx: Optional[Context] = None) -> None:
    """Sort a 1-D CPU tensor of a SortableKey type in place (lib.rs:99-101)
    on the GPU: copied to the device, sorted there and copied back; returns
    once ``t`` is sorted.  A pinned tensor (``pin_memory()``) streams at the
    link rate."""
    if t.device.type != "cpu":
        raise ValueError("sort_host takes a CPU tensor; use sort for device tensors")
    if t.dim() != 1 or not t.is_contiguous():
        raise ValueError("slice must be a contiguous 1-D tensor")
    code = _dtype_code(t.dtype)
    dev = torch.device("cuda", torch.cuda.current_device())
    ctx = ctx or _default_context(dev)
    ctx.check(ctx._lib.ts_sort_host(ctx.handle, code, ctypes.c_void_p(t.data_ptr()), t.numel(), _stream(ctx.device)))


@dataclass(frozen=True)
class KeyField:
    """A key stored inside each record: ``dtype`` at byte ``offset``.

    The device counterpart of a key closure such as ``|r| r.score``: the key
    is read from every record exactly once, in index order (by_key.rs:197-200).
    """

    offset: int
    dtype: torch.dtype


KeyLike = Union[KeyField, Callable[[torch.Tensor], torch.Tensor], None]


def _record_bytes(records: torch.Tensor) -> int:
    n = records.shape[0] if records.dim() > 0 else 0
    return records.element_size() * (records.numel() // n if n else int(torch.Size(records.shape[1:]).numel()))


class SortByKeyBuffer:
    """Reusable device memory for ``sort_by_key_with_buffer`` (by_key.rs:51-159).

    Holds the workspace of the last sort and its ``order()``: after a sort,
    position ``i`` of the sorted slice holds the element that was at
    ``order()[i]``.  ``apply_order`` puts another slice of the same length in
    the same order.
    """

    def __init__(self, device: Union[int, torch.device, None] = None):
        self._device = device
        self._ctx: Optional[Context] = None
        self._order = None  # u32 storage, grow-only
        self._len = 0

    @classmethod
    def new(cls, device=None) -> "SortByKeyBuffer":
        return cls(device)

    @classmethod
    def with_capacity(cls, length: int, device=None, key_dtype: torch.dtype = torch.uint64,
                      record_bytes: int = 16) -> "SortByKeyBuffer":
        b = cls(device)
        b._ensure(torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device),
                  length)
        b._ctx.reserve(key_dtype, length, record_bytes)
        return b

    def _ensure_context(self, device: torch.device) -> None:
        if self._ctx is None or self._ctx.device.index != (device.index if device.index is not None else
                                                           torch.cuda.current_device()):
            self._ctx = Context(device)
            self._order = None

    def _ensure(self, device: torch.device, n: int) -> None:
        self._ensure_context(device)
        if self._order is None or self._order.numel() < n:
            self._order = torch.empty(max(n, 1), dtype=torch.uint32, device=self._ctx.device)

    def capacity(self) -> int:
        """How many elements it can order without growing its order array."""
        return 0 if self._order is None else int(self._order.numel())

    def order(self) -> torch.Tensor:
        """The order of the last sort (u32, length of that slice; empty before)."""
        if self._order is None:
            dev = self._ctx.device if self._ctx else "cuda"
            return torch.empty(0, dtype=torch.uint32, device=dev)
        return self._order[: self._len]

    def apply_order(self, other: torch.Tensor, scratch=None) -> None:
        """Put ``other`` in the order of the last sort (by_key.rs:136-144).

        ``scratch`` exists for signature parity; the buffer's own workspace
        serves as the staging memory.
        """
        del scratch
        _slice(other, "other")
        n = other.shape[0] if other.dim() > 0 else 0
        if self._ctx is None:
            if n != 0:
                raise TurbosortPanic("apply_order: the slice's length differs from the last sorted slice's")
            return
        ctx = self._ctx
        order = ctypes.c_void_p(self._order.data_ptr() if self._order is not None else None)
        ctx.check(ctx._lib.ts_apply_order(ctx.handle, order, self._len,
                                          ctypes.c_void_p(other.data_ptr()), n, _record_bytes(other) if n else 1,
                                          _stream(other.device)))

    def __repr__(self) -> str:
        return f"SortByKeyBuffer {{ capacity: {self.capacity()} }}"


def sort_by_key_with_buffer(records: torch.Tensor, buffer: SortByKeyBuffer, key: KeyLike = None) -> None:
    """Stable sort of ``records`` by a key (lib.rs:196-202, by_key.rs:175-235).

    ``key`` is a ``KeyField`` (a key stored in each record), a callable that
    maps the records to a 1-D key tensor (evaluated once), or ``None`` when
    ``records`` is itself a 1-D key tensor.
    """
    _slice(records, "records")
    n = records.shape[0] if records.dim() > 0 else 0
    if n > U32_MAX:
        _sort_by_key_unordered(records, buffer, key, n)
        return
    buffer._ensure(records.device, n)
    ctx = buffer._ctx
    stream = _stream(records.device)
    order_ptr = ctypes.c_void_p(buffer._order.data_ptr())
    if isinstance(key, KeyField):
        code = _dtype_code(key.dtype)
        rc = ctx._lib.ts_sort_by_key(ctx.handle, code, ctypes.c_void_p(records.data_ptr()), n,
                                     _record_bytes(records), key.offset, order_ptr, 1, stream)
        ctx.check(rc)
    else:
        keys = records if key is None else key(records)
        _slice(keys, "key tensor")
        if keys.dim() != 1 or keys.shape[0] != n:
            raise ValueError("the key must give one value per element")
        code = _dtype_code(keys.dtype)
        rc = ctx._lib.ts_sort_by_key(ctx.handle, code, ctypes.c_void_p(keys.data_ptr()), n, keys.element_size(), 0,
                                     order_ptr, 1 if key is None else 0, stream)
        ctx.check(rc)
        if key is not None and n:
            ctx.check(ctx._lib.ts_apply_order(ctx.handle, order_ptr, n, ctypes.c_void_p(records.data_ptr()), n,
                                              _record_bytes(records), stream))
    buffer._len = n


U32_MAX = 0xFFFFFFFF


def _sort_by_key_unordered(records: torch.Tensor, buffer: SortByKeyBuffer, key: KeyLike, n: int) -> None:
    """More elements than a u32 order can index (by_key.rs:190-194): the
    records are sorted stably by the key, and the buffer's order is left
    empty, so that a later apply_order of a non-empty slice is a length
    mismatch."""
    buffer._ensure_context(records.device)
    ctx = buffer._ctx
    stream = _stream(records.device)
    if isinstance(key, KeyField):
        code = _dtype_code(key.dtype)
        rb = _record_bytes(records)
        rc = ctx._lib.ts_sort_by_key_unordered(ctx.handle, code, ctypes.c_void_p(records.data_ptr()), rb, key.offset,
                                               ctypes.c_void_p(records.data_ptr()), n, rb, stream)
    else:
        keys = records if key is None else key(records)
        _slice(keys, "key tensor")
        if keys.dim() != 1 or keys.shape[0] != n:
            raise ValueError("the key must give one value per element")
        code = _dtype_code(keys.dtype)
        rc = ctx._lib.ts_sort_by_key_unordered(ctx.handle, code, ctypes.c_void_p(keys.data_ptr()), keys.element_size(),
                                               0, ctypes.c_void_p(records.data_ptr()), n, _record_bytes(records),
                                               stream)
    ctx.check(rc)
    buffer._len = 0


def sort_by_key(records: torch.Tensor, key: KeyLike = None) -> None:
    """Stable sort by a key with a fresh buffer (lib.rs:152-158)."""
    sort_by_key_with_buffer(records, SortByKeyBuffer(records.device), key)


def segmented_sort(keys: torch.Tensor, offsets: torch.Tensor, values: Optional[torch.Tensor] = None, *,
                   max_segment_len: int = 0, ctx: Optional[Context] = None) -> None:
    """Sort every segment ``keys[offsets[s]:offsets[s+1]]`` stably, in place.

    ``values`` (u32 or i32, same length as ``keys``) move with their keys.  Per
    segment this equals ``sort_by_key_with_buffer`` keyed by the key itself
    plus ``apply_order`` on the values (by_key.rs:136-235).  ``offsets`` holds
    n_segments + 1 non-decreasing u64 (or i64) positions on the keys' device.
    ``max_segment_len`` bounds the segment lengths (0: unknown); a bound no
    larger than ``segmented_lds_max_len`` keeps the call free of host waits.
    """
    _slice(keys, "keys")
    _slice(offsets, "offsets")
    code = _dtype_code(keys.dtype)
    if offsets.dtype not in (torch.uint64, torch.int64) or offsets.dim() != 1:
        raise TypeError("offsets must be a 1-D u64 or i64 tensor")
    if offsets.device != keys.device:
        raise ValueError("offsets must live on the keys' device")
    vp = None
    if values is not None:
        _slice(values, "values")
        if values.dtype not in (torch.uint32, torch.int32):
            raise TypeError("values must be u32 or i32")
        if values.numel() != keys.numel() or values.device != keys.device:
            raise ValueError("values must match the keys' length and device")
        vp = ctypes.c_void_p(values.data_ptr())
    nseg = max(offsets.numel() - 1, 0)
    ctx = ctx or _default_context(keys.device)
    ctx.check(ctx._lib.ts_segmented_sort_pairs(ctx.handle, code, ctypes.c_void_p(keys.data_ptr()), vp, keys.numel(),
                                               ctypes.c_void_p(offsets.data_ptr()), nseg, int(max_segment_len),
                                               _stream(keys.device)))


def segmented_lds_max_len(dtype: torch.dtype, with_values: bool = True) -> int:
    """Longest segment one workgroup sorts for this key type."""
    return int(_native.load().ts_segmented_lds_max_len(_dtype_code(dtype), int(bool(with_values))))


@dataclass
class LearnedConfig:
    """Settings of ``learned_sort`` (ts_learned_cfg)."""

    buckets_log2: int = 16
    sample_per_mille: int = 5
    train: bool = True
    iters: int = 40
    lr: float = 0.02
    points: int = 16384
    verify: bool = True


def learned_sort(t: torch.Tensor, *, config: Optional[LearnedConfig] = None, ctx: Optional[Context] = None) -> dict:
    """Sort a 1-D tensor of 4- or 8-byte keys in place through the learned partitioner.

    The result equals ``sort(t)``: a monotone CDF model fitted on the device to
    a sample buckets the keys, and every bucket is then sorted.  Returns the
    run's report (sample size, bucket overflow, model loss before and after
    fitting).
    """
    _slice(t)
    code = _dtype_code(t.dtype)
    c = config or LearnedConfig()
    cfg = _native.LearnedCfg(c.buckets_log2, c.sample_per_mille, int(bool(c.train)), c.iters, c.lr, c.points,
                             int(bool(c.verify)))
    rep = _native.LearnedReport()
    ctx = ctx or _default_context(t.device)
    ctx.check(ctx._lib.ts_learned_sort(ctx.handle, code, ctypes.c_void_p(t.data_ptr()), t.numel(), ctypes.byref(cfg),
                                       ctypes.byref(rep), _stream(t.device)))
    return {name: getattr(rep, name) for name, _ in rep._fields_}


def pwl_fit(u: torch.Tensor, target: torch.Tensor, theta: torch.Tensor, *, iters: int = 0, lr: float = 0.05,
            ctx: Optional[Context] = None):
    """One run of the learned model's training kernel.

    ``u`` (knot units, sorted) and ``target`` are float32 points; ``theta``
    (1024 float32) is updated in place.  Returns (gradient at the initial
    theta, loss before, loss after ``iters`` Adam steps).
    """
    for x, nm in ((u, "u"), (target, "target"), (theta, "theta")):
        _slice(x, nm)
        if x.dtype != torch.float32:
            raise TypeError(f"{nm} must be float32")
    if theta.numel() != 1024 or u.numel() != target.numel():
        raise ValueError("theta has 1024 entries; u and target have equal lengths")
    grad = torch.empty(1024, dtype=torch.float32, device=theta.device)
    loss = torch.zeros(2, dtype=torch.float32, device=theta.device)
    ctx = ctx or _default_context(theta.device)
    ctx.check(ctx._lib.ts_pwl_fit(ctx.handle, ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(target.data_ptr()),
                                  u.numel(), ctypes.c_void_p(theta.data_ptr()), ctypes.c_void_p(grad.data_ptr()),
                                  ctypes.c_void_p(loss.data_ptr()), int(iters), float(lr), _stream(theta.device)))
    torch.cuda.synchronize(theta.device)
    lh = loss.cpu()
    return grad, float(lh[0]), float(lh[1])


def fill_splitmix64(out: torch.Tensor, seed: int, *, ctx: Optional[Context] = None) -> None:
    """Synthetic keys on the device: output i of SplitMix64 with state ``seed``.

    uint32 tensors take the high 32 bits; uint64 the full output.
    """
    _slice(out, "out")
    kind = {torch.uint32: 0, torch.uint64: 1}.get(out.dtype)
    if kind is None:
        raise TypeError("fill_splitmix64 fills uint32 or uint64 tensors")
    ctx = ctx or _default_context(out.device)
    ctx.check(ctx._lib.ts_fill_splitmix64(ctx.handle, ctypes.c_void_p(out.data_ptr()), out.numel(), seed, kind,
                                          _stream(out.device)))


def fill_splitmix64_zero_extended(out: torch.Tensor, seed: int, *, ctx: Optional[Context] = None) -> None:
    """uint64 keys holding the zero-extended high 32 bits of SplitMix64."""
    _slice(out, "out")
    if out.dtype != torch.uint64:
        raise TypeError("needs a uint64 tensor")
    ctx = ctx or _default_context(out.device)
    ctx.check(ctx._lib.ts_fill_splitmix64(ctx.handle, ctypes.c_void_p(out.data_ptr()), out.numel(), seed, 2,
                                          _stream(out.device)))


def verify_sorted(t: torch.Tensor, expect=None, *, ctx: Optional[Context] = None) -> None:
    """The output judge (ts_verify_sorted): raises ``RuntimeError`` unless no
    adjacent pair of ``t`` is out of order and, when ``expect`` = (sum, xor)
    of ``checksum`` on the input is given, ``t`` holds the same multiset."""
    _slice(t)
    ctx = ctx or _default_context(t.device)
    exp = None
    if expect is not None:
        exp = (ctypes.c_uint64 * 2)(int(expect[0]), int(expect[1]))
    ctx.check(ctx._lib.ts_verify_sorted(ctx.handle, _dtype_code(t.dtype), ctypes.c_void_p(t.data_ptr()), t.numel(),
                                        exp, _stream(t.device)))


def checksum(t: torch.Tensor, *, ctx: Optional[Context] = None):
    """(multiset sum, multiset xor, adjacent pairs out of order) of a key tensor."""
    _slice(t)
    ctx = ctx or _default_context(t.device)
    sx = (ctypes.c_uint64 * 2)()
    bad = ctypes.c_uint64()
    ctx.check(ctx._lib.ts_checksum(ctx.handle, _dtype_code(t.dtype), ctypes.c_void_p(t.data_ptr()), t.numel(), sx,
                                   ctypes.byref(bad), _stream(t.device)))
    return int(sx[0]), int(sx[1]), int(bad.value)
