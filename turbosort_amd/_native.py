"""ctypes binding of libtsort.so (the C ABI declared in include/tsort.h).

The library is built in-tree (``python -m turbosort_amd.build``).  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtsort.so")

# Every symbol include/tsort.h declares.
EXPORTS = [
    "ts_ctx_create", "ts_ctx_destroy", "ts_ctx_reserve", "ts_ctx_workspace_bytes", "ts_ctx_last_error",
    "ts_status_string", "ts_sort", "ts_sort_with_buffer", "ts_sort_by_key", "ts_apply_order",
    "ts_fill_splitmix64", "ts_checksum", "ts_ctx_set_profiling", "ts_ctx_kernel_times",
    "ts_sbk_create", "ts_sbk_destroy", "ts_sbk_capacity", "ts_sbk_order", "ts_sbk_sort", "ts_sbk_apply_order",
    "ts_sbk_last_error", "ts_segmented_sort_pairs", "ts_segmented_lds_max_len",
    "ts_learned_cfg_default", "ts_learned_sort", "ts_pwl_fit",
    "ts_segmented_variants", "ts_segmented_variant_name", "ts_segmented_set_variant", "ts_segmented_get_variant",
    "ts_ctx_set_ranking", "ts_ctx_ranking", "ts_sort_host", "ts_ctx_set_split_rounds", "ts_ctx_split_rounds",
    "ts_ctx_set_divert_min", "ts_ctx_divert_min", "ts_ctx_check", "ts_ctx_last_sort_info", "ts_plan_status_bytes",
    "ts_verify_sorted", "ts_sort_with_buffer_static", "ts_sort_by_key_unordered",
]


class LearnedCfg(ctypes.Structure):
    """ts_learned_cfg"""
    _fields_ = [("buckets_log2", ctypes.c_uint32), ("sample_per_mille", ctypes.c_uint32),
                ("train", ctypes.c_uint32), ("iters", ctypes.c_uint32), ("lr", ctypes.c_float),
                ("points", ctypes.c_uint32), ("verify", ctypes.c_uint32)]


class LearnedReport(ctypes.Structure):
    """ts_learned_report"""
    _fields_ = [("sample", ctypes.c_uint64), ("buckets", ctypes.c_uint32), ("overflow_buckets", ctypes.c_uint32),
                ("split_rounds", ctypes.c_uint32), ("overflow_keys", ctypes.c_uint64), ("largest_bucket", ctypes.c_uint64),
                ("loss_before", ctypes.c_float), ("loss_after", ctypes.c_float), ("loss_uniform", ctypes.c_float),
                ("verified", ctypes.c_uint32)]

class SortInfo(ctypes.Structure):
    """ts_sort_info"""
    _fields_ = [("live_digits", ctypes.c_uint32), ("executed_passes", ctypes.c_uint32),
                ("skipped_digits", ctypes.c_uint32), ("finish_shift", ctypes.c_uint32),
                ("fallback", ctypes.c_uint32), ("reversed", ctypes.c_uint32)]


TS_OK = 0
TS_ERR_BUFFER_TOO_SMALL = 1
TS_ERR_LENGTH_MISMATCH = 2
TS_ERR_TOO_LONG = 3
TS_ERR_INVALID_ARG = 4
TS_ERR_HIP = 5
TS_ERR_VERIFY_FAILED = 6
TS_ERR_NO_DEVICE = 7

_lib = None


def load() -> ctypes.CDLL:
    """Load libtsort.so and declare its signatures (no GPU call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m turbosort_amd.build` "
            "(turbosort_amd has no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, i32, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64
    L.ts_ctx_create.argtypes = [ctypes.POINTER(vp), i32]
    L.ts_ctx_create.restype = i32
    L.ts_ctx_destroy.argtypes = [vp]
    L.ts_ctx_destroy.restype = None
    L.ts_ctx_reserve.argtypes = [vp, i32, sz, sz]
    L.ts_ctx_reserve.restype = i32
    L.ts_ctx_workspace_bytes.argtypes = [vp]
    L.ts_ctx_workspace_bytes.restype = sz
    L.ts_ctx_last_error.argtypes = [vp]
    L.ts_ctx_last_error.restype = ctypes.c_char_p
    L.ts_status_string.argtypes = [i32]
    L.ts_status_string.restype = ctypes.c_char_p
    L.ts_sort.argtypes = [vp, i32, vp, sz, vp]
    L.ts_sort.restype = i32
    L.ts_sort_with_buffer.argtypes = [vp, i32, vp, sz, vp, sz, vp]
    L.ts_sort_with_buffer.restype = i32
    L.ts_sort_by_key.argtypes = [vp, i32, vp, sz, sz, sz, vp, i32, vp]
    L.ts_sort_by_key.restype = i32
    L.ts_apply_order.argtypes = [vp, vp, sz, vp, sz, sz, vp]
    L.ts_apply_order.restype = i32
    L.ts_sort_by_key_unordered.argtypes = [vp, i32, vp, sz, sz, vp, sz, sz, vp]
    L.ts_sort_by_key_unordered.restype = i32
    L.ts_segmented_sort_pairs.argtypes = [vp, i32, vp, vp, sz, vp, sz, sz, vp]
    L.ts_segmented_sort_pairs.restype = i32
    L.ts_segmented_lds_max_len.argtypes = [i32, i32]
    L.ts_segmented_lds_max_len.restype = sz
    L.ts_segmented_variants.argtypes = [i32, i32, i32]
    L.ts_segmented_variants.restype = i32
    L.ts_segmented_variant_name.argtypes = [i32, i32, i32, i32]
    L.ts_segmented_variant_name.restype = ctypes.c_char_p
    L.ts_segmented_set_variant.argtypes = [vp, i32, i32, i32, i32]
    L.ts_segmented_set_variant.restype = i32
    L.ts_segmented_get_variant.argtypes = [vp, i32, i32, i32]
    L.ts_segmented_get_variant.restype = i32
    L.ts_learned_cfg_default.argtypes = [ctypes.POINTER(LearnedCfg)]
    L.ts_learned_cfg_default.restype = None
    L.ts_learned_sort.argtypes = [vp, i32, vp, sz, ctypes.POINTER(LearnedCfg), ctypes.POINTER(LearnedReport), vp]
    L.ts_learned_sort.restype = i32
    L.ts_pwl_fit.argtypes = [vp, vp, vp, sz, vp, vp, vp, i32, ctypes.c_float, vp]
    L.ts_pwl_fit.restype = i32
    L.ts_fill_splitmix64.argtypes = [vp, vp, sz, u64, i32, vp]
    L.ts_fill_splitmix64.restype = i32
    L.ts_checksum.argtypes = [vp, i32, vp, sz, ctypes.POINTER(u64), ctypes.POINTER(u64), vp]
    L.ts_checksum.restype = i32
    L.ts_ctx_set_profiling.argtypes = [vp, i32]
    L.ts_ctx_set_profiling.restype = i32
    L.ts_sort_host.argtypes = [vp, i32, vp, ctypes.c_size_t, vp]
    L.ts_sort_host.restype = i32
    L.ts_ctx_set_ranking.argtypes = [vp, i32]
    L.ts_ctx_set_ranking.restype = i32
    L.ts_ctx_ranking.argtypes = [vp]
    L.ts_ctx_ranking.restype = i32
    L.ts_ctx_set_split_rounds.argtypes = [vp, i32]
    L.ts_ctx_set_split_rounds.restype = i32
    L.ts_ctx_split_rounds.argtypes = [vp]
    L.ts_ctx_split_rounds.restype = i32
    L.ts_ctx_set_divert_min.argtypes = [vp, sz]
    L.ts_ctx_set_divert_min.restype = i32
    L.ts_ctx_divert_min.argtypes = [vp]
    L.ts_ctx_divert_min.restype = sz
    L.ts_ctx_check.argtypes = [vp, vp]
    L.ts_ctx_check.restype = i32
    L.ts_ctx_last_sort_info.argtypes = [vp, ctypes.POINTER(SortInfo), vp]
    L.ts_ctx_last_sort_info.restype = i32
    L.ts_plan_status_bytes.argtypes = [i32, sz, i32]
    L.ts_plan_status_bytes.restype = sz
    L.ts_verify_sorted.argtypes = [vp, i32, vp, sz, ctypes.POINTER(u64), vp]
    L.ts_verify_sorted.restype = i32
    L.ts_sort_with_buffer_static.argtypes = [vp, i32, vp, sz, vp, sz, vp]
    L.ts_sort_with_buffer_static.restype = i32
    L.ts_ctx_kernel_times.argtypes = [vp, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_float),
                                      ctypes.POINTER(ctypes.c_double), i32]
    L.ts_ctx_kernel_times.restype = i32
    L.ts_sbk_create.argtypes = [ctypes.POINTER(vp), i32]
    L.ts_sbk_create.restype = i32
    L.ts_sbk_destroy.argtypes = [vp]
    L.ts_sbk_destroy.restype = None
    L.ts_sbk_capacity.argtypes = [vp]
    L.ts_sbk_capacity.restype = sz
    L.ts_sbk_order.argtypes = [vp, ctypes.POINTER(sz)]
    L.ts_sbk_order.restype = vp
    L.ts_sbk_sort.argtypes = [vp, i32, vp, sz, sz, sz, i32, vp]
    L.ts_sbk_sort.restype = i32
    L.ts_sbk_apply_order.argtypes = [vp, vp, sz, sz, vp]
    L.ts_sbk_apply_order.restype = i32
    L.ts_sbk_last_error.argtypes = [vp]
    L.ts_sbk_last_error.restype = ctypes.c_char_p
    _lib = L
    return L
