"""ctypes binding of the C-ABI declared in include/acpp_mi355x.h.

Loading fails loudly when the HIP library has not been built: there is no
CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, c_char_p, c_double, c_float, c_int, c_int8, c_int16, c_int32,
                    c_int64, c_size_t, c_uint8, c_uint16, c_uint32, c_uint64, c_void_p)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libacpp_mi355x.so")

# element type tag -> ctypes scalar
INT_TAGS = {
    "i8": c_int8, "i16": c_int16, "i32": c_int32, "i64": c_int64,
    "u8": c_uint8, "u16": c_uint16, "u32": c_uint32, "u64": c_uint64,
}
FLOAT_TAGS = {"f32": c_float, "f64": c_double}
SCALAR_TAGS = {**INT_TAGS, **FLOAT_TAGS}
DOT_TAGS = {"i32": c_int32, "i64": c_int64, "u32": c_uint32, "u64": c_uint64,
            "f32": c_float, "f64": c_double}

SUCCESS = 0
ERROR_INVALID_ARGUMENT = 1
ERROR_UNSUPPORTED = 2
ERROR_OUT_OF_MEMORY = 3
ERROR_HIP = 4
ERROR_WATCHDOG = 5
VARIANT_BLOCK_LOOKBACK = 1
VARIANT_NONTEMPORAL = 2
VARIANT_TILE_PER_GROUP = 4
VARIANT_FORCE_GIVE_UP = 8


class LaunchConfig(ctypes.Structure):
    """acpp_launch_config: zero fields mean "library's choice"."""
    _fields_ = [("work_group_size", c_uint32), ("items_per_thread", c_uint32),
                ("groups_per_cu", c_uint32), ("flags", c_uint32)]


class AcppError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"acpp status {status}: {message}")
        self.status = status


def _signatures() -> dict[str, tuple[list, object]]:
    S = {}
    vp = c_void_p
    cfg = POINTER(LaunchConfig)
    S["acpp_last_error_message"] = ([], c_char_p)
    S["acpp_version"] = ([], c_char_p)
    S["acpp_device_compute_units"] = ([c_int, POINTER(c_int)], c_int)
    S["acpp_allocation_cache_create"] = ([c_int, POINTER(vp)], c_int)
    S["acpp_allocation_cache_purge"] = ([vp], c_int)
    S["acpp_allocation_cache_destroy"] = ([vp], c_int)
    S["acpp_allocation_group_create"] = ([vp, c_int, POINTER(vp)], c_int)
    S["acpp_allocation_group_obtain"] = ([vp, c_size_t, c_size_t, POINTER(vp)], c_int)
    S["acpp_allocation_group_release"] = ([vp], c_int)
    S["acpp_allocation_group_destroy"] = ([vp], c_int)
    S["acpp_allocation_group_watchdog_trips"] = ([vp, POINTER(c_uint64)], c_int)
    S["acpp_check_asynchronous_errors"] = ([c_int], c_int)
    S["acpp_exclusive_scan_i32_noalloc"] = ([vp, vp, c_uint64, vp, c_int32, c_int32, cfg], c_int)
    S["acpp_stdpar_should_offload"] = ([c_int, c_int, c_int, c_int, c_uint64, vp, c_uint32, POINTER(c_int)], c_int)
    S["acpp_stdpar_set_offload_policy"] = ([c_int], c_int)
    S["acpp_stdpar_offload_threshold"] = ([c_int, c_int, c_int, c_int, POINTER(c_uint64)], c_int)
    for tag, ct in SCALAR_TAGS.items():
        S[f"acpp_reduce_{tag}"] = ([vp, vp, vp, c_uint64, vp, ct, c_int32, cfg], c_int)
        S[f"acpp_reduction_{tag}"] = ([vp, vp, vp, c_uint64, vp, c_int32, c_int, cfg], c_int)
        S[f"acpp_reduction_identity_{tag}"] = ([vp, vp, vp, c_uint64, vp, ct, c_int32, c_int, cfg], c_int)
        S[f"acpp_inclusive_scan_{tag}"] = ([vp, vp, vp, c_uint64, vp, c_int32, vp, cfg], c_int)
        S[f"acpp_exclusive_scan_{tag}"] = ([vp, vp, vp, c_uint64, vp, ct, c_int32, cfg], c_int)
    for tag, ct in SCALAR_TAGS.items():
        S[f"acpp_transform_reduce_{tag}"] = ([vp, vp, vp, c_uint64, vp, ct, c_int32, c_int32, cfg], c_int)
        S[f"acpp_transform_inclusive_scan_{tag}"] = (
            [vp, vp, vp, c_uint64, vp, c_int32, c_int32, vp, cfg], c_int)
        S[f"acpp_transform_exclusive_scan_{tag}"] = (
            [vp, vp, vp, c_uint64, vp, ct, c_int32, c_int32, cfg], c_int)
    for tag, ct in DOT_TAGS.items():
        S[f"acpp_transform_reduce_dot_{tag}"] = ([vp, vp, vp, c_uint64, vp, vp, ct, cfg], c_int)
        S[f"acpp_transform_reduce_binary_{tag}"] = (
            [vp, vp, vp, c_uint64, vp, vp, ct, c_int32, c_int32, cfg], c_int)
    S["acpp_reduce_bf16"] = ([vp, vp, vp, c_uint64, vp, c_float, c_int32, cfg], c_int)
    S["acpp_transform_reduce_dot_bf16"] = ([vp, vp, vp, c_uint64, vp, vp, c_float, cfg], c_int)
    for tag in ("f32", "bf16"):
        S[f"acpp_reduce_compensated_{tag}"] = ([vp, vp, vp, c_uint64, vp, c_float, cfg], c_int)
        S[f"acpp_transform_reduce_dot_compensated_{tag}"] = (
            [vp, vp, vp, c_uint64, vp, vp, c_float, cfg], c_int)
    for tag in ("i32", "i64", "f32", "f64"):
        ct = SCALAR_TAGS[tag]
        S[f"acpp_segmented_inclusive_scan_{tag}"] = (
            [vp, vp, vp, c_uint64, vp, vp, c_uint64, c_int32, cfg], c_int)
        S[f"acpp_segmented_reduce_{tag}"] = (
            [vp, vp, vp, c_uint64, vp, c_uint64, vp, ct, c_int32, cfg], c_int)
        S[f"acpp_segmented_plan_inclusive_scan_{tag}"] = (
            [vp, vp, c_uint64, vp, vp, c_uint64, c_int32, cfg, POINTER(vp)], c_int)
        S[f"acpp_segmented_plan_reduce_{tag}"] = (
            [vp, vp, c_uint64, vp, c_uint64, vp, ct, c_int32, cfg, POINTER(vp)], c_int)
    S["acpp_segmented_plan_launch"] = ([vp, vp], c_int)
    S["acpp_segmented_plan_destroy"] = ([vp], c_int)
    S["acpp_histogram_u32"] = ([vp, vp, vp, c_uint64, vp, c_uint32, c_int, cfg], c_int)
    for op in ("add", "min", "max"):
        for tag in ("i32", "u32", "i64", "u64", "f32", "f64"):
            S[f"acpp_atomic_fetch_{op}_strided_{tag}"] = (
                [vp, vp, vp, c_uint64, vp, c_uint32, c_int, cfg], c_int)
    for tag in ("i8", "u8", "i16", "u16", "i32", "u32", "i64", "u64", "f32", "f64"):
        S[f"acpp_reduction_multi_{tag}"] = ([vp, vp, c_uint32, vp, c_uint64, vp, vp, c_int, cfg], c_int)
        S[f"acpp_reduction_multi_identity_{tag}"] = ([vp, vp, c_uint32, vp, c_uint64, vp, vp, vp, vp, cfg],
                                                     c_int)
    for tag in ("i8", "u8", "i16", "u16", "i32", "u32", "i64", "u64", "f32", "f64"):
        ct = SCALAR_TAGS[tag]
        S[f"acpp_copy_if_{tag}"] = ([vp, vp, vp, c_uint64, vp, c_int, ct, ct, vp, cfg], c_int)
        S[f"acpp_copy_if_stencil_{tag}"] = ([vp, vp, vp, c_uint64, vp, vp, vp, cfg], c_int)
    S["acpp_cost_model_create"] = ([c_int, vp, vp, vp, c_double, c_double, POINTER(vp)], c_int)
    S["acpp_cost_model_destroy"] = ([vp], c_int)
    S["acpp_cost_model_choose_segmented"] = ([vp, vp, vp, c_uint64, c_int, c_int, c_int32, POINTER(LaunchConfig),
                                              POINTER(c_float)], c_int)
    S["acpp_mlp_forward"] = ([vp, vp, vp, c_uint32, vp], c_int)
    S["acpp_mlp_gradient"] = ([vp, vp, vp, vp, vp, c_uint32, vp, vp], c_int)
    S["acpp_mlp_adam_step"] = ([vp, vp, vp, vp, vp, c_uint32, vp, vp, c_uint32, c_float, c_float, vp],
                               c_int)
    return S


MLP_INPUTS = 16
MLP_HIDDEN = 64
MLP_PARAMS = 5313

SIGNATURES = _signatures()

_lib = None


def load() -> ctypes.CDLL:
    """Load the HIP library (once) and attach argument types."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: the HIP library has not been built "
            "(run `python -c 'import __graft_entry__ as g; g.build()'` at the repository root)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (argtypes, restype) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != SUCCESS:
        msg = load().acpp_last_error_message().decode(errors="replace")
        raise AcppError(status, msg)
