"""CPU oracle for the device-wide primitives -- TEST INFRASTRUCTURE ONLY.

This package is the checker: only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product library
never calls it.  It wraps ``oracle/liboracle_acpp.so`` (built from
``oracle/oracle.c`` by ``oracle/Makefile``), a C restatement of the
reference's algorithms; each C function cites the reference file:line it
follows.

Pinning status (what checks each function against something other than its
own formula) is listed in the header of tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_double, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle_acpp.so")

_NP_TAG = {
    np.dtype(np.int8): "i8", np.dtype(np.int16): "i16", np.dtype(np.int32): "i32",
    np.dtype(np.int64): "i64", np.dtype(np.uint8): "u8", np.dtype(np.uint16): "u16",
    np.dtype(np.uint32): "u32", np.dtype(np.uint64): "u64", np.dtype(np.float32): "f32",
    np.dtype(np.float64): "f64",
}
_CT = {
    "i8": ctypes.c_int8, "i16": ctypes.c_int16, "i32": ctypes.c_int32, "i64": ctypes.c_int64,
    "u8": ctypes.c_uint8, "u16": ctypes.c_uint16, "u32": ctypes.c_uint32, "u64": ctypes.c_uint64,
    "f32": ctypes.c_float, "f64": ctypes.c_double,
}

# operator codes (== __acpp_sscp_algorithm_op)
PLUS, MULTIPLY, MIN, MAX, BIT_AND, BIT_OR, BIT_XOR, LOGICAL_AND, LOGICAL_OR = range(9)
# atomic rmw codes (StdAtomicRemapperPass.cpp:54)
RMW_AND, RMW_OR, RMW_XOR, RMW_ADD, RMW_SUB, RMW_MIN, RMW_MAX = range(7)

_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing; build it with `make -C oracle`")
        lib = ctypes.CDLL(LIB_PATH)
        for n in ("oracle_sum_f32_fp64", "oracle_sum_bf16_fp64", "oracle_dot_f32_fp64",
                  "oracle_dot_bf16_fp64", "oracle_abs_sum_f32_fp64"):
            getattr(lib, n).restype = c_double
        _lib = lib
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data if a is not None else 0


def _tag(a: np.ndarray) -> str:
    return _NP_TAG[a.dtype]


def _ok(rc: int) -> None:
    if rc != 0:
        raise ValueError("oracle rejected the call (invalid operator?)")


def set_fault(k: int) -> None:
    load().oracle_set_fault(c_int(k))


def max_threads() -> int:
    return load().oracle_max_threads()


def num_procs() -> int:
    """omp_get_num_procs(): the host device's max_compute_units in the reference."""
    return load().oracle_num_procs()


def identity(dtype, op: int):
    t = _NP_TAG[np.dtype(dtype)]
    v = _CT[t]()
    _ok(getattr(load(), f"oracle_identity_{t}")(c_int(op), ctypes.byref(v)))
    return np.dtype(dtype).type(v.value)


def reduce(x: np.ndarray, init, op: int = PLUS):
    x = np.ascontiguousarray(x)
    t = _tag(x)
    out = _CT[t]()
    _ok(getattr(load(), f"oracle_reduce_{t}")(c_void_p(_p(x)), c_uint64(x.size), _CT[t](init),
                                              c_int(op), ctypes.byref(out)))
    return x.dtype.type(out.value)


def reduction(x: np.ndarray, out_value, op: int, initialize_to_identity: bool):
    x = np.ascontiguousarray(x)
    t = _tag(x)
    out = _CT[t](out_value)
    _ok(getattr(load(), f"oracle_reduction_{t}")(c_void_p(_p(x)), c_uint64(x.size),
                                                 ctypes.byref(out), c_int(op),
                                                 c_int(1 if initialize_to_identity else 0)))
    return x.dtype.type(out.value)


def inclusive_scan(x: np.ndarray, op: int = PLUS, init=None) -> np.ndarray:
    x = np.ascontiguousarray(x)
    t = _tag(x)
    y = np.empty_like(x)
    init_c = _CT[t](init) if init is not None else None
    _ok(getattr(load(), f"oracle_inclusive_scan_{t}")(
        c_void_p(_p(x)), c_uint64(x.size), c_void_p(_p(y)), c_int(op),
        ctypes.byref(init_c) if init_c is not None else None))
    return y


def exclusive_scan(x: np.ndarray, init, op: int = PLUS) -> np.ndarray:
    x = np.ascontiguousarray(x)
    t = _tag(x)
    y = np.empty_like(x)
    _ok(getattr(load(), f"oracle_exclusive_scan_{t}")(c_void_p(_p(x)), c_uint64(x.size),
                                                      c_void_p(_p(y)), _CT[t](init), c_int(op)))
    return y


def dot(a: np.ndarray, b: np.ndarray, init):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    t = _tag(a)
    out = _CT[t]()
    _ok(getattr(load(), f"oracle_dot_{t}")(c_void_p(_p(a)), c_void_p(_p(b)), c_uint64(a.size),
                                           _CT[t](init), ctypes.byref(out)))
    return a.dtype.type(out.value)


def reduce_omp_order(x: np.ndarray, init, op: int = PLUS, threads: int = 0):
    x = np.ascontiguousarray(x)
    t = _tag(x)
    out = _CT[t](init)
    _ok(getattr(load(), f"oracle_reduce_omp_order_{t}")(
        c_void_p(_p(x)), c_uint64(x.size), _CT[t](init), c_int(op), ctypes.byref(out),
        c_int(threads)))
    return x.dtype.type(out.value)


def dot_omp_order(a: np.ndarray, b: np.ndarray, init, threads: int = 0):
    t = _tag(a)
    out = _CT[t](init)
    _ok(getattr(load(), f"oracle_dot_omp_order_{t}")(
        c_void_p(_p(a)), c_void_p(_p(b)), c_uint64(a.size), _CT[t](init), ctypes.byref(out),
        c_int(threads)))
    return a.dtype.type(out.value)


def scan_omp_order(x: np.ndarray, exclusive: bool, init=None, op: int = PLUS,
                   threads: int = 0, out: np.ndarray | None = None) -> np.ndarray:
    x = np.ascontiguousarray(x)
    t = _tag(x)
    y = np.empty_like(x) if out is None else out
    has_init = init is not None
    _ok(getattr(load(), f"oracle_scan_omp_order_{t}")(
        c_void_p(_p(x)), c_uint64(x.size), c_void_p(_p(y)), c_int(1 if exclusive else 0),
        c_int(1 if has_init else 0), _CT[t](init if has_init else 0), c_int(op), c_int(threads)))
    return y


def group_reduce(x: np.ndarray, local_size: int, op: int = PLUS) -> np.ndarray:
    x = np.ascontiguousarray(x)
    t = _tag(x)
    y = np.empty_like(x)
    _ok(getattr(load(), f"oracle_group_reduce_{t}")(c_void_p(_p(x)), c_uint64(x.size),
                                                    c_uint64(local_size), c_int(op),
                                                    c_void_p(_p(y))))
    return y


def group_inclusive_scan(x: np.ndarray, local_size: int, op: int = PLUS) -> np.ndarray:
    x = np.ascontiguousarray(x)
    t = _tag(x)
    y = np.empty_like(x)
    _ok(getattr(load(), f"oracle_group_inclusive_scan_{t}")(c_void_p(_p(x)), c_uint64(x.size),
                                                            c_uint64(local_size), c_int(op),
                                                            c_void_p(_p(y))))
    return y


def group_exclusive_scan(x: np.ndarray, local_size: int, init, op: int = PLUS) -> np.ndarray:
    x = np.ascontiguousarray(x)
    t = _tag(x)
    y = np.empty_like(x)
    _ok(getattr(load(), f"oracle_group_exclusive_scan_{t}")(
        c_void_p(_p(x)), c_uint64(x.size), c_uint64(local_size), _CT[t](init), c_int(op),
        c_void_p(_p(y))))
    return y


def atomic_fold(rmw_op: int, operands: np.ndarray, initial):
    operands = np.ascontiguousarray(operands)
    t = _tag(operands)
    out = _CT[t]()
    _ok(getattr(load(), f"oracle_atomic_fold_{t}")(c_int(rmw_op), c_void_p(_p(operands)),
                                                   c_uint64(operands.size), _CT[t](initial),
                                                   ctypes.byref(out)))
    return operands.dtype.type(out.value)


def segmented_inclusive_scan(x: np.ndarray, offsets: np.ndarray, op: int = PLUS) -> np.ndarray:
    x = np.ascontiguousarray(x)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    t = _tag(x)
    y = np.zeros_like(x)
    _ok(getattr(load(), f"oracle_segmented_inclusive_scan_{t}")(
        c_void_p(_p(x)), c_void_p(_p(offsets)), c_uint64(offsets.size - 1), c_void_p(_p(y)),
        c_int(op)))
    return y


def segmented_reduce(x: np.ndarray, offsets: np.ndarray, init, op: int = PLUS) -> np.ndarray:
    x = np.ascontiguousarray(x)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    t = _tag(x)
    out = np.empty(offsets.size - 1, dtype=x.dtype)
    _ok(getattr(load(), f"oracle_segmented_reduce_{t}")(
        c_void_p(_p(x)), c_void_p(_p(offsets)), c_uint64(offsets.size - 1), _CT[t](init),
        c_int(op), c_void_p(_p(out))))
    return out


def segmented_reduce_f32_fp64(x: np.ndarray, offsets: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    out = np.empty(offsets.size - 1, dtype=np.float64)
    _ok(load().oracle_segmented_reduce_f32_fp64(c_void_p(_p(x)), c_void_p(_p(offsets)),
                                                c_uint64(offsets.size - 1), c_void_p(_p(out))))
    return out


def sum_fp64(x: np.ndarray) -> float:
    """FP64 oracle of a float32 array or of bf16 bit patterns (uint16)."""
    x = np.ascontiguousarray(x)
    if x.dtype == np.float32:
        return load().oracle_sum_f32_fp64(c_void_p(_p(x)), c_uint64(x.size))
    if x.dtype == np.uint16:
        return load().oracle_sum_bf16_fp64(c_void_p(_p(x)), c_uint64(x.size))
    raise TypeError("sum_fp64 takes float32 or bf16 bit patterns (uint16)")


def abs_sum_fp64(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, dtype=np.float32)
    return load().oracle_abs_sum_f32_fp64(c_void_p(_p(x)), c_uint64(x.size))


def dot_fp64(a: np.ndarray, b: np.ndarray) -> float:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype == np.float32:
        return load().oracle_dot_f32_fp64(c_void_p(_p(a)), c_void_p(_p(b)), c_uint64(a.size))
    if a.dtype == np.uint16:
        return load().oracle_dot_bf16_fp64(c_void_p(_p(a)), c_void_p(_p(b)), c_uint64(a.size))
    raise TypeError("dot_fp64 takes float32 or bf16 bit patterns (uint16)")


def histogram(keys: np.ndarray, num_bins: int, bins: np.ndarray | None = None,
              threads: int = 0) -> np.ndarray:
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    b = np.zeros(num_bins, dtype=np.uint32) if bins is None else bins.copy()
    _ok(load().oracle_histogram_u32(c_void_p(_p(keys)), c_uint64(keys.size), c_void_p(_p(b)),
                                    c_uint32(num_bins), c_int(threads)))
    return b


def fetch_max_strided(values: np.ndarray, num_slots: int, slots: np.ndarray | None = None,
                      threads: int = 0) -> np.ndarray:
    values = np.ascontiguousarray(values, dtype=np.int32)
    s = np.full(num_slots, np.iinfo(np.int32).min, dtype=np.int32) if slots is None \
        else slots.copy()
    _ok(load().oracle_fetch_max_strided_i32(c_void_p(_p(values)), c_uint64(values.size),
                                            c_void_p(_p(s)), c_uint32(num_slots), c_int(threads)))
    return s


# copy_if predicate codes (== ACPP_PRED_*)
PRED_NONZERO, PRED_EQ, PRED_NE, PRED_LT, PRED_LE, PRED_GT, PRED_GE, PRED_MASK_EQ, PRED_STENCIL = range(9)


def copy_if(x: np.ndarray, pred: int, operand=0, mask=0, stencil: np.ndarray | None = None):
    """(selected elements, count) -- oracle_copy_if_<t> (algorithm.hpp:178-243)."""
    x = np.ascontiguousarray(x)
    tag = _tag(x)
    ct = _CT[tag]
    out = np.zeros(max(x.size, 1), dtype=x.dtype)
    cnt = ctypes.c_uint64(0)
    st = None if stencil is None else np.ascontiguousarray(stencil, dtype=np.uint8)
    fn = getattr(load(), f"oracle_copy_if_{tag}")
    _ok(fn(c_void_p(_p(x)), c_uint64(x.size), c_void_p(_p(st) if st is not None else 0), c_int(pred),
           ct(operand), ct(mask), c_void_p(_p(out)), ctypes.byref(cnt)))
    return out[:cnt.value].copy(), int(cnt.value)


# element transform codes (== acpp_unary_op)
IDENTITY, NEGATE, ABS, SQUARE, TWICE, BIT_NOT, IS_ZERO, IS_NONZERO = range(8)


def transform(x: np.ndarray, code: int) -> np.ndarray:
    """y[i] = t(x[i]) for a transform code; integer results wrap."""
    x = np.ascontiguousarray(x)
    t = _tag(x)
    y = np.empty_like(x)
    _ok(getattr(load(), f"oracle_transform_{t}")(c_void_p(_p(x)), c_uint64(x.size),
                                                 c_void_p(_p(y)), c_int(code)))
    return y


def transform_reduce(x: np.ndarray, init, op: int, code: int):
    """Unary transform_reduce (numeric.hpp:193-212): reduce over t(x[i])."""
    return reduce(transform(x, code), init, op)


def transform_inclusive_scan(x: np.ndarray, op: int, code: int, init=None) -> np.ndarray:
    """transform_inclusive_scan (numeric.hpp:299-317; scan.hpp:85-132)."""
    return inclusive_scan(transform(x, code), op, init)


def transform_exclusive_scan(x: np.ndarray, init, op: int, code: int) -> np.ndarray:
    """transform_exclusive_scan (numeric.hpp:319-327; scan.hpp:85-132)."""
    return exclusive_scan(transform(x, code), init, op)


# pair transform codes (== acpp_binary_op)
B_MULTIPLIES, B_PLUS, B_MINUS, B_ABS_DIFF, B_SQ_DIFF, B_MIN, B_MAX, B_BIT_AND, B_BIT_OR, B_BIT_XOR = range(10)


def transform_reduce_binary(a: np.ndarray, b: np.ndarray, init, op: int, code: int):
    """Binary transform_reduce (numeric.hpp:170-191): serial fold of
    op over t(a[i], b[i]) starting from init."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    t = _tag(a)
    out = _CT[t]()
    _ok(getattr(load(), f"oracle_transform_reduce_binary_{t}")(
        c_void_p(_p(a)), c_void_p(_p(b)), c_uint64(a.size), _CT[t](init), c_int(op), c_int(code),
        ctypes.byref(out)))
    return a.dtype.type(out.value)


# group data movement operation codes (oracle_group_move_<t>)
MOVE_BROADCAST, MOVE_SHIFT_LEFT, MOVE_SHIFT_RIGHT, MOVE_PERMUTE_XOR, MOVE_SELECT, MOVE_ANY, MOVE_ALL, \
    MOVE_NONE, MOVE_JOINT_ANY, MOVE_JOINT_ALL, MOVE_JOINT_NONE = range(11)


def group_move(level: int, which: int, param: int, x: np.ndarray, src: np.ndarray | None,
               local_size: int) -> np.ndarray:
    """Group data movement (group_functions.hpp:57-140, 240-276) within
    sub-groups of 64 (level 0) or work-groups (level 1)."""
    x = np.ascontiguousarray(x)
    t = _tag(x)
    s = np.zeros(x.size, dtype=np.uint32) if src is None else np.ascontiguousarray(src, dtype=np.uint32)
    out = np.empty_like(x)
    _ok(getattr(load(), f"oracle_group_move_{t}")(c_int(level), c_int(which), ctypes.c_int64(param),
                                                  c_void_p(_p(x)), c_void_p(_p(s)), c_uint64(x.size),
                                                  c_uint64(local_size), c_void_p(_p(out))))
    return out
