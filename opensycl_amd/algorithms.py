"""Python mirror of ``acpp::algorithms`` (include/hipSYCL/algorithms/numeric.hpp).

Every function has the reference's argument order -- queue, scratch
allocation group, input range, output, init, operator -- with torch device
tensors standing for the iterator ranges.  Calls are asynchronous on the
queue's stream, like the reference's (doc/algorithms.md:35-37); use
``queue.wait()`` before reading results on the host.

The device work is done by the HIP library through its C-ABI
(``opensycl_amd._abi``); nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes
from typing import Optional

from . import _abi
from ._abi import LaunchConfig, check

# Operator codes (== __acpp_sscp_algorithm_op, builtin_config.hpp:27-37)
PLUS = 0
MULTIPLY = 1
MIN = 2
MAX = 3
BIT_AND = 4
BIT_OR = 5
BIT_XOR = 6
LOGICAL_AND = 7
LOGICAL_OR = 8

_OP_NAMES = {
    "plus": PLUS, "multiplies": MULTIPLY, "multiply": MULTIPLY, "minimum": MIN, "min": MIN,
    "maximum": MAX, "max": MAX, "bit_and": BIT_AND, "bit_or": BIT_OR, "bit_xor": BIT_XOR,
    "logical_and": LOGICAL_AND, "logical_or": LOGICAL_OR,
}

# Element transform codes (acpp_unary_op): the unary callables of
# transform_reduce and the transform scans (numeric.hpp:193-212, 299-327).
IDENTITY = 0
NEGATE = 1
ABS = 2
SQUARE = 3
TWICE = 4
BIT_NOT = 5
IS_ZERO = 6
IS_NONZERO = 7

_UNARY_NAMES = {
    "identity": IDENTITY, "negate": NEGATE, "abs": ABS, "square": SQUARE, "twice": TWICE,
    "bit_not": BIT_NOT, "is_zero": IS_ZERO, "is_nonzero": IS_NONZERO,
}

ALLOCATION_DEVICE = 0
ALLOCATION_SHARED = 1
ALLOCATION_HOST = 2


def _torch():
    import torch
    return torch


def _tag(dtype) -> str:
    torch = _torch()
    table = {
        torch.int8: "i8", torch.int16: "i16", torch.int32: "i32", torch.int64: "i64",
        torch.uint8: "u8", torch.uint16: "u16", torch.uint32: "u32", torch.uint64: "u64",
        torch.float32: "f32", torch.float64: "f64", torch.bfloat16: "bf16",
    }
    if dtype not in table:
        raise TypeError(f"unsupported element type {dtype}")
    return table[dtype]


def _op(op) -> int:
    if isinstance(op, str):
        if op not in _OP_NAMES:
            raise ValueError(f"unknown operator {op!r}")
        return _OP_NAMES[op]
    return int(op)


def _unary(transform) -> int:
    if isinstance(transform, str):
        if transform not in _UNARY_NAMES:
            raise ValueError(f"unknown element transform {transform!r}")
        return _UNARY_NAMES[transform]
    return int(transform)


def _cfg(config):
    if config is None:
        return None
    if isinstance(config, LaunchConfig):
        return ctypes.byref(config)
    if isinstance(config, dict):
        return ctypes.byref(LaunchConfig(**config))
    wg, ipt, gpc, flags = (tuple(config) + (0, 0, 0, 0))[:4]
    return ctypes.byref(LaunchConfig(wg, ipt, gpc, flags))


class Queue:
    """In-order queue on a HIP stream (torch stream or raw hipStream_t)."""

    def __init__(self, stream=None, device: int = 0):
        torch = _torch()
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device)
        self.stream = stream

    @property
    def handle(self) -> int:
        s = self.stream
        if isinstance(s, int):
            return s
        return s.cuda_stream

    def wait(self) -> None:
        """Wait for the queue's work, then raise the device's asynchronous
        errors (queue::wait_and_throw, include/hipSYCL/sycl/queue.hpp:
        378-383): AcppError with status ERROR_WATCHDOG if a kernel gave up
        waiting for another work-group, whose results are then wrong."""
        if isinstance(self.stream, int):
            _torch().cuda.synchronize(self.device)
        else:
            self.stream.synchronize()
        check_asynchronous_errors(self.device)


def check_asynchronous_errors(device: int = 0) -> None:
    """acpp_check_asynchronous_errors: raise (once) for a watchdog give-up
    on `device` since the last check.  Non-blocking: synchronise first."""
    check(_abi.load().acpp_check_asynchronous_errors(device))


class AllocationCache:
    """util::allocation_cache (allocation_cache.hpp:42-127)."""

    def __init__(self, allocation_type: int = ALLOCATION_DEVICE):
        lib = _abi.load()
        h = ctypes.c_void_p()
        check(lib.acpp_allocation_cache_create(allocation_type, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def purge(self) -> None:
        check(_abi.load().acpp_allocation_cache_purge(self._h))

    def close(self) -> None:
        if self._h:
            h, self._h = self._h, None
            check(_abi.load().acpp_allocation_cache_destroy(h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AllocationGroup:
    """util::allocation_group (allocation_cache.hpp:142-181)."""

    def __init__(self, cache: AllocationCache, device: int = 0):
        lib = _abi.load()
        h = ctypes.c_void_p()
        check(lib.acpp_allocation_group_create(cache.handle, device, ctypes.byref(h)))
        self._h = h
        self._cache = cache
        self._plans = 0             # live SegmentedPlans captured from this group
        self._close_pending = False

    @property
    def handle(self):
        return self._h

    def obtain(self, nbytes: int, alignment: int = 256) -> int:
        p = ctypes.c_void_p()
        check(_abi.load().acpp_allocation_group_obtain(self._h, nbytes, alignment,
                                                       ctypes.byref(p)))
        return p.value

    def release(self) -> None:
        """Return the group's regions to the cache now.  Refused while a
        SegmentedPlan captured from this group is alive: its graph still
        uses them, and replays would corrupt whichever call got them next."""
        if self._plans > 0:
            raise RuntimeError("AllocationGroup.release: a SegmentedPlan captured from this group is alive; "
                               "close the plan first")
        check(_abi.load().acpp_allocation_group_release(self._h))

    def watchdog_trips(self) -> int:
        """Look-backs that gave up waiting in the scratch regions this group
        holds (blocking; call after the group's work has completed).  A
        non-zero count means some results are wrong; expected to be 0."""
        v = ctypes.c_uint64()
        check(_abi.load().acpp_allocation_group_watchdog_trips(self._h, ctypes.byref(v)))
        return v.value

    def close(self) -> None:
        """Return the group's regions to the cache.  While a SegmentedPlan
        captured from this group is alive, its graph still uses them: the
        release then waits until the last such plan is closed."""
        if self._plans > 0:
            self._close_pending = True
            return
        if self._h:
            h, self._h = self._h, None
            # destroyed even when it reports an asynchronous error
            check(_abi.load().acpp_allocation_group_destroy(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compute_units(device: int = 0) -> int:
    v = ctypes.c_int()
    check(_abi.load().acpp_device_compute_units(device, ctypes.byref(v)))
    return v.value


def _ptr(t) -> int:
    if t is None:
        return 0
    if not t.is_cuda:
        raise ValueError("tensors must live in device memory")
    return t.data_ptr()


def _check_contig(*ts):
    for t in ts:
        if t is not None and not t.is_contiguous():
            raise ValueError("tensors must be contiguous")


def _check_out(out, dtype, count: int, what: str) -> None:
    """The C-ABI receives raw pointers only, so an output too short or of a
    narrower type would be written past its end: reject it before launch."""
    if out is None:
        raise ValueError(f"{what}: output is missing")
    if out.dtype != dtype:
        raise TypeError(f"{what}: output must be {dtype}, not {out.dtype}")
    if out.numel() < count:
        raise ValueError(f"{what}: output holds {out.numel()} elements, needs {count}")


def _acc_dtype(x):
    """Result type of a reduction over x: float32 for bf16 input."""
    torch = _torch()
    return torch.float32 if x.dtype == torch.bfloat16 else x.dtype


def _scalar(tag: str, v):
    ct = _abi.SCALAR_TAGS.get(tag, ctypes.c_float)
    return ct(v)


def reduce(q: Queue, scratch: AllocationGroup, x, out, init=0, op=PLUS, config=None) -> None:
    """algorithms::reduce (numeric.hpp:226-249): out[0] = op(init, fold(x))."""
    _check_contig(x, out)
    _check_out(out, _acc_dtype(x), 1, "reduce")
    tag = _tag(x.dtype)
    fn = getattr(_abi.load(), f"acpp_reduce_{tag}")
    check(fn(scratch.handle, q.handle, _ptr(x), x.numel(), _ptr(out), _scalar(tag, init),
             _op(op), _cfg(config)))


def reduce_compensated(q: Queue, scratch: AllocationGroup, x, out, init=0.0,
                       config=None) -> None:
    """Compensated fp32 sum of fp32 or bf16 input into an fp32 out[0]."""
    _check_contig(x, out)
    tag = _tag(x.dtype)
    if tag not in ("f32", "bf16"):
        raise TypeError("compensated sums take float32 or bfloat16 input")
    _check_out(out, _torch().float32, 1, "reduce_compensated")
    fn = getattr(_abi.load(), f"acpp_reduce_compensated_{tag}")
    check(fn(scratch.handle, q.handle, _ptr(x), x.numel(), _ptr(out), ctypes.c_float(init),
             _cfg(config)))


# Pair transform codes of the binary transform_reduce (== acpp_binary_op)
B_MULTIPLIES, B_PLUS, B_MINUS, B_ABS_DIFF, B_SQ_DIFF, B_MIN, B_MAX, B_BIT_AND, B_BIT_OR, B_BIT_XOR = range(10)
_BINARY_NAMES = {
    "multiplies": B_MULTIPLIES, "plus": B_PLUS, "minus": B_MINUS, "abs_diff": B_ABS_DIFF,
    "sq_diff": B_SQ_DIFF, "minimum": B_MIN, "maximum": B_MAX, "bit_and": B_BIT_AND, "bit_or": B_BIT_OR,
    "bit_xor": B_BIT_XOR,
}


def transform_reduce(q: Queue, scratch: AllocationGroup, x1, x2, out, init=0,
                     compensated: bool = False, config=None, reduce_op=PLUS,
                     transform=B_MULTIPLIES) -> None:
    """Binary transform_reduce (numeric.hpp:170-223): out[0] = reduce_op(init,
    reduce_op over transform(x1[i], x2[i])).  The default (plus, multiplies)
    is the dot product, also for bf16 and in the compensated form; other
    pairs (reduce_op: plus, multiplies, minimum, maximum; transform: a pair
    transform code or name) take 32/64-bit integers, float32 and float64."""
    _check_contig(x1, x2, out)
    if x1.numel() > x2.numel() or x1.dtype != x2.dtype:
        raise ValueError("second range must have the first range's type and length")
    tag = _tag(x1.dtype)
    _check_out(out, _torch().float32 if compensated else _acc_dtype(x1), 1, "transform_reduce")
    code = _BINARY_NAMES[transform] if isinstance(transform, str) else int(transform)
    op = _op(reduce_op)
    lib = _abi.load()
    if (op, code) != (PLUS, B_MULTIPLIES):
        if compensated or tag == "bf16":
            raise TypeError("compensated and bf16 transform_reduce take plus with multiplies only")
        fn = getattr(lib, f"acpp_transform_reduce_binary_{tag}")
        check(fn(scratch.handle, q.handle, _ptr(x1), x1.numel(), _ptr(x2), _ptr(out), _scalar(tag, init), op,
                 code, _cfg(config)))
        return
    if compensated:
        if tag not in ("f32", "bf16"):
            raise TypeError("compensated dot products take float32 or bfloat16 input")
        fn = getattr(lib, f"acpp_transform_reduce_dot_compensated_{tag}")
        init_c = ctypes.c_float(init)
    else:
        fn = getattr(lib, f"acpp_transform_reduce_dot_{tag}")
        init_c = ctypes.c_float(init) if tag == "bf16" else _scalar(tag, init)
    check(fn(scratch.handle, q.handle, _ptr(x1), x1.numel(), _ptr(x2), _ptr(out), init_c,
             _cfg(config)))


def transform_reduce_unary(q: Queue, scratch: AllocationGroup, x, out, init=0, op=PLUS,
                           transform=IDENTITY, config=None) -> None:
    """Unary transform_reduce (numeric.hpp:193-212):
    out[0] = op(init, fold_op(transform(x[i]))); empty x leaves out as is."""
    _check_contig(x, out)
    _check_out(out, x.dtype, 1, "transform_reduce")
    tag = _tag(x.dtype)
    fn = getattr(_abi.load(), f"acpp_transform_reduce_{tag}")
    check(fn(scratch.handle, q.handle, _ptr(x), x.numel(), _ptr(out), _scalar(tag, init),
             _op(op), _unary(transform), _cfg(config)))


def reduction(q: Queue, scratch: AllocationGroup, x, out, op=PLUS,
              initialize_to_identity: bool = False, config=None, identity=None) -> None:
    """parallel_for(range{n}, sycl::reduction(out, [identity,] op
    [, initialize_to_identity])) whose kernel combines x[i]
    (reduction.hpp:165-236): with initialize_to_identity out[0] = op(identity,
    fold), else op(out[0], fold).  Without a given identity the operator's
    own is used."""
    _check_contig(x, out)
    _check_out(out, x.dtype, 1, "reduction")
    tag = _tag(x.dtype)
    lib = _abi.load()
    flag = 1 if initialize_to_identity else 0
    if identity is None:
        check(getattr(lib, f"acpp_reduction_{tag}")(scratch.handle, q.handle, _ptr(x), x.numel(), _ptr(out),
                                                    _op(op), flag, _cfg(config)))
    else:
        check(getattr(lib, f"acpp_reduction_identity_{tag}")(scratch.handle, q.handle, _ptr(x), x.numel(),
                                                             _ptr(out), _scalar(tag, identity), _op(op),
                                                             flag, _cfg(config)))


def reduction_multi(q: Queue, scratch: AllocationGroup, inputs, outs, ops,
                    initialize_to_identity=False, config=None, identities=None) -> None:
    """Up to four sycl::reduction objects in one parallel_for
    (handler.hpp:799-975): outs[k][0] = fold of inputs[k] with ops[k]
    (starting from the identity, or from outs[k][0]).  Inputs share one
    length and element type; they may be the same tensor.
    initialize_to_identity: one flag for all, or one per reduction;
    identities: None, or one given identity per reduction."""
    k = len(inputs)
    if not (1 <= k <= 4) or len(outs) != k or len(ops) != k:
        raise ValueError("one to four (input, output, operator) triples")
    n = inputs[0].numel()
    tag = _tag(inputs[0].dtype)
    for x, o in zip(inputs, outs):
        _check_contig(x, o)
        if x.numel() != n or _tag(x.dtype) != tag or _tag(o.dtype) != tag:
            raise ValueError("inputs must share length and element type with the outputs")
        _check_out(o, x.dtype, 1, "reduction_multi")
    arr_in = (ctypes.c_void_p * k)(*[_ptr(x) for x in inputs])
    arr_out = (ctypes.c_void_p * k)(*[_ptr(o) for o in outs])
    arr_ops = (ctypes.c_int32 * k)(*[_op(o) for o in ops])
    lib = _abi.load()
    if identities is None and isinstance(initialize_to_identity, bool):
        fn = getattr(lib, f"acpp_reduction_multi_{tag}")
        check(fn(scratch.handle, q.handle, k, arr_in, n, arr_out, arr_ops, int(initialize_to_identity),
                 _cfg(config)))
        return
    flags = list(initialize_to_identity) if not isinstance(initialize_to_identity, bool) \
        else [initialize_to_identity] * k
    if len(flags) != k or (identities is not None and len(identities) != k):
        raise ValueError("one initialize_to_identity flag and identity per reduction")
    arr_flags = (ctypes.c_int * k)(*[int(bool(f)) for f in flags])
    arr_id = None
    if identities is not None:
        ct = _abi.SCALAR_TAGS[tag]
        arr_id = (ct * k)(*identities)
    fn = getattr(lib, f"acpp_reduction_multi_identity_{tag}")
    check(fn(scratch.handle, q.handle, k, arr_in, n, arr_out, arr_ops, arr_id, arr_flags, _cfg(config)))


def inclusive_scan(q: Queue, scratch: AllocationGroup, x, d, op=PLUS, init=None,
                   config=None) -> None:
    """algorithms::inclusive_scan (numeric.hpp:253-279)."""
    _check_contig(x, d)
    _check_out(d, x.dtype, x.numel(), "inclusive_scan")
    tag = _tag(x.dtype)
    fn = getattr(_abi.load(), f"acpp_inclusive_scan_{tag}")
    init_p = None
    if init is not None:
        init_c = _scalar(tag, init)
        init_p = ctypes.addressof(init_c)
    check(fn(scratch.handle, q.handle, _ptr(x), x.numel(), _ptr(d), _op(op), init_p,
             _cfg(config)))


def exclusive_scan(q: Queue, scratch: AllocationGroup, x, d, init=0, op=PLUS,
                   config=None) -> None:
    """algorithms::exclusive_scan (numeric.hpp:281-297)."""
    _check_contig(x, d)
    _check_out(d, x.dtype, x.numel(), "exclusive_scan")
    tag = _tag(x.dtype)
    fn = getattr(_abi.load(), f"acpp_exclusive_scan_{tag}")
    check(fn(scratch.handle, q.handle, _ptr(x), x.numel(), _ptr(d), _scalar(tag, init),
             _op(op), _cfg(config)))


def transform_inclusive_scan(q: Queue, scratch: AllocationGroup, x, d, op=PLUS,
                             transform=IDENTITY, init=None, config=None) -> None:
    """algorithms::transform_inclusive_scan (numeric.hpp:299-317): the
    inclusive scan of transform(x[i])."""
    _check_contig(x, d)
    _check_out(d, x.dtype, x.numel(), "transform_inclusive_scan")
    tag = _tag(x.dtype)
    fn = getattr(_abi.load(), f"acpp_transform_inclusive_scan_{tag}")
    init_p = None
    if init is not None:
        init_c = _scalar(tag, init)
        init_p = ctypes.addressof(init_c)
    check(fn(scratch.handle, q.handle, _ptr(x), x.numel(), _ptr(d), _op(op), _unary(transform),
             init_p, _cfg(config)))


def transform_exclusive_scan(q: Queue, scratch: AllocationGroup, x, d, init=0, op=PLUS,
                             transform=IDENTITY, config=None) -> None:
    """algorithms::transform_exclusive_scan (numeric.hpp:319-327): the
    exclusive scan of transform(x[i]) starting from init."""
    _check_contig(x, d)
    _check_out(d, x.dtype, x.numel(), "transform_exclusive_scan")
    tag = _tag(x.dtype)
    fn = getattr(_abi.load(), f"acpp_transform_exclusive_scan_{tag}")
    check(fn(scratch.handle, q.handle, _ptr(x), x.numel(), _ptr(d), _scalar(tag, init),
             _op(op), _unary(transform), _cfg(config)))


# ---------------------------------------------------------------------------
# Bulk atomics (atomic_ref fetch_add / fetch_min / fetch_max in a loop)
# ---------------------------------------------------------------------------
ATOMICS_AUTO = 0
ATOMICS_PRIVATIZED = 1
ATOMICS_GLOBAL = 2


def histogram(q: Queue, scratch: AllocationGroup, keys, bins, strategy: int = ATOMICS_AUTO,
              config=None) -> None:
    """bins[k] += number of keys equal to k (keys >= len(bins) are ignored)."""
    _check_contig(keys, bins)
    if _tag(keys.dtype) != "u32" or _tag(bins.dtype) != "u32":
        raise TypeError("histogram takes uint32 keys and uint32 bins")
    check(_abi.load().acpp_histogram_u32(scratch.handle, q.handle, _ptr(keys), keys.numel(),
                                         _ptr(bins), bins.numel(), strategy, _cfg(config)))


def atomic_fetch_strided(q: Queue, scratch: AllocationGroup, op: str, values, target,
                         strategy: int = ATOMICS_AUTO, config=None) -> None:
    """target[i % len(target)] = op(target[i % len(target)], values[i]) for op in
    add / min / max."""
    _check_contig(values, target)
    if op not in ("add", "min", "max"):
        raise ValueError("op must be add, min or max")
    tag = _tag(values.dtype)
    if values.dtype != target.dtype:
        raise TypeError("values and target must have the same type")
    if target.numel() < 1:
        raise ValueError("target must hold at least one slot")
    fn = getattr(_abi.load(), f"acpp_atomic_fetch_{op}_strided_{tag}")
    check(fn(scratch.handle, q.handle, _ptr(values), values.numel(), _ptr(target), target.numel(),
             strategy, _cfg(config)))


# ---------------------------------------------------------------------------
# copy_if (stream compaction)
# ---------------------------------------------------------------------------
PRED_NONZERO, PRED_EQ, PRED_NE, PRED_LT, PRED_LE, PRED_GT, PRED_GE, PRED_MASK_EQ = range(8)


_COPY_IF_TAGS = ("i8", "u8", "i16", "u16", "i32", "u32", "i64", "u64", "f32", "f64")


def copy_if(q: Queue, scratch: AllocationGroup, x, d, pred: int, operand=0, mask=0, num_copied=None,
            config=None) -> None:
    """algorithms::copy_if (algorithm.hpp:178-243) with a data predicate:
    d[0..count) = the x[i] with pred(x[i]); num_copied[0] = count (a uint64
    device tensor, optional)."""
    _check_contig(x, d)
    _check_out(d, x.dtype, x.numel(), "copy_if")
    if num_copied is not None:
        _check_out(num_copied, _torch().uint64, 1, "copy_if num_copied")
    tag = _tag(x.dtype)
    if tag not in _COPY_IF_TAGS:
        raise TypeError(f"copy_if takes integer, f32 or f64 elements, not {tag}")
    fn = getattr(_abi.load(), f"acpp_copy_if_{tag}")
    check(fn(scratch.handle, q.handle, _ptr(x), x.numel(), _ptr(d), int(pred), _scalar(tag, operand),
             _scalar(tag, mask), _ptr(num_copied) if num_copied is not None else None, _cfg(config)))


def copy_if_stencil(q: Queue, scratch: AllocationGroup, x, stencil, d, num_copied=None,
                    config=None) -> None:
    """copy_if with the predicate evaluated by the caller: copies x[i] where
    stencil[i] != 0 (stencil: uint8 or bool tensor of x's length)."""
    _check_contig(x, d, stencil)
    if stencil.numel() < x.numel() or stencil.element_size() != 1:
        raise ValueError("stencil must be a one-byte tensor at least as long as x")
    _check_out(d, x.dtype, x.numel(), "copy_if_stencil")
    if num_copied is not None:
        _check_out(num_copied, _torch().uint64, 1, "copy_if_stencil num_copied")
    tag = _tag(x.dtype)
    if tag not in _COPY_IF_TAGS:
        raise TypeError(f"copy_if takes integer, f32 or f64 elements, not {tag}")
    fn = getattr(_abi.load(), f"acpp_copy_if_stencil_{tag}")
    check(fn(scratch.handle, q.handle, _ptr(x), x.numel(), _ptr(stencil), _ptr(d),
             _ptr(num_copied) if num_copied is not None else None, _cfg(config)))


# ---------------------------------------------------------------------------
# Segmented scans and reductions
# ---------------------------------------------------------------------------
def _seg_check(x, offsets):
    if _tag(offsets.dtype) != "u64":
        raise TypeError("offsets must be uint64")
    if offsets.numel() < 1:
        raise ValueError("offsets must hold num_segments + 1 entries")
    _check_contig(x, offsets)


def segmented_inclusive_scan(q: Queue, scratch: AllocationGroup, x, d, offsets, op=PLUS,
                             config=None) -> None:
    """d[i] = inclusive scan of x within the segment [offsets[s], offsets[s+1])."""
    _seg_check(x, offsets)
    _check_contig(d)
    _check_out(d, x.dtype, x.numel(), "segmented_inclusive_scan")
    tag = _tag(x.dtype)
    fn = getattr(_abi.load(), f"acpp_segmented_inclusive_scan_{tag}")
    check(fn(scratch.handle, q.handle, _ptr(x), x.numel(), _ptr(d), _ptr(offsets),
             offsets.numel() - 1, _op(op), _cfg(config)))


def segmented_reduce(q: Queue, scratch: AllocationGroup, x, offsets, out, init=0, op=PLUS,
                     config=None) -> None:
    """out[s] = op(init, fold of x over [offsets[s], offsets[s+1]))."""
    _seg_check(x, offsets)
    _check_contig(out)
    _check_out(out, x.dtype, offsets.numel() - 1, "segmented_reduce")
    tag = _tag(x.dtype)
    fn = getattr(_abi.load(), f"acpp_segmented_reduce_{tag}")
    check(fn(scratch.handle, q.handle, _ptr(x), x.numel(), _ptr(offsets), offsets.numel() - 1,
             _ptr(out), _scalar(tag, init), _op(op), _cfg(config)))


class SegmentedPlan:
    """A segmented scan or reduction captured into a hipGraph once and
    replayed with ``launch``.

    The captured graph holds raw pointers to the input, output and offsets
    tensors and to scratch regions of the allocation group, so the plan
    keeps references to all of them (and to the group's cache) for as long
    as it lives: a replay never touches freed or reissued memory."""

    def __init__(self, handle, scratch, keep=()):
        self._h = handle
        self._scratch = scratch
        scratch._plans += 1
        self._keep = tuple(keep)

    @staticmethod
    def inclusive_scan(scratch: AllocationGroup, x, d, offsets, op=PLUS, config=None):
        _seg_check(x, offsets)
        _check_contig(d)
        _check_out(d, x.dtype, x.numel(), "segmented plan inclusive_scan")
        tag = _tag(x.dtype)
        h = ctypes.c_void_p()
        fn = getattr(_abi.load(), f"acpp_segmented_plan_inclusive_scan_{tag}")
        check(fn(scratch.handle, _ptr(x), x.numel(), _ptr(d), _ptr(offsets), offsets.numel() - 1,
                 _op(op), _cfg(config), ctypes.byref(h)))
        return SegmentedPlan(h, scratch, (x, d, offsets))

    @staticmethod
    def reduce(scratch: AllocationGroup, x, offsets, out, init=0, op=PLUS, config=None):
        _seg_check(x, offsets)
        _check_contig(out)
        _check_out(out, x.dtype, offsets.numel() - 1, "segmented plan reduce")
        tag = _tag(x.dtype)
        h = ctypes.c_void_p()
        fn = getattr(_abi.load(), f"acpp_segmented_plan_reduce_{tag}")
        check(fn(scratch.handle, _ptr(x), x.numel(), _ptr(offsets), offsets.numel() - 1,
                 _ptr(out), _scalar(tag, init), _op(op), _cfg(config), ctypes.byref(h)))
        return SegmentedPlan(h, scratch, (x, out, offsets))

    def launch(self, q: Queue) -> None:
        check(_abi.load().acpp_segmented_plan_launch(self._h, q.handle))

    def close(self) -> None:
        if self._h:
            h, self._h = self._h, None
            check(_abi.load().acpp_segmented_plan_destroy(h))
        self._keep = ()
        g, self._scratch = self._scratch, None
        if g is not None:
            g._plans -= 1
            if g._close_pending and g._plans == 0:
                g.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
