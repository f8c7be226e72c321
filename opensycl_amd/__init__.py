"""MI355X-native device-wide primitives with AdaptiveCpp's acpp::algorithms API.

The compute path is the HIP library ``lib/libacpp_mi355x.so`` (gfx950),
reached through the C-ABI in ``include/acpp_mi355x.h``.
"""
from . import _abi  # noqa: F401
from ._abi import AcppError, ERROR_WATCHDOG, VARIANT_FORCE_GIVE_UP  # noqa: F401
from .algorithms import (ALLOCATION_DEVICE, ALLOCATION_HOST, ALLOCATION_SHARED,  # noqa: F401
                         ATOMICS_AUTO, ATOMICS_GLOBAL, ATOMICS_PRIVATIZED, atomic_fetch_strided,
                         histogram, SegmentedPlan, segmented_inclusive_scan, segmented_reduce,
                         BIT_AND, BIT_OR, BIT_XOR, LOGICAL_AND, LOGICAL_OR, MAX, MIN, MULTIPLY,
                         PLUS, AllocationCache, AllocationGroup, Queue, compute_units,
                         exclusive_scan, inclusive_scan, reduce, reduce_compensated, reduction,
                         transform_reduce)
from .algorithms import (PRED_EQ, PRED_GE, PRED_GT, PRED_LE, PRED_LT, PRED_MASK_EQ,  # noqa: F401
                         PRED_NE, PRED_NONZERO, check_asynchronous_errors, copy_if, copy_if_stencil,
                         reduction_multi)
from .algorithms import (ABS, BIT_NOT, IDENTITY, IS_NONZERO, IS_ZERO, NEGATE,  # noqa: F401
                         SQUARE, TWICE, transform_exclusive_scan, transform_inclusive_scan,
                         transform_reduce_unary)
from .algorithms import (B_ABS_DIFF, B_BIT_AND, B_BIT_OR, B_BIT_XOR, B_MAX, B_MIN,  # noqa: F401
                         B_MINUS, B_MULTIPLIES, B_PLUS, B_SQ_DIFF)
from .cost_model import CostModel  # noqa: F401

__version__ = "0.1.0"
