"""Pins the oracle's copy_if (oracle_copy_if_<t>, a restatement of
include/hipSYCL/algorithms/algorithm.hpp:178-243) against the reference's
own test cases and numpy's boolean indexing, with a planted fault."""
import numpy as np
import pytest

import oracle as O

TYPES = [np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64]


@pytest.fixture(autouse=True)
def _no_fault():
    O.set_fault(0)
    yield
    O.set_fault(0)


# tests/pstl/copy_if.cpp:25-64: data[i] = gen(i), predicate x % 2 == 0,
# compared with std::copy_if; the expected counts follow from the generators.
@pytest.mark.parametrize("n,gen,count", [
    (0, lambda i: i, 0),            # par_unseq_empty (46-48)
    (1, lambda i: i + 3, 0),        # par_unseq_single_element (50-52)
    (1000, lambda i: 1, 0),         # par_unseq_none (54-56)
    (1000, lambda i: 2 * i, 1000),  # par_unseq_all (58-60)
    (1000, lambda i: i, 500),       # par_unseq_half (62-64)
])
def test_reference_cases(n, gen, count):
    x = np.array([gen(i) for i in range(n)], dtype=np.int32)
    out, c = O.copy_if(x, O.PRED_MASK_EQ, 0, 1)   # (x & 1) == 0  <=>  x % 2 == 0
    assert c == count
    assert np.array_equal(out, np.array([v for v in x if v % 2 == 0], dtype=np.int32))


def _np_pred(x, pred, operand, mask, stencil):
    return {
        O.PRED_NONZERO: x != 0, O.PRED_EQ: x == operand, O.PRED_NE: x != operand,
        O.PRED_LT: x < operand, O.PRED_LE: x <= operand, O.PRED_GT: x > operand,
        O.PRED_GE: x >= operand,
        O.PRED_MASK_EQ: (x & mask) == operand if np.issubdtype(x.dtype, np.integer) else None,
        O.PRED_STENCIL: stencil != 0 if stencil is not None else None,
    }[pred]


@pytest.mark.parametrize("dtype", TYPES)
@pytest.mark.parametrize("pred", range(9))
def test_matches_numpy_boolean_indexing(dtype, pred):
    if pred == O.PRED_MASK_EQ and not np.issubdtype(dtype, np.integer):
        pytest.skip("mask predicates are integer-only")
    rng = np.random.default_rng(pred)
    n = 5003
    if np.issubdtype(dtype, np.integer):
        x = rng.integers(0, 50, n).astype(dtype)
        operand, mask = dtype(20), dtype(3)
    else:
        x = np.round(rng.standard_normal(n) * 3).astype(dtype)
        operand, mask = dtype(1.0), dtype(0)
    stencil = rng.integers(0, 2, n).astype(np.uint8) if pred == O.PRED_STENCIL else None
    out, c = O.copy_if(x, pred, operand, mask, stencil)
    ref = x[_np_pred(x, pred, operand, mask, stencil)]
    assert c == ref.size
    assert np.array_equal(out, ref)


def test_planted_count_fault_is_caught():
    x = np.arange(100, dtype=np.int32)
    O.set_fault(6)
    _, c = O.copy_if(x, O.PRED_GE, 50)
    O.set_fault(0)
    assert c != int(np.sum(x >= 50))
