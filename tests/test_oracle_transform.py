"""Pins the oracle's element transforms and the transform forms built on them
(oracle_transform_*, oracle.transform_reduce / transform_*_scan).

Checks that are not the oracle's own formula:
  - numpy ufuncs with wrapping integer arithmetic (np.negative, np.abs,
    np.multiply, np.add, np.invert, comparisons) for every element type;
  - closed forms: sums of -i, i^2, 2i, ~i over i < n;
  - the reference's pstl transform tests in closed form
    (tests/golden/reference_vectors.json: pstl_transform_scans,
    pstl_transform_reduce);
  - a planted mistake (fault 7: abs returns x unchanged) that the numpy check
    catches.
Pinning status: pinned for every code and element type.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))
INT_TYPES = [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64]
FP_TYPES = [np.float32, np.float64]


@pytest.fixture(autouse=True)
def _no_fault():
    O.set_fault(0)
    yield
    O.set_fault(0)


def _np_transform(x, code):
    with np.errstate(over="ignore"):
        if code == O.IDENTITY:
            return x.copy()
        if code == O.NEGATE:
            return np.negative(x)
        if code == O.ABS:
            return np.abs(x)
        if code == O.SQUARE:
            return np.multiply(x, x)
        if code == O.TWICE:
            return np.add(x, x)
        if code == O.BIT_NOT:
            return np.invert(x)
        if code == O.IS_ZERO:
            return (x == 0).astype(x.dtype)
        return (x != 0).astype(x.dtype)


def _inputs(dtype, n=4099, seed=3):
    rng = np.random.default_rng(seed)
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        x = rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
        x[:4] = [0, 1, info.max, info.min]  # includes the most negative value (abs wraps)
        x[rng.random(n) < 0.1] = 0
        return x
    x = (rng.standard_normal(n) * 100).astype(dtype)
    x[:4] = [0.0, -0.0, np.inf, -np.inf]
    x[rng.random(n) < 0.1] = 0
    return x


@pytest.mark.parametrize("dtype", INT_TYPES + FP_TYPES)
def test_transform_matches_numpy(dtype):
    x = _inputs(dtype)
    for code in range(8):
        if code == O.BIT_NOT and dtype in FP_TYPES:
            with pytest.raises(ValueError):
                O.transform(x, code)
            continue
        got = O.transform(x, code)
        exp = _np_transform(x, code)
        assert got.dtype == x.dtype
        # bit-exact, -0.0 included
        assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), (dtype, code)


def test_transform_rejects_unknown_code():
    with pytest.raises(ValueError):
        O.transform(np.zeros(4, np.int32), 8)
    with pytest.raises(ValueError):
        O.transform(np.zeros(4, np.int32), -1)


def test_closed_forms():
    n = 100000
    x = np.arange(n, dtype=np.int64)
    assert O.transform_reduce(x, 0, O.PLUS, O.NEGATE) == -(n * (n - 1) // 2)
    assert O.transform_reduce(x, 0, O.PLUS, O.SQUARE) == (n - 1) * n * (2 * n - 1) // 6
    assert O.transform_reduce(x, 0, O.PLUS, O.TWICE) == n * (n - 1)
    assert O.transform_reduce(x, 0, O.PLUS, O.BIT_NOT) == -(n * (n + 1) // 2)
    assert O.transform_reduce(x, 0, O.PLUS, O.IS_ZERO) == 1
    assert O.transform_reduce(x, 0, O.PLUS, O.IS_NONZERO) == n - 1
    c = x - n // 2
    assert O.transform_reduce(c, 0, O.PLUS, O.ABS) == 2 * sum(range(1, n // 2)) + n // 2
    # wrapping square in int8: 100 * 100 = 10000 = 16 (mod 256)
    assert O.transform(np.array([100, -128, 127], np.int8), O.SQUARE).tolist() == [16, 0, 1]


@pytest.mark.parametrize("n", GOLDEN["pstl_transform_scans"]["sizes"])
def test_golden_pstl_transform_scans(n):
    i = np.arange(n, dtype=np.int64)
    data = i.astype(np.int32)
    m32 = (1 << 32)
    wrap = lambda v: ((v % m32 + m32) % m32).astype(np.uint32).view(np.int32)  # noqa: E731
    unary_plus = O.TWICE  # plus(x, x)
    assert np.array_equal(O.transform_inclusive_scan(data, O.PLUS, unary_plus), wrap(i * (i + 1)))
    assert np.array_equal(O.transform_inclusive_scan(data, O.PLUS, unary_plus, np.int32(3)),
                          wrap(3 + i * (i + 1)))
    assert np.array_equal(O.transform_exclusive_scan(data, np.int32(3), O.PLUS, unary_plus),
                          wrap(3 + i * (i - 1)))
    u = i.astype(np.uint64)
    unary_mul = O.SQUARE  # multiplies(x, x)
    zeros = np.zeros(n, np.uint64)
    assert np.array_equal(O.transform_inclusive_scan(u, O.MULTIPLY, unary_mul), zeros)
    assert np.array_equal(O.transform_inclusive_scan(u, O.MULTIPLY, unary_mul, np.uint64(3)), zeros)
    ex = zeros.copy()
    if n:
        ex[0] = 3
    assert np.array_equal(O.transform_exclusive_scan(u, np.uint64(3), O.MULTIPLY, unary_mul), ex)


@pytest.mark.parametrize("case", GOLDEN["pstl_transform_reduce"]["cases"])
def test_golden_pstl_transform_reduce(case):
    dt = np.int64 if case.get("dtype") == "int64" else np.int32
    x = np.arange(case["n"], dtype=dt)
    assert O.transform_reduce(x, dt(case["init"]), O.PLUS, O.IDENTITY) == case["expected"]


def test_planted_fault_is_caught():
    x = _inputs(np.int32)
    O.set_fault(7)
    got = O.transform(x, O.ABS)
    assert not np.array_equal(got, _np_transform(x, O.ABS))
    xf = _inputs(np.float32)
    assert not np.array_equal(O.transform(xf, O.ABS), _np_transform(xf, O.ABS))


# ---------------------------------------------------------------------------
# Binary transform_reduce (numeric.hpp:170-191): pinned against numpy's
# wrapping ufuncs and closed forms; a planted |x-y| -> x-y must fail.
# ---------------------------------------------------------------------------
BIN_TYPES = [np.int32, np.int64, np.uint32, np.uint64, np.float32, np.float64]


def np_pair(code, a, b):
    with np.errstate(over="ignore"):
        if code == O.B_MULTIPLIES:
            return np.multiply(a, b, dtype=a.dtype)
        if code == O.B_PLUS:
            return np.add(a, b, dtype=a.dtype)
        if code == O.B_MINUS:
            return np.subtract(a, b, dtype=a.dtype)
        if code == O.B_ABS_DIFF:
            d = np.subtract(a, b, dtype=a.dtype)
            return np.abs(d) if np.issubdtype(a.dtype, np.signedinteger) or \
                np.issubdtype(a.dtype, np.floating) else d
        if code == O.B_SQ_DIFF:
            d = np.subtract(a, b, dtype=a.dtype)
            return np.multiply(d, d, dtype=a.dtype)
        if code == O.B_MIN:
            return np.where(a < b, a, b)
        if code == O.B_MAX:
            return np.where(a > b, a, b)
        if code == O.B_BIT_AND:
            return np.bitwise_and(a, b)
        if code == O.B_BIT_OR:
            return np.bitwise_or(a, b)
        return np.bitwise_xor(a, b)


def np_fold(op, v, init):
    x = np.concatenate([np.array([init], dtype=v.dtype), v])
    with np.errstate(over="ignore"):
        return {O.PLUS: lambda: np.add.reduce(x, dtype=x.dtype),
                O.MULTIPLY: lambda: np.multiply.reduce(x, dtype=x.dtype),
                O.MIN: lambda: np.minimum.reduce(x), O.MAX: lambda: np.maximum.reduce(x)}[op]()


@pytest.mark.parametrize("dt", BIN_TYPES)
def test_binary_transform_reduce_vs_numpy(dt):
    rng = np.random.default_rng(11)
    codes = range(10) if np.issubdtype(dt, np.integer) else range(7)
    for n in (0, 1, 5, 1000):
        if np.issubdtype(dt, np.integer):
            info = np.iinfo(dt)
            a = rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)
            b = rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)
        else:
            # integer-valued, so every fold order gives the same bits
            a = rng.integers(-50, 50, n).astype(dt)
            b = rng.integers(-50, 50, n).astype(dt)
        for code in codes:
            for op in (O.PLUS, O.MIN, O.MAX) + ((O.MULTIPLY,) if np.issubdtype(dt, np.integer) else ()):
                got = O.transform_reduce_binary(a, b, dt(3), op, code)
                want = np_fold(op, np_pair(code, a, b), dt(3))
                assert got == want, (dt, n, code, op)


def check_binary_closed_forms():
    i = np.arange(1, 1001, dtype=np.int64)
    ok = O.transform_reduce_binary(i, 2 * i, np.int64(0), O.PLUS, O.B_ABS_DIFF) == 1000 * 1001 // 2
    ok &= O.transform_reduce_binary(i, 2 * i, np.int64(0), O.PLUS, O.B_SQ_DIFF) == 1000 * 1001 * 2001 // 6
    ok &= O.transform_reduce_binary(i, 3 * i, np.int64(7), O.MAX, O.B_ABS_DIFF) == 2000
    f = i.astype(np.float64)
    ok &= O.transform_reduce_binary(f, 2 * f, 0.0, O.PLUS, O.B_ABS_DIFF) == 500500.0
    return bool(ok)


def test_binary_transform_reduce_closed_forms():
    assert check_binary_closed_forms()
    O.set_fault(15)
    try:
        assert not check_binary_closed_forms()
    finally:
        O.set_fault(0)


def test_binary_transform_reduce_dot_closed_form():
    # plus with multiplies is the dot product (numeric.hpp:215-223):
    # sum over i < n of i (i + 1) = (n-1) n (n+1) / 3, and oracle_dot agrees
    for n in (1, 10, 1000):
        a = np.arange(n, dtype=np.int64)
        assert O.transform_reduce_binary(a, a + 1, np.int64(0), O.PLUS, O.B_MULTIPLIES) == \
            (n - 1) * n * (n + 1) // 3
        assert O.transform_reduce_binary(a, a + 1, np.int64(0), O.PLUS, O.B_MULTIPLIES) == O.dot(a, a + 1, 0)


def test_binary_transform_reduce_rejects():
    a = np.zeros(3, dtype=np.float32)
    with pytest.raises(ValueError):
        O.transform_reduce_binary(a, a, np.float32(0), O.PLUS, O.B_BIT_AND)
    with pytest.raises(ValueError):
        O.transform_reduce_binary(a.astype(np.int32), a.astype(np.int32), 0, O.PLUS, 10)
