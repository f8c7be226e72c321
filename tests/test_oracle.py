"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Every oracle function is checked against something that is not its own
formula:
  - the known-answer values the reference's own tests hold
    (tests/golden/reference_vectors.json, copied from the cited tests);
  - independent library routines: numpy's wrapping add/multiply/min/max/
    bitwise ufunc reductions and accumulations, np.bincount, math.fsum;
  - closed forms (sums of i, sums of squares) and brute force on tiny inputs.
A deliberate-mistake switch in the oracle (oracle_set_fault) shows that the
checks catch plausible errors: a dropped last element, an inclusive scan in
place of an exclusive one, an ignored init, an off-by-one group bound and a
lost histogram key.

Pinning status per oracle function:
  reduce, reduction, dot, inclusive/exclusive scan, group reduce/scans,
  atomic fold, segmented scan/reduce, histogram, fetch_max, FP64 sums and the
  FP64 absolute sum: pinned (golden vectors + numpy / fsum / closed forms).
  *_omp_order for integers: pinned (equal to numpy for every op).
  *_omp_order for floating point: pinned by crafted inputs whose result under
  the reference's evaluation order has a closed form (section "OpenMP
  evaluation order" below): values at 2^24 that absorb later ones expose how
  many elements a work-item folds serially (8), the shape of the 128-wide
  tree (stride halving), the stages over partials and the scan's tile length
  (1024 x 16); signed zeros under sycl::minimum / maximum, which return their
  second operand when both are zeros, expose every operand order (work-item
  fold op(acc, x), tree op(l[i], l[i+s]), final op(init, r)).  Planted
  faults 8-14 restate each of these wrongly and must fail.  The one part left
  open is the look-back's grouping of several aggregates in a multi-threaded
  scan, which the reference leaves to thread timing; the crafted inputs keep
  every such aggregate exact, so their result does not depend on it.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "reference_vectors.json")))
INT_TYPES = [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64]
FP_TYPES = [np.float32, np.float64]
INT_OPS = [O.PLUS, O.MULTIPLY, O.MIN, O.MAX, O.BIT_AND, O.BIT_OR, O.BIT_XOR, O.LOGICAL_AND,
           O.LOGICAL_OR]


@pytest.fixture(autouse=True)
def _no_fault():
    O.set_fault(0)
    yield
    O.set_fault(0)


def rand_ints(dtype, n, seed=0):
    info = np.iinfo(dtype)
    rng = np.random.default_rng(seed)
    return rng.integers(info.min, info.max, size=n, dtype=dtype, endpoint=True)


# ---------------------------------------------------------------------------
# numpy restatements used as independent references (library routines)
# ---------------------------------------------------------------------------
def np_fold(op, x, init):
    dt = x.dtype
    x = np.concatenate([np.array([init], dtype=dt), x])
    if op == O.PLUS:
        return np.add.reduce(x, dtype=dt)
    if op == O.MULTIPLY:
        return np.multiply.reduce(x, dtype=dt)
    if op == O.MIN:
        return np.minimum.reduce(x)
    if op == O.MAX:
        return np.maximum.reduce(x)
    if op == O.BIT_AND:
        return np.bitwise_and.reduce(x)
    if op == O.BIT_OR:
        return np.bitwise_or.reduce(x)
    if op == O.BIT_XOR:
        return np.bitwise_xor.reduce(x)
    if op == O.LOGICAL_AND:
        return dt.type(np.logical_and.reduce(x))
    return dt.type(np.logical_or.reduce(x))


def np_scan(op, x):
    dt = x.dtype
    if op == O.PLUS:
        return np.add.accumulate(x, dtype=dt)
    if op == O.MULTIPLY:
        return np.multiply.accumulate(x, dtype=dt)
    if op == O.MIN:
        return np.minimum.accumulate(x)
    if op == O.MAX:
        return np.maximum.accumulate(x)
    if op == O.BIT_AND:
        return np.bitwise_and.accumulate(x)
    if op == O.BIT_OR:
        return np.bitwise_or.accumulate(x)
    if op == O.BIT_XOR:
        return np.bitwise_xor.accumulate(x)
    if op == O.LOGICAL_AND:
        r = np.logical_and.accumulate(x != 0).astype(dt)
    else:
        r = np.logical_or.accumulate(x != 0).astype(dt)
    # a single element keeps its raw value under std::inclusive_scan (no op applied)
    if x.size:
        r[0] = x[0]
    return r


# ---------------------------------------------------------------------------
# golden vectors from the reference's tests
# ---------------------------------------------------------------------------
def check_cbs_group_sums():
    g = GOLDEN["cbs_reduce_for"]
    x = np.arange(1024, dtype=np.int32)
    y = O.group_reduce(x, g["local_size"], O.PLUS)
    return list(y[:: g["local_size"]]) == g["expected_group_sums"]


def test_golden_cbs_group_reduce():
    assert check_cbs_group_sums()


def test_golden_cbs_accumulator_partials():
    g = GOLDEN["cbs_accumulator_for"]
    x = np.arange(1024, dtype=np.int32)
    L = g["local_size"]
    lid1 = [O.reduce(x[b * L + 1:(b + 1) * L:2], 0) for b in range(4)]
    lid0 = [O.reduce(x[b * L:(b + 1) * L], 0) for b in range(4)]
    assert lid1 == g["expected_lid1"]
    assert lid0 == g["expected_lid0"]


def test_golden_babelstream_dot():
    g = GOLDEN["babelstream_dot"]
    a = np.arange(1024, dtype=np.int32)
    gsz = g["num_groups"] * g["local_size"]
    parts = []
    for grp in range(g["num_groups"]):
        idx = np.array([j for j in range(1024) if (j % gsz) // g["local_size"] == grp])
        parts.append(int(O.dot(a[idx], a[idx], 0)))
    assert parts == g["expected_group_partials"]


def test_golden_sycl_reduction():
    x = np.arange(1024, dtype=np.int32)
    g = GOLDEN["sycl_reduction_accessor"]
    assert O.reduction(x, 12345, O.PLUS, True) == g["expected_sum"]
    assert O.reduction(x, -7, O.MAX, True) == g["expected_max"]
    gi = GOLDEN["sycl_reduction_incremental"]
    v = O.reduction(x, 0, O.PLUS, False)
    assert v == gi["expected_after_first"]
    assert O.reduction(x, v, O.PLUS, False) == gi["expected_after_second"]


def check_reduction_products():
    g = GOLDEN["sycl_reduction_multiplies_generator"]
    ok = True
    for n_s, expected in g["expected_products"].items():
        n = int(n_s)
        i = np.arange(n)
        x = np.where(i > 1500, 1, np.where(i % 100 == 0, 2, 1)).astype(np.int64)
        ok &= O.reduction(x, 0, O.MULTIPLY, True) == expected
        ok &= O.reduce(x, 1, O.MULTIPLY) == expected
    return ok


def test_golden_reduction_products():
    assert check_reduction_products()


def check_pstl_reduce():
    ok = True
    for c in GOLDEN["pstl_reduce"]["cases"]:
        dt = np.int64 if c.get("dtype") == "int64" else np.int32
        x = np.arange(c["n"], dtype=dt)
        ok &= O.reduce(x, c["init"]) == c["expected"]
    return ok


def test_golden_pstl_reduce():
    assert check_pstl_reduce()


def test_golden_std_atomic():
    g = GOLDEN["std_atomic_fetch"]
    codes = {"add": O.RMW_ADD, "sub": O.RMW_SUB, "or": O.RMW_OR, "and": O.RMW_AND,
             "xor": O.RMW_XOR}
    for dt in (np.int32, np.int64, np.uint32, np.uint64):
        for c in g["cases"]:
            got = O.atomic_fold(codes[c["op"]], np.array([c["operand"]], dtype=dt), g["initial"])
            assert got == c["expected"]


def test_golden_sycl_atomic_fetch_op():
    g = GOLDEN["sycl_atomic_fetch_op"]["expected_int32"]
    operands = -np.arange(1, 2048, dtype=np.int32)
    codes = {"add": O.RMW_ADD, "sub": O.RMW_SUB, "min": O.RMW_MIN, "max": O.RMW_MAX,
             "and": O.RMW_AND, "or": O.RMW_OR}
    for name, code in codes.items():
        assert O.atomic_fold(code, operands, 0) == g[name], name


@pytest.mark.parametrize("dt", [np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64])
def test_sycl_atomic_fold_brute_force(dt):
    # Serial fold of tests/sycl/atomic.cpp:112-123 with Python integers/floats.
    vals = [-i for i in range(2048)]
    if np.issubdtype(dt, np.integer):
        bits = np.iinfo(dt).bits
        mask = (1 << bits) - 1
        signed = np.iinfo(dt).min < 0

        def wrap(v):
            v &= mask
            return v - (1 << bits) if signed and v >> (bits - 1) else v
        ops = {O.RMW_ADD: lambda a, b: wrap(a + b), O.RMW_SUB: lambda a, b: wrap(a - b),
               O.RMW_AND: lambda a, b: wrap(a & b), O.RMW_OR: lambda a, b: wrap(a | b),
               O.RMW_XOR: lambda a, b: wrap(a ^ b), O.RMW_MIN: min, O.RMW_MAX: max}
        conv = [wrap(v) for v in vals]
    else:
        ops = {O.RMW_ADD: lambda a, b: a + b, O.RMW_SUB: lambda a, b: a - b,
               O.RMW_MIN: min, O.RMW_MAX: max}
        conv = [float(v) for v in vals]
    operands = np.array(conv[1:], dtype=dt)
    for code, f in ops.items():
        exp = conv[0]
        for v in conv[1:]:
            exp = f(exp, v)
        assert O.atomic_fold(code, operands, conv[0]) == dt(exp)


# ---------------------------------------------------------------------------
# independent library routines
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dt", INT_TYPES)
@pytest.mark.parametrize("op", INT_OPS)
def test_reduce_and_scans_vs_numpy_int(dt, op):
    for n in (1, 2, 127, 1000, 4099):
        x = rand_ints(dt, n, seed=n)
        if op in (O.LOGICAL_AND, O.LOGICAL_OR):
            x[np.random.default_rng(n).random(n) < 0.3] = 0
        init = dt(7)
        assert O.reduce(x, init, op) == np_fold(op, x, init)
        assert np.array_equal(O.inclusive_scan(x, op), np_scan(op, x))
        with_init = O.inclusive_scan(x, op, init)
        ref = np_scan(op, np.concatenate([np.array([init], dtype=dt), x]))[1:]
        assert np.array_equal(with_init, ref)
        ex = O.exclusive_scan(x, init, op)
        assert ex[0] == init
        assert np.array_equal(ex[1:], ref[:-1])


@pytest.mark.parametrize("dt", FP_TYPES)
@pytest.mark.parametrize("op", [O.MIN, O.MAX])
def test_fp_minmax_vs_numpy(dt, op):
    x = np.random.default_rng(3).standard_normal(5000).astype(dt)
    assert O.reduce(x, dt(np.inf) if op == O.MIN else dt(-np.inf), op) == np_fold(op, x, x[0])
    assert np.array_equal(O.inclusive_scan(x, op), np_scan(op, x))


def test_fp_serial_sum_is_left_fold():
    # std::accumulate order: compare with a float32 left fold done by numpy's
    # sequential accumulate (np.add.accumulate is strictly sequential).
    x = np.random.default_rng(4).standard_normal(10000).astype(np.float32)
    assert O.reduce(x, np.float32(0)) == np.add.accumulate(
        np.concatenate([[np.float32(0)], x]), dtype=np.float32)[-1]
    assert np.array_equal(O.inclusive_scan(x), np.add.accumulate(x, dtype=np.float32))


def test_identities_are_identities():
    for dt in INT_TYPES + FP_TYPES:
        ops = INT_OPS if np.issubdtype(dt, np.integer) else [O.PLUS, O.MULTIPLY, O.MIN, O.MAX]
        xs = rand_ints(dt, 64) if np.issubdtype(dt, np.integer) else \
            np.array([-0.0, 0.0, 1.5, -3.25, 1e30, -1e-30], dtype=dt)
        for op in ops:
            e = O.identity(dt, op)
            for v in xs:
                if op in (O.LOGICAL_AND, O.LOGICAL_OR):
                    assert O.reduce(np.array([v], dtype=dt), e, op) == dt(bool(v))
                else:
                    r = O.reduce(np.array([v], dtype=dt), e, op)
                    assert r == v and np.signbit(r) == np.signbit(v)


def test_fp64_oracles_vs_fsum():
    rng = np.random.default_rng(5)
    x = rng.standard_normal(200003).astype(np.float32)
    x[rng.random(x.size) < 0.01] *= np.float32(1e4)
    assert O.sum_fp64(x) == math.fsum(x.astype(np.float64))
    y = rng.standard_normal(x.size).astype(np.float32)
    assert O.dot_fp64(x, y) == pytest.approx(
        math.fsum((x.astype(np.float64) * y.astype(np.float64)).tolist()), rel=1e-15, abs=1e-9)
    b = (x.view(np.uint32) >> 16).astype(np.uint16)
    widened = (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    assert O.sum_fp64(b) == math.fsum(widened)


def test_group_scans_vs_numpy():
    for L in (25, 144, 256):
        x = rand_ints(np.int32, 4 * L, seed=L)
        inc = O.group_inclusive_scan(x, L)
        exc = O.group_exclusive_scan(x, L, np.int32(5))
        red = O.group_reduce(x, L)
        for b in range(4):
            seg = x[b * L:(b + 1) * L]
            c = np.add.accumulate(seg, dtype=np.int32)
            assert np.array_equal(inc[b * L:(b + 1) * L], c)
            assert exc[b * L] == 5
            assert np.array_equal(exc[b * L + 1:(b + 1) * L], (c[:-1] + np.int32(5)).astype(np.int32))
            assert np.all(red[b * L:(b + 1) * L] == c[-1])


def test_segmented_vs_numpy():
    rng = np.random.default_rng(6)
    lens = rng.integers(0, 50, 200)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    x = rng.integers(-100, 100, int(off[-1])).astype(np.int64)
    s = O.segmented_inclusive_scan(x, off)
    r = O.segmented_reduce(x, off, np.int64(3))
    for i in range(len(lens)):
        seg = x[off[i]:off[i + 1]]
        assert np.array_equal(s[off[i]:off[i + 1]], np.cumsum(seg))
        assert r[i] == 3 + seg.sum()
    f = rng.standard_normal(int(off[-1])).astype(np.float32)
    d = O.segmented_reduce_f32_fp64(f, off)
    for i in range(len(lens)):
        assert d[i] == math.fsum(f[off[i]:off[i + 1]].astype(np.float64))


def check_histogram():
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 300, 100000).astype(np.uint32)
    return np.array_equal(O.histogram(keys, 256), np.bincount(keys[keys < 256], minlength=256))


def test_histogram_and_fetch_max_vs_numpy():
    assert check_histogram()
    v = rand_ints(np.int32, 1024 * 37 + 5, seed=8)
    pad = np.full(1024 * 38, np.iinfo(np.int32).min, dtype=np.int32)
    pad[:v.size] = v
    assert np.array_equal(O.fetch_max_strided(v, 1024), pad.reshape(38, 1024).max(axis=0))


def test_closed_forms_ladder():
    # size ladder of tests/pstl/inclusive_scan.cpp:158-176: 0, 1, 127, 1000, 2^20
    for n in (0, 1, 127, 1000, 1 << 20):
        i = np.arange(n, dtype=np.int64)
        x = i.astype(np.int32)
        inc = O.inclusive_scan(x, O.PLUS, np.int32(3))
        exc = O.exclusive_scan(x, np.int32(3))
        assert np.array_equal(inc, ((i * (i + 1) // 2 + 3) & 0xffffffff).astype(np.uint32).view(np.int32))
        assert np.array_equal(exc, ((i * (i - 1) // 2 + 3) & 0xffffffff).astype(np.uint32).view(np.int32))
        m = O.inclusive_scan(i.astype(np.uint64), O.MULTIPLY, np.uint64(3))
        assert np.all(m == 0)


# ---------------------------------------------------------------------------
# OpenMP evaluation order of the reference
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dt", [np.int32, np.int64, np.uint32, np.int8])
def test_omp_order_int_equals_serial(dt):
    for n in (1, 127, 1000, 16385, 300000):
        x = rand_ints(dt, n, seed=n)
        for op in (O.PLUS, O.MIN, O.MAX, O.BIT_XOR):
            assert O.reduce_omp_order(x, dt(1), op, 3) == O.reduce(x, dt(1), op)
        assert np.array_equal(O.scan_omp_order(x, False, None, O.PLUS, 4), O.inclusive_scan(x))
        assert np.array_equal(O.scan_omp_order(x, True, dt(2), O.PLUS, 4),
                              O.exclusive_scan(x, dt(2)))
        assert np.array_equal(O.scan_omp_order(x, False, dt(2), O.PLUS, 4),
                              O.inclusive_scan(x, O.PLUS, dt(2)))
    a = rand_ints(dt, 5000, 1)
    b = rand_ints(dt, 5000, 2)
    assert O.dot_omp_order(a, b, dt(0), 3) == O.dot(a, b, dt(0))


def test_omp_order_fp_close_to_fp64():
    x = np.random.default_rng(9).standard_normal(1 << 20).astype(np.float32)
    exact = O.sum_fp64(x)
    r = float(O.reduce_omp_order(x, np.float32(0)))
    assert abs(r - exact) <= 1e-5 * O.abs_sum_fp64(x)


# Crafted inputs.  In float32 every integer up to 2^24 is exact and the
# spacing above 2^24 is 2, so 2^24 + 1 rounds back to 2^24 (tie to even):
# ones folded serially into 2^24 are lost, ones summed among themselves first
# are kept.  fl32() is numpy's round-to-nearest-even conversion from float64.
T24 = float(1 << 24)


def fl32(v):
    return np.float32(np.float64(v))


def check_omp_reduce_chunking():
    # 1024 elements = one 128-wide group of 8-element work-items.  Work-item 0
    # folds 2^24 and its 7 ones serially (the ones are lost); the other 127
    # items hold 8 ones each, and the tree adds 2^24 + 8 * 127 exactly.
    # 16 per item would give 2^24 + 1008, a serial fold 2^24.
    x = np.ones(1024, dtype=np.float32)
    x[0] = T24
    ok = O.reduce_omp_order(x, np.float32(0), O.PLUS, 2) == fl32(T24 + 1016)
    a = np.ones(1024, dtype=np.float32)
    a[0] = 4096.0
    ok &= O.dot_omp_order(a, a.copy(), np.float32(0), 2) == fl32(T24 + 1016)
    return bool(ok)


def check_omp_reduce_tree_shape():
    # One group; work-items 1 and 65 hold a one each, item 0 holds 2^24.  The
    # stride-halving tree first joins items i and i + 64, so the two ones meet
    # (2) before reaching 2^24: 2^24 + 2.  Adjacent pairs would add each one
    # to 2^24 on its own and lose both.
    x = np.zeros(1024, dtype=np.float32)
    x[0], x[8 * 1], x[8 * 65] = T24, 1.0, 1.0
    ok = O.reduce_omp_order(x, np.float32(0), O.PLUS, 2) == fl32(T24 + 2)
    # The same pattern one stage up: 130 groups; the partials of groups 0, 1
    # and 65 are 2^24, 1 and 1 and meet in the second stage's 128-wide tree.
    y = np.zeros(1024 * 130, dtype=np.float32)
    y[0], y[1024 * 1], y[1024 * 65] = T24, 1.0, 1.0
    ok &= O.reduce_omp_order(y, np.float32(0), O.PLUS, 2) == fl32(T24 + 2)
    return bool(ok)


def check_omp_operand_order():
    # sycl::minimum is x < y ? x : y: on two zeros it returns the second
    # operand, and on a zero and the padding identity (the largest float) it
    # returns the zero.  The reference combines every pair as op(earlier,
    # later) -- op(acc, x) in a work-item, op(l[i], l[i + s]) in the tree,
    # op(init, r) at the end -- so over zeros the result is the zero the
    # tree's last combine takes as its second operand.  Following l[0] <-
    # l[1] <- l[3] <- ... <- l[63] <- l[127] back to a leaf: with full groups
    # that is the last element of work-item 127, i.e. the last input element
    # (n = 1024, and n = 1024 * 128 whose 128 partials fill the second stage).
    # With n = 1000 items 125..127 are padding, l[63] keeps its own value at
    # the stride-64 step, and the result is the last element of item 63,
    # x[511].  (sycl::maximum cannot show this: its padding identity,
    # numeric_limits<float>::min(), beats every zero.)
    ok = True
    for n, decisive in ((1024, 1023), (1024 * 128, 1024 * 128 - 1), (1000, 511)):
        for zero, other in ((-0.0, 0.0), (0.0, -0.0)):
            x = np.full(n, other, dtype=np.float32)
            x[decisive] = zero
            r = O.reduce_omp_order(x, np.float32(other), O.MIN, 2)
            ok &= bool(r == 0 and np.signbit(r) == np.signbit(np.float32(zero)))
    # the quirk itself: maximum over zeros returns the padding value
    r = O.reduce_omp_order(np.zeros(1024, dtype=np.float32), np.float32(-1.0), O.MAX, 2)
    ok &= bool(r == np.finfo(np.float32).tiny)
    return ok


def check_omp_scan_tile_length():
    # 2^24 then ones.  Tile 0 (16384 elements, scanned serially) absorbs its
    # ones: y = 2^24.  Later tiles scan their own ones exactly and add the
    # exact prefix of earlier tiles, so y[i] = fl32(2^24 + i - 16383).  With
    # 8192-element tiles the rise would start at 8192.
    n = 3 * 16384
    x = np.ones(n, dtype=np.float32)
    x[0] = T24
    i = np.arange(n, dtype=np.float64)
    want = (T24 + np.maximum(i - 16383, 0)).astype(np.float32)
    ok = np.array_equal(O.scan_omp_order(x, False, None, O.PLUS, 3), want)
    ex = O.scan_omp_order(x, True, np.float32(0), O.PLUS, 3)
    want_ex = want.copy()
    want_ex[0] = 0.0
    ok &= np.array_equal(ex, want_ex)
    return bool(ok)


def test_omp_reduce_ones_above_2e24():
    # 2^25 ones: every partial of the reference's order (8 per item, trees of
    # 128) is an exact power of two, so the sum is exact; a serial fold stops
    # at 2^24.
    x = np.ones(1 << 25, dtype=np.float32)
    assert O.reduce_omp_order(x, np.float32(0)) == np.float32(1 << 25)
    O.set_fault(9)
    assert O.reduce_omp_order(x, np.float32(0)) == np.float32(1 << 24)


def test_omp_scan_ones_above_2e24():
    # Every tile aggregate and prefix is exact, so y[i] = fl32(i + 1): the
    # rounding of the exact count, ties to even, past 2^24.  A serial scan
    # would stay at 2^24.
    n = (1 << 24) + 3 * 16384 + 5
    x = np.ones(n, dtype=np.float32)
    want = (np.arange(n, dtype=np.float64) + 1).astype(np.float32)
    assert np.array_equal(O.scan_omp_order(x, False, None, O.PLUS, 4), want)
    assert want[(1 << 24) + 2] == np.float32(T24 + 4)


def test_omp_dot_rounds_each_product():
    # transform then reduce: the product is rounded to float32 before the
    # add (numeric.hpp:204-223).  (-1)(1) + (1 + 2^-12)^2: the rounded product
    # 1 + 2^-11 gives 2^-11; a fused multiply-add would keep the 2^-24.
    a = np.array([-1.0, 1.0 + 2.0 ** -12], dtype=np.float32)
    b = np.array([1.0, 1.0 + 2.0 ** -12], dtype=np.float32)
    assert O.dot_omp_order(a, b, np.float32(0), 1) == np.float32(2.0 ** -11)
    assert O.dot_fp64(a, b) == 2.0 ** -11 + 2.0 ** -24


@pytest.mark.parametrize("check", [check_omp_reduce_chunking, check_omp_reduce_tree_shape,
                                   check_omp_operand_order, check_omp_scan_tile_length])
def test_omp_order_crafted(check):
    assert check()


@pytest.mark.parametrize("fault,check", [
    (8, check_omp_reduce_chunking),
    (9, check_omp_reduce_chunking),
    (10, check_omp_reduce_tree_shape),
    (11, check_omp_operand_order),
    (12, check_omp_scan_tile_length),
    (13, check_omp_operand_order),
    (14, check_omp_operand_order),
])
def test_omp_order_planted_mistakes_are_caught(fault, check):
    O.set_fault(fault)
    try:
        assert not check()
    finally:
        O.set_fault(0)


def test_abs_sum_fp64_vs_fsum():
    rng = np.random.default_rng(10)
    x = rng.standard_normal(300007).astype(np.float32)
    x[rng.random(x.size) < 0.01] *= np.float32(-1e4)
    assert O.abs_sum_fp64(x) == math.fsum(np.abs(x.astype(np.float64)).tolist())
    # closed form: sum of |i - 1000| over 0..2999 = 1000*1001/2 + 2000*1999/2
    i = np.arange(3000, dtype=np.float32) - np.float32(1000)
    assert O.abs_sum_fp64(i) == 1000 * 1001 / 2 + 1999 * 2000 / 2


def test_omp_order_reference_max_identity_quirk():
    # numeric.hpp:53 pads maximum with numeric_limits<float>::min(); with all
    # inputs negative the OpenMP order returns that padding value, while the
    # serial semantics (and the HIP kernels) return the true maximum.
    x = -np.arange(1, 200, dtype=np.float32)
    assert O.reduce_omp_order(x, np.float32(-1e30), O.MAX) == np.finfo(np.float32).tiny
    assert O.reduce(x, np.float32(-1e30), O.MAX) == np.float32(-1)


# ---------------------------------------------------------------------------
# the checks have teeth: plant mistakes in the oracle
# ---------------------------------------------------------------------------
def check_inclusive_closed_form():
    x = np.arange(1000, dtype=np.int64)
    return np.array_equal(O.inclusive_scan(x), x * (x + 1) // 2)


def check_exclusive_closed_form():
    x = np.arange(1000, dtype=np.int64)
    return np.array_equal(O.exclusive_scan(x, np.int64(0)), x * (x - 1) // 2)


@pytest.mark.parametrize("fault,check", [
    (1, check_inclusive_closed_form),
    (2, check_exclusive_closed_form),
    (3, check_pstl_reduce),
    (4, check_cbs_group_sums),
    (5, check_histogram),
])
def test_planted_mistakes_are_caught(fault, check):
    assert check()
    O.set_fault(fault)
    try:
        assert not check()
    finally:
        O.set_fault(0)
    assert check()
