"""GPU parity: unary transform_reduce and transform_inclusive_scan /
transform_exclusive_scan (numeric.hpp:193-212, 299-327) against the oracle.

Integer results are bit-exact for every element type, operator and
transform.  Floating-point inputs are small integers, so sums, minima and
maxima of their transforms are exact in any order and compare bit-exactly;
products use values near 1 and compare within 4 * n * eps relative.
"""
import json
import os

import numpy as np
import pytest

import opensycl_amd as acpp
import oracle as O
from gpu_util import ALL_OPS, FP_OPS, FP_TYPES, INT_TYPES, dev, host, op_input, rand

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))
INT_CODES = list(range(8))
FP_CODES = [0, 1, 2, 3, 4, 6, 7]
SIZES = [1, 127, 1000, 8193, 100003]


def _fp_input(dtype, op, n, seed):
    rng = np.random.default_rng(seed)
    if op == acpp.MULTIPLY:
        return (1.0 + rng.standard_normal(n) * 1e-3).astype(dtype)
    x = np.round(rng.standard_normal(n) * 8).astype(dtype)
    x[rng.random(n) < 0.05] = 0
    return x


def _reduce(torch, q, scratch, x, init, op, code, config=None):
    xd = dev(torch, x)
    out = torch.full((1,), 77, dtype=xd.dtype, device="cuda")
    acpp.transform_reduce_unary(q, scratch, xd, out, init, op, code, config=config)
    q.wait()
    return host(out)[0]


def _scan(torch, q, scratch, x, op, code, kind, init=None, config=None):
    xd = dev(torch, x)
    yd = torch.empty_like(xd)
    if kind == "inclusive":
        acpp.transform_inclusive_scan(q, scratch, xd, yd, op, code, init, config=config)
    else:
        acpp.transform_exclusive_scan(q, scratch, xd, yd, init, op, code, config=config)
    q.wait()
    return host(yd)


def _close(got, exp, n, dtype):
    eps = float(np.finfo(dtype).eps)
    got = np.asarray(got, np.float64)
    exp = np.asarray(exp, np.float64)
    same_inf = np.isinf(exp) & (got == exp)
    with np.errstate(invalid="ignore"):  # inf - inf where both overflowed
        ok = same_inf | (np.abs(got - exp) <= 4 * n * eps * np.abs(exp))
    return bool(np.all(ok))


@pytest.mark.parametrize("dtype", INT_TYPES)
def test_transform_reduce_int_exact(gpu, q, scratch, dtype):
    torch, _ = gpu
    for op in ALL_OPS:
        for n in (1, 1000, 100003):
            x = op_input(dtype, op, n, seed=n + op)
            init = dtype(3)
            for code in INT_CODES:
                got = _reduce(torch, q, scratch, x, init, op, code)
                assert got == O.transform_reduce(x, init, op, code), (dtype, op, n, code)


# 1-byte sums large enough for the grid-stride main loop (2^25 + 12345
# elements), where the bytes are summed four at a time: the plain sum under
# several launch shapes (the last with non-temporal loads, flag 2) and from
# misaligned starts (scalar head), and every transform, against the oracle.
# Every byte value occurs at every position of a 16-byte vector.
def test_byte_sums_main_loop(gpu, q, scratch):
    torch, _ = gpu
    for dtype in (np.int8, np.uint8):
        info = np.iinfo(dtype)
        rng = np.random.default_rng(5)
        x = rng.integers(info.min, info.max + 1, (1 << 25) + 12345).astype(dtype)
        x[:4096] = np.repeat(np.arange(256), 16).astype(np.uint8).view(dtype)
        init = dtype(-7 if dtype == np.int8 else 200)
        for code in INT_CODES:
            assert _reduce(torch, q, scratch, x, init, acpp.PLUS, code) == \
                O.transform_reduce(x, init, acpp.PLUS, code), (dtype, code)
        xd = dev(torch, x)
        out = torch.zeros(1, dtype=xd.dtype, device="cuda")
        for off in (0, 1, 5):
            want = O.reduce(x[off:], init, acpp.PLUS)
            for cfg in (None, (512, 2), (256, 8), (1024, 16), (1024, 4, 0, 2)):
                acpp.reduce(q, scratch, xd[off:], out, init, acpp.PLUS, config=cfg)
                q.wait()
                assert host(out)[0] == want, (dtype, off, cfg)
    assert scratch.watchdog_trips() == 0


@pytest.mark.parametrize("dtype", FP_TYPES)
def test_transform_reduce_fp(gpu, q, scratch, dtype):
    torch, _ = gpu
    for op in FP_OPS:
        for n in (1, 1000, 100003):
            x = _fp_input(dtype, op, n, seed=n + op)
            init = dtype(3) if op != acpp.MULTIPLY else dtype(1)
            for code in FP_CODES:
                got = _reduce(torch, q, scratch, x, init, op, code)
                exp = O.transform_reduce(x, init, op, code)
                if op == acpp.MULTIPLY:
                    assert _close(got, exp, n, dtype), (dtype, n, code, got, exp)
                else:
                    assert got == exp, (dtype, op, n, code, got, exp)


@pytest.mark.parametrize("dtype", INT_TYPES)
def test_transform_scans_int_exact(gpu, q, scratch, dtype):
    torch, _ = gpu
    for op in ALL_OPS:
        for n in SIZES:
            x = op_input(dtype, op, n, seed=7 * n + op)
            init = dtype(3)
            for code in INT_CODES:
                got = _scan(torch, q, scratch, x, op, code, "inclusive")
                assert np.array_equal(got, O.transform_inclusive_scan(x, op, code)), (dtype, op, n, code)
                got = _scan(torch, q, scratch, x, op, code, "inclusive", init)
                assert np.array_equal(got, O.transform_inclusive_scan(x, op, code, init)), \
                    (dtype, op, n, code, "init")
                got = _scan(torch, q, scratch, x, op, code, "exclusive", init)
                assert np.array_equal(got, O.transform_exclusive_scan(x, init, op, code)), \
                    (dtype, op, n, code, "excl")


PACKED_OPS = [acpp.PLUS, acpp.MULTIPLY, acpp.MIN, acpp.MAX, acpp.BIT_AND, acpp.BIT_OR, acpp.BIT_XOR]


# 8-bit transform scans run the packed-pair kernel (8 rows: default, 1024- and
# 256-wide) with the transform applied to four bytes per 32-bit operation,
# or the generic kernel (4 rows).  Every byte value occurs at every position
# of a 16-byte vector; the sizes give one partial tile, three full 1024 x 8
# tiles plus a ragged tail, and twelve full 256 x 8 tiles plus a tail.
@pytest.mark.parametrize("cfg", [None, (1024, 8), (256, 8), (1024, 4)])
def test_packed_byte_transform_scans(gpu, q, scratch, cfg):
    torch, _ = gpu
    for dtype in (np.int8, np.uint8):
        info = np.iinfo(dtype)
        for op in PACKED_OPS:
            for n in (1, 17, 1000, 3 * 131072 + 77):
                rng = np.random.default_rng(n + 31 * op)
                if op == acpp.MULTIPLY:
                    x = rng.choice(np.array([1, 3, 5, 7, 255 if dtype == np.uint8 else -1], dtype=dtype), n)
                else:
                    x = rng.integers(info.min, info.max + 1, n).astype(dtype)
                    if n >= 4096:
                        x[:4096] = np.repeat(np.arange(256), 16).astype(np.uint8).view(dtype)
                init = dtype(-3 if dtype == np.int8 else 250)
                for code in INT_CODES:
                    got = _scan(torch, q, scratch, x, op, code, "inclusive", config=cfg)
                    assert np.array_equal(got, O.transform_inclusive_scan(x, op, code)), (dtype, op, n, code, cfg)
                    got = _scan(torch, q, scratch, x, op, code, "inclusive", init, config=cfg)
                    assert np.array_equal(got, O.transform_inclusive_scan(x, op, code, init)), \
                        (dtype, op, n, code, cfg, "init")
                    got = _scan(torch, q, scratch, x, op, code, "exclusive", init, config=cfg)
                    assert np.array_equal(got, O.transform_exclusive_scan(x, init, op, code)), \
                        (dtype, op, n, code, cfg, "excl")
    assert scratch.watchdog_trips() == 0


@pytest.mark.parametrize("dtype", FP_TYPES)
def test_transform_scans_fp(gpu, q, scratch, dtype):
    torch, _ = gpu
    for op in FP_OPS:
        for n in SIZES:
            x = _fp_input(dtype, op, n, seed=3 * n + op)
            init = dtype(3) if op != acpp.MULTIPLY else dtype(1)
            for code in FP_CODES:
                for kind in ("inclusive", "exclusive"):
                    got = _scan(torch, q, scratch, x, op, code, kind, init)
                    if kind == "inclusive":
                        exp = O.transform_inclusive_scan(x, op, code, init)
                    else:
                        exp = O.transform_exclusive_scan(x, init, op, code)
                    if op == acpp.MULTIPLY:
                        assert _close(got, exp, n, dtype), (dtype, n, code, kind)
                    else:
                        assert np.array_equal(got, exp), (dtype, op, n, code, kind)


def test_transform_misaligned_inputs(gpu, q, scratch):
    # views that start off a 16-byte boundary take the scalar / partial paths
    torch, _ = gpu
    for dtype in (np.int8, np.int16, np.int32, np.int64):
        x = rand(dtype, 100003, seed=11)
        for off in (1, 3):
            xd = dev(torch, x)[off:]
            out = torch.zeros(1, dtype=xd.dtype, device="cuda")
            acpp.transform_reduce_unary(q, scratch, xd, out, 0, acpp.PLUS, acpp.SQUARE)
            yd = torch.empty_like(xd)
            acpp.transform_inclusive_scan(q, scratch, xd, yd, acpp.PLUS, acpp.NEGATE)
            q.wait()
            assert host(out)[0] == O.transform_reduce(x[off:], dtype(0), O.PLUS, O.SQUARE)
            assert np.array_equal(host(yd), O.transform_inclusive_scan(x[off:], O.PLUS, O.NEGATE))


def test_golden_pstl_transform_scans(gpu, q, scratch):
    # tests/pstl/transform_inclusive_scan.cpp / transform_exclusive_scan.cpp:
    # data[i] = i, unary_op = op(x, x), plus with init 3 and multiplies with 3ull
    torch, _ = gpu
    for n in GOLDEN["pstl_transform_scans"]["sizes"]:
        if n == 0:
            continue
        i = np.arange(n, dtype=np.int64)
        w = lambda v: (v & 0xffffffff).astype(np.uint32).view(np.int32)  # noqa: E731
        x = i.astype(np.int32)
        assert np.array_equal(_scan(torch, q, scratch, x, acpp.PLUS, acpp.TWICE, "inclusive"), w(i * (i + 1)))
        assert np.array_equal(_scan(torch, q, scratch, x, acpp.PLUS, acpp.TWICE, "inclusive", np.int32(3)),
                              w(3 + i * (i + 1)))
        assert np.array_equal(_scan(torch, q, scratch, x, acpp.PLUS, acpp.TWICE, "exclusive", np.int32(3)),
                              w(3 + i * (i - 1)))
        u = i.astype(np.uint64)
        assert np.all(_scan(torch, q, scratch, u, acpp.MULTIPLY, acpp.SQUARE, "inclusive") == 0)
        assert np.all(_scan(torch, q, scratch, u, acpp.MULTIPLY, acpp.SQUARE, "inclusive", np.uint64(3)) == 0)
        ex = _scan(torch, q, scratch, u, acpp.MULTIPLY, acpp.SQUARE, "exclusive", np.uint64(3))
        assert ex[0] == 3 and np.all(ex[1:] == 0)


@pytest.mark.parametrize("case", GOLDEN["pstl_transform_reduce"]["cases"])
def test_golden_pstl_transform_reduce(gpu, q, scratch, case):
    torch, _ = gpu
    dt = np.int64 if case.get("dtype") == "int64" else np.int32
    n = case["n"]
    xd = torch.arange(n, dtype=torch.int64 if dt == np.int64 else torch.int32, device="cuda")
    out = torch.full((1,), case["init"], dtype=xd.dtype, device="cuda")
    acpp.transform_reduce_unary(q, scratch, xd, out, dt(case["init"]), acpp.PLUS, acpp.IDENTITY)
    q.wait()
    assert host(out)[0] == case["expected"]


def test_transform_large_properties(gpu, q, scratch):
    # size-independent checks at 2^26 int32: the negated scan is the negation
    # of the scan (linearity), and the sum of squares matches the oracle.
    torch, _ = gpu
    n = 1 << 26
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    xd = torch.randint(-(1 << 15), (1 << 15) + 1, (n,), dtype=torch.int32, device="cuda", generator=g)
    yd = torch.empty_like(xd)
    zd = torch.empty_like(xd)
    acpp.transform_inclusive_scan(q, scratch, xd, yd, acpp.PLUS, acpp.NEGATE)
    acpp.inclusive_scan(q, scratch, xd, zd)
    out = torch.zeros(1, dtype=torch.int64, device="cuda")
    x64 = xd.to(torch.int64)
    acpp.transform_reduce_unary(q, scratch, x64, out, 0, acpp.PLUS, acpp.SQUARE)
    q.wait()
    assert torch.equal(yd, -zd)
    assert host(out)[0] == O.transform_reduce(host(x64), np.int64(0), O.PLUS, O.SQUARE)


def test_transform_errors_and_empty(gpu, q, scratch):
    torch, _ = gpu
    xf = torch.ones(16, dtype=torch.float32, device="cuda")
    out = torch.full((1,), 5.0, dtype=torch.float32, device="cuda")
    with pytest.raises(acpp.AcppError) as e:
        acpp.transform_reduce_unary(q, scratch, xf, out, 0.0, acpp.PLUS, acpp.BIT_NOT)
    assert e.value.status == 2
    with pytest.raises(acpp.AcppError) as e:
        acpp.transform_inclusive_scan(q, scratch, xf, xf, acpp.PLUS, 8)
    assert e.value.status == 1
    with pytest.raises(acpp.AcppError) as e:
        acpp.transform_reduce_unary(q, scratch, xf, out, 0.0, acpp.PLUS, acpp.SQUARE,
                                    config=(256, 8))
    assert e.value.status == 1
    # empty input leaves *out untouched (numeric.hpp:165-169)
    acpp.transform_reduce_unary(q, scratch, xf[:0], out, 1.0, acpp.PLUS, acpp.SQUARE)
    q.wait()
    assert host(out)[0] == 5.0


def test_transform_reduce_launch_configs(gpu, q, scratch):
    torch, _ = gpu
    x = rand(np.int32, 3 * 1024 * 1024 + 5, seed=2)
    exp = O.transform_reduce(x, np.int32(0), O.PLUS, O.ABS)
    for cfg in [(64, 4), (256, 4, 1), (1024, 4, 2), (512, 4, 0, 0)]:
        assert _reduce(torch, q, scratch, x, np.int32(0), acpp.PLUS, acpp.ABS, config=cfg) == exp, cfg


# ---------------------------------------------------------------------------
# Binary transform_reduce (numeric.hpp:170-191) with pair transforms.
# ---------------------------------------------------------------------------
BIN_INT_TYPES = [np.int32, np.int64, np.uint32, np.uint64]


@pytest.mark.parametrize("dtype", BIN_INT_TYPES + FP_TYPES)
def test_binary_transform_reduce(gpu, q, scratch, dtype):
    # integers: full-range values, bit-exact for every pair transform and
    # reduction; floating point: integers in [-2, 2], whose pair transforms
    # are at most 16 in magnitude, so every sum of up to 100003 of them is
    # below 2^24 and exact in any order
    torch, _ = gpu
    is_int = np.issubdtype(dtype, np.integer)
    codes = range(10) if is_int else range(7)
    ops = (acpp.PLUS, acpp.MULTIPLY, acpp.MIN, acpp.MAX) if is_int else (acpp.PLUS, acpp.MIN, acpp.MAX)
    for n in ((1, 1000, (1 << 20) + 3) if is_int else (1, 1000, 100003)):
        rng = np.random.default_rng(n)
        if is_int:
            a, b = rand(dtype, n + 1, seed=n), rand(dtype, n + 1, seed=n + 1)
        else:
            a = rng.integers(-2, 3, n + 1).astype(dtype)
            b = rng.integers(-2, 3, n + 1).astype(dtype)
        ad, bd = dev(torch, a), dev(torch, b)
        for code in codes:
            for op in ops:
                for off in (0, 1):  # views one element in: the scalar head
                    out = torch.full((1,), 77, dtype=ad.dtype, device="cuda")
                    acpp.transform_reduce(q, scratch, ad[off:off + n], bd[off:off + n], out, dtype(5),
                                          reduce_op=op, transform=code)
                    q.wait()
                    want = O.transform_reduce_binary(a[off:off + n], b[off:off + n], dtype(5), op, code)
                    assert host(out)[0] == want, (dtype, n, code, op, off)
    # empty input leaves out untouched
    out = torch.full((1,), 77, dtype=ad.dtype, device="cuda")
    acpp.transform_reduce(q, scratch, ad[:0], bd[:0], out, dtype(5), reduce_op=acpp.MAX, transform=acpp.B_ABS_DIFF)
    q.wait()
    assert host(out)[0] == 77


def test_binary_transform_reduce_rejects(gpu, q, scratch):
    torch, _ = gpu
    x = torch.zeros(8, dtype=torch.float32, device="cuda")
    o = torch.zeros(1, dtype=torch.float32, device="cuda")
    with pytest.raises(acpp.AcppError):
        acpp.transform_reduce(q, scratch, x, x, o, 0.0, transform=acpp.B_BIT_AND, reduce_op=acpp.PLUS)
    xi = torch.zeros(8, dtype=torch.int32, device="cuda")
    oi = torch.zeros(1, dtype=torch.int32, device="cuda")
    with pytest.raises(acpp.AcppError):
        acpp.transform_reduce(q, scratch, xi, xi, oi, 0, transform=acpp.B_MINUS, reduce_op=acpp.BIT_XOR)
