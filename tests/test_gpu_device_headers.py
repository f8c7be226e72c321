"""GPU parity: group collectives and atomic_ref from the device headers.

The kernels live in tests/csrc/device_kernels.hip (built into
tests/lib/libacpp_test_kernels.so) and follow the reference's own tests:
group sizes 25/144/256 over 100/576/1024 work-items
(tests/sycl/group_functions/group_functions.hpp:254-301), the CBS golden
sums (tests/compiler/cbs/reduce_for.cpp), the babelstream dot partials
(tests/compiler/cbs/babelstream_dot.cpp), the 2048-element atomic folds
(tests/sycl/atomic.cpp:81-123) and the std::atomic cases
(tests/pstl/std_atomic.cpp:30-136).
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from gpu_util import ALL_OPS, dev, host, op_input

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))
TAGS = {np.int8: 0, np.int32: 1, np.int64: 2, np.uint32: 3, np.float32: 4, np.float64: 5,
        np.float16: 6}
ATOMIC_TAGS = {np.int32: 1, np.int64: 2, np.uint32: 3, np.float32: 4, np.float64: 5, np.uint64: 6}
CASES = [(25, 100), (144, 576), (256, 1024), (64, 4096), (1024, 8192), (200, 2000)]


@pytest.fixture(scope="module")
def tk():
    path = os.path.join(HERE, "lib", "libacpp_test_kernels.so")
    if not os.path.exists(path):
        pytest.fail("tests/lib/libacpp_test_kernels.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.tk_group_collective.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_uint64,
                                        ctypes.c_uint32, ctypes.c_int, vp, vp]
    lib.tk_joint.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_uint64, ctypes.c_uint64,
                             ctypes.c_uint32, ctypes.c_int, vp, vp]
    lib.tk_babelstream_dot.argtypes = [vp, vp, vp, ctypes.c_uint64, ctypes.c_uint32,
                                       ctypes.c_uint32, vp]
    lib.tk_atomic_fold.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_uint64, ctypes.c_int, vp]
    lib.tk_atomic_single.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
    lib.tk_atomic_fold_narrow.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_uint64,
                                          ctypes.c_uint32, vp]
    lib.tk_atomic_single_narrow.argtypes = [ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, vp, vp]
    lib.tk_reducer.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_uint64, vp, ctypes.c_uint32,
                               ctypes.c_uint32, ctypes.c_int, vp]
    lib.tk_atomic_fold_local.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_uint32,
                                         ctypes.c_uint32, vp]
    lib.tk_atomic_pointer.argtypes = [vp, vp, ctypes.c_uint64, vp]
    return lib


def _stream(torch):
    return torch.cuda.current_stream().cuda_stream


def _collective(torch, tk, x, local, kind, op, init=None):
    xd = dev(torch, x)
    yd = torch.empty_like(xd)
    init_c = None
    if init is not None:
        init_c = np.array([init], dtype=x.dtype)
    rc = tk.tk_group_collective(TAGS[x.dtype.type], kind, xd.data_ptr(), yd.data_ptr(), x.size,
                                local, op, init_c.ctypes.data if init_c is not None else None,
                                _stream(torch))
    assert rc == 0
    torch.cuda.synchronize()
    return host(yd)


def _expected(x, local, kind, op, init):
    """Oracle for each kind; sub-group kinds split the work-group into waves."""
    if kind in (3, 4, 5, 13, 14, 15):
        # sub-groups: consecutive chunks of 64 within each work-group
        out = np.empty_like(x)
        for b in range(0, x.size, local):
            for w in range(b, b + local, 64):
                e = min(w + 64, b + local)
                seg = x[w:e]
                if kind in (3, 13):
                    out[w:e] = O.group_reduce(seg, seg.size, op)
                elif kind in (4, 14):
                    out[w:e] = O.inclusive_scan(seg, op)
                else:
                    out[w:e] = O.exclusive_scan(seg, init, op)
        return out
    if kind in (0, 10):
        return O.group_reduce(x, local, op)
    if kind in (1, 11):
        out = np.empty_like(x)
        for b in range(0, x.size, local):
            out[b:b + local] = O.inclusive_scan(x[b:b + local], op)
        return out
    if kind in (2, 12, 8):
        out = np.empty_like(x)
        for b in range(0, x.size, local):
            out[b:b + local] = O.exclusive_scan(x[b:b + local], init, op)
        return out
    if kind == 6:
        return _apply_right(O.group_reduce(x, local, op), init, op)
    if kind == 7:
        out = np.empty_like(x)
        for b in range(0, x.size, local):
            out[b:b + local] = O.inclusive_scan(x[b:b + local], op, init)
        return out
    raise AssertionError(kind)


def _apply_right(r, init, op):
    # reduce_over_group(g, x, init, op) == op(reduction, init)
    return np.array([O.reduce(np.array([init], dtype=r.dtype), v, op) for v in r], dtype=r.dtype)


def _inputs(dtype, n, op, seed):
    rng = np.random.default_rng(seed)
    if np.issubdtype(dtype, np.floating):
        if op == 1:
            return (1.0 + rng.standard_normal(n) * 1e-3).astype(dtype)
        return np.round(rng.standard_normal(n) * 64).astype(dtype)  # exactly summable
    info = np.iinfo(dtype)
    x = rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
    if op in (7, 8):
        x[rng.random(n) < 0.5] = 0
    return x


@pytest.mark.parametrize("dtype", [np.int8, np.int32, np.int64, np.uint32, np.float32, np.float64])
def test_group_collectives(gpu, tk, dtype):
    torch, _ = gpu
    ops = [0, 1, 2, 3] if np.issubdtype(dtype, np.floating) else list(range(9))
    for local, glob in CASES:
        for op in ops:
            x = _inputs(dtype, glob, op, seed=local * 10 + op)
            init = dtype(3) if op != 4 else dtype(-1) if np.issubdtype(dtype, np.signedinteger) else dtype(7)
            for kind in (0, 1, 2, 3, 4, 5, 6, 7, 10, 11, 12, 13, 14, 15):
                got = _collective(torch, tk, x, local, kind, op, init)
                exp = _expected(x, local, kind, op, init)
                if np.issubdtype(dtype, np.floating) and op == 1:
                    # products: every multiplication rounds (|err| <= n eps |p|)
                    eps = np.finfo(dtype).eps
                    ok = np.all(np.abs(got.astype(np.float64) - exp.astype(np.float64))
                                <= 2 * local * eps * np.abs(exp.astype(np.float64)) + 1e-300)
                    assert ok, (dtype, local, op, kind)
                else:
                    assert np.array_equal(got, exp), (dtype, local, op, kind)


def test_group_collectives_f16(gpu, tk):
    """The f16 builtins (__acpp_sscp_*_group_*_f16) and the templates on
    _Float16.  Sums use integers of magnitude <= 8, whose partial sums stay
    below 2048 and are exact in half precision in any order, so they and
    min/max compare bit-exactly against the float32 oracle rounded to half;
    products of values near 1 round once per multiplication and compare
    within 2 * local * eps(half) relative."""
    torch, _ = gpu
    eps = float(np.finfo(np.float16).eps)
    for local, glob in CASES:
        for op in (0, 1, 2, 3):
            rng = np.random.default_rng(local * 10 + op)
            if op == 1:
                x = (1.0 + rng.standard_normal(glob) * 1e-2).astype(np.float16)
            else:
                x = np.clip(np.round(rng.standard_normal(glob) * 3), -8, 8).astype(np.float16)
            init = np.float16(3)
            for kind in (0, 1, 2, 3, 4, 5, 6, 7, 10, 11, 12, 13, 14, 15):
                got = _collective(torch, tk, x, local, kind, op, init).astype(np.float64)
                exp = _expected(x.astype(np.float32), local, kind, op, np.float32(init))
                exp = exp.astype(np.float16).astype(np.float64)
                if op == 1:
                    assert np.all(np.abs(got - exp) <= 2 * local * eps * np.abs(exp)), (local, kind)
                else:
                    assert np.array_equal(got, exp), (local, op, kind)


def test_exclusive_scan_known_identity(gpu, tk):
    torch, _ = gpu
    x = _inputs(np.int32, 1024, 0, 1)
    got = _collective(torch, tk, x, 256, 8, 0)
    exp = _expected(x, 256, 2, 0, np.int32(0))
    assert np.array_equal(got, exp)


def test_golden_cbs_group_sums(gpu, tk):
    torch, _ = gpu
    g = GOLDEN["cbs_reduce_for"]
    x = np.arange(1024, dtype=np.int32)
    got = _collective(torch, tk, x, g["local_size"], 0, 0)
    assert list(got[::g["local_size"]]) == g["expected_group_sums"]


def test_golden_babelstream_dot(gpu, tk):
    torch, _ = gpu
    g = GOLDEN["babelstream_dot"]
    a = dev(torch, np.arange(1024, dtype=np.int32))
    sums = torch.zeros(g["num_groups"], dtype=torch.int32, device="cuda")
    assert tk.tk_babelstream_dot(a.data_ptr(), a.data_ptr(), sums.data_ptr(), 1024,
                                 g["num_groups"], g["local_size"], _stream(torch)) == 0
    torch.cuda.synchronize()
    assert list(host(sums)) == g["expected_group_partials"]


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.float32])
def test_joint_algorithms(gpu, tk, dtype):
    torch, _ = gpu
    tag = TAGS[dtype]
    for local, seg in ((256, 1000), (64, 64), (128, 5000), (1024, 3000)):
        nseg = 7
        x = _inputs(dtype, seg * nseg, 0, seed=seg)
        init = dtype(5)
        for kind in (0, 1, 2, 3, 4):
            xd = dev(torch, x)
            yd = torch.zeros_like(xd)
            init_c = np.array([init], dtype=dtype)
            assert tk.tk_joint(tag, kind, xd.data_ptr(), yd.data_ptr(), x.size, seg, local, 0,
                               init_c.ctypes.data, _stream(torch)) == 0
            torch.cuda.synchronize()
            y = host(yd)
            for s in range(nseg):
                xs = x[s * seg:(s + 1) * seg]
                ys = y[s * seg:(s + 1) * seg]
                if kind == 0:
                    assert ys[0] == O.reduce(xs, init), (local, seg, kind)
                elif kind == 4:
                    assert ys[0] == O.reduce(xs, dtype(0)), (local, seg, kind)
                elif kind == 1:
                    assert np.array_equal(ys, O.inclusive_scan(xs)), (local, seg, kind)
                elif kind == 2:
                    assert np.array_equal(ys, O.exclusive_scan(xs, init)), (local, seg, kind)
                else:
                    assert np.array_equal(ys, O.inclusive_scan(xs, 0, init)), (local, seg, kind)


RMW = {"and": 0, "or": 1, "xor": 2, "add": 3, "sub": 4, "min": 5, "max": 6}


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.uint32, np.uint64, np.float32,
                                   np.float64])
def test_atomic_ref_fold(gpu, tk, dtype):
    # tests/sycl/atomic.cpp:81-123: element i = -i converted to T
    torch, _ = gpu
    n = 2048
    vals = (-np.arange(n, dtype=np.int64)).astype(dtype)
    names = ["add", "sub", "min", "max"] + ([] if np.issubdtype(dtype, np.floating)
                                            else ["and", "or", "xor"])
    for name in names:
        d = dev(torch, vals)
        assert tk.tk_atomic_fold(ATOMIC_TAGS[dtype], RMW[name], d.data_ptr(), n, 0,
                                 _stream(torch)) == 0
        torch.cuda.synchronize()
        exp = O.atomic_fold({"and": O.RMW_AND, "or": O.RMW_OR, "xor": O.RMW_XOR,
                             "add": O.RMW_ADD, "sub": O.RMW_SUB, "min": O.RMW_MIN,
                             "max": O.RMW_MAX}[name], vals[1:], vals[0])
        assert host(d)[0] == exp, (dtype, name)
    if dtype == np.int32:
        g = GOLDEN["sycl_atomic_fetch_op"]["expected_int32"]
        for name, expected in g.items():
            d = dev(torch, vals)
            tk.tk_atomic_fold(1, RMW[name], d.data_ptr(), n, 0, _stream(torch))
            torch.cuda.synchronize()
            assert host(d)[0] == expected


@pytest.mark.parametrize("dtype", [np.int32, np.float32])
def test_atomic_builtin_abi_fold(gpu, tk, dtype):
    torch, _ = gpu
    n = 2048
    vals = (-np.arange(n, dtype=np.int64)).astype(dtype)
    names = ["add", "sub", "min", "max"] + ([] if dtype == np.float32 else ["and", "or", "xor"])
    for name in names:
        d = dev(torch, vals)
        assert tk.tk_atomic_fold(ATOMIC_TAGS[dtype], RMW[name], d.data_ptr(), n, 1,
                                 _stream(torch)) == 0
        torch.cuda.synchronize()
        exp = O.atomic_fold({"and": O.RMW_AND, "or": O.RMW_OR, "xor": O.RMW_XOR,
                             "add": O.RMW_ADD, "sub": O.RMW_SUB, "min": O.RMW_MIN,
                             "max": O.RMW_MAX}[name], vals[1:], vals[0])
        assert host(d)[0] == exp, (dtype, name)


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.uint32, np.uint64])
def test_std_atomic_cases(gpu, tk, dtype):
    # tests/pstl/std_atomic.cpp: initial 42; fetch ops return the previous value
    torch, _ = gpu
    g = GOLDEN["std_atomic_fetch"]
    for c in g["cases"]:
        d = dev(torch, np.array([g["initial"]], dtype=dtype))
        prev = torch.zeros(5, dtype=d.dtype, device="cuda")
        operand = np.array([c["operand"]], dtype=dtype)
        assert tk.tk_atomic_single(ATOMIC_TAGS[dtype], RMW[c["op"]], d.data_ptr(),
                                   operand.ctypes.data, prev.data_ptr(), _stream(torch)) == 0
        torch.cuda.synchronize()
        assert host(d)[0] == c["expected"]
        assert host(prev)[0] == g["initial"]
    # fetch_max / fetch_min (commented out as C++26 in the reference test)
    for op, operand, expected in ((6, 45, 45), (5, 41, 41)):
        d = dev(torch, np.array([42], dtype=dtype))
        prev = torch.zeros(5, dtype=d.dtype, device="cuda")
        tk.tk_atomic_single(ATOMIC_TAGS[dtype], op, d.data_ptr(),
                            np.array([operand], dtype=dtype).ctypes.data, prev.data_ptr(),
                            _stream(torch))
        torch.cuda.synchronize()
        assert host(d)[0] == expected and host(prev)[0] == 42
    # load/store 42 -> 43, exchange -> 43, compare_exchange (expected 0 fails,
    # expected 42 succeeds, data becomes 12)
    for op, operand, expected in ((7, 0, 43), (8, 0, 43), (9, 12, 12)):
        d = dev(torch, np.array([42], dtype=dtype))
        prev = torch.zeros(5, dtype=d.dtype, device="cuda")
        tk.tk_atomic_single(ATOMIC_TAGS[dtype], op, d.data_ptr(),
                            np.array([operand], dtype=dtype).ctypes.data, prev.data_ptr(),
                            _stream(torch))
        torch.cuda.synchronize()
        assert host(d)[0] == expected, op
        if op == 9:
            p = host(prev)
            assert list(p[1:5]) == [42, 42, 0, 1]


NARROW_TAGS = {np.int8: 7, np.int16: 8, np.uint8: 9, np.uint16: 10}


@pytest.mark.parametrize("dtype", [np.int8, np.int16, np.uint8, np.uint16])
@pytest.mark.parametrize("slots", [1, 3, 4])
def test_narrow_atomic_builtins_fold(gpu, tk, dtype, slots):
    # __acpp_sscp_atomic_fetch_<op>_{i8,i16,u8,u16} (sscp/builtins/atomic.hpp:
    # 110-311): 4096 work-items fold their element into data[i % slots];
    # with 3 or 4 slots, neighbouring elements of one 32-bit word change
    # concurrently and must not disturb each other.
    torch, _ = gpu
    n = 4096
    rng = np.random.default_rng(slots)
    info = np.iinfo(dtype)
    vals = rng.integers(info.min, info.max, n, endpoint=True, dtype=dtype)
    names = ["add", "sub", "min", "max"]
    if dtype in (np.int8, np.int16):
        names += ["and", "or", "xor"]
    for name in names:
        d = dev(torch, vals)
        assert tk.tk_atomic_fold_narrow(NARROW_TAGS[dtype], RMW[name], d.data_ptr(), n, slots,
                                        _stream(torch)) == 0
        torch.cuda.synchronize()
        got = host(d)
        code = {"and": O.RMW_AND, "or": O.RMW_OR, "xor": O.RMW_XOR, "add": O.RMW_ADD,
                "sub": O.RMW_SUB, "min": O.RMW_MIN, "max": O.RMW_MAX}[name]
        for k in range(slots):
            exp = O.atomic_fold(code, vals[slots + k::slots], vals[k])
            assert got[k] == exp, (dtype, name, k)
        assert np.array_equal(got[slots:], vals[slots:])  # operands untouched


@pytest.mark.parametrize("dtype", [np.int8, np.int16])
def test_narrow_atomic_builtins_single(gpu, tk, dtype):
    # load / store / exchange / compare-exchange on one element of a word
    # whose other elements must keep their values
    torch, _ = gpu
    for k in range(4):
        init = np.array([10, 20, 30, 40], dtype=dtype)
        d = dev(torch, init)
        prev = torch.zeros(6, dtype=d.dtype, device="cuda")
        assert tk.tk_atomic_single_narrow(NARROW_TAGS[dtype], d.data_ptr(), k, -7, prev.data_ptr(),
                                          _stream(torch)) == 0
        torch.cuda.synchronize()
        x0 = int(init[k])
        # load -> x0; store x0 + 1; exchange(1) -> x0 + 1; store x0 + 2;
        # cas(expected 0) fails and sees x0 + 2; cas(expected x0 + 2) stores -7
        assert list(host(prev)) == [x0, x0 + 1, x0 + 2, x0 + 2, 0, 1], (dtype, k)
        want = init.copy()
        want[k] = -7
        assert np.array_equal(host(d), want), (dtype, k)


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.uint32, np.uint64, np.float32,
                                   np.float64])
def test_atomic_ref_local_space_fold(gpu, tk, dtype):
    # atomic_ref<T, relaxed, work_group, local_space> on an LDS value
    # (atomic_ref.hpp:58-61 allows the local space): each work-group folds
    # its own slice, starting from the slice's first element.
    torch, _ = gpu
    groups, local = 37, 256
    vals = (-np.arange(groups * local, dtype=np.int64) % 1000 - 500).astype(dtype)
    names = ["add", "sub", "min", "max"] + ([] if np.issubdtype(dtype, np.floating)
                                            else ["and", "or", "xor"])
    for name in names:
        d = dev(torch, vals)
        out = torch.empty(groups, dtype=d.dtype, device="cuda")
        assert tk.tk_atomic_fold_local(ATOMIC_TAGS[dtype], RMW[name], d.data_ptr(), out.data_ptr(), groups,
                                       local, _stream(torch)) == 0
        torch.cuda.synchronize()
        got = host(out)
        rmw = {"and": O.RMW_AND, "or": O.RMW_OR, "xor": O.RMW_XOR, "add": O.RMW_ADD, "sub": O.RMW_SUB,
               "min": O.RMW_MIN, "max": O.RMW_MAX}[name]
        for g in range(groups):
            sl = vals[g * local:(g + 1) * local]
            assert got[g] == O.atomic_fold(rmw, sl[1:], sl[0]), (dtype, name, g)


def test_atomic_ref_pointer(gpu, tk):
    # atomic_ref<int*>: fetch_add(1) moves the pointer by sizeof(int); every
    # work-item sees a distinct old value and the final one is base + n.
    torch, _ = gpu
    n = 100003
    base = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    slot = torch.tensor([base.data_ptr()], dtype=torch.int64, device="cuda")
    seen = torch.zeros(n, dtype=torch.int64, device="cuda")
    assert tk.tk_atomic_pointer(slot.data_ptr(), seen.data_ptr(), n, _stream(torch)) == 0
    torch.cuda.synchronize()
    assert int(slot.item()) == base.data_ptr() + 4 * n
    offs = (host(seen) - base.data_ptr()) // 4
    assert np.array_equal(np.sort(offs), np.arange(n))


# acpp::reducer inside a user kernel (sycl::reduction's device side,
# reduction.hpp:61-160): every work-item folds its grid-stride share, the
# group's total reaches *out through combine_into.  *out starts at a
# non-identity value, so the result is op(old, fold of x) as for a
# reduction without initialize_to_identity.
@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.uint32, np.uint64, np.float32, np.float64])
def test_reducer_in_user_kernel(gpu, tk, dtype):
    torch, _ = gpu
    ops = ALL_OPS if np.issubdtype(dtype, np.integer) else [0, 1, 2, 3, 7, 8]
    for op in ops:
        for n, groups, local in ((1, 1, 64), (1000, 3, 256), (100003, 40, 512), (1 << 20, 256, 1024)):
            x = op_input(dtype, op, n, seed=n + 17 * op)
            if np.issubdtype(dtype, np.floating) and op == 0:
                x = np.round(x * 64).astype(dtype)  # small integers: sums are exact
            old = dtype(3) if op not in (7, 8) else dtype(1)
            if np.issubdtype(dtype, np.floating) and op in (7, 8):
                # logical operators on floats give 0 or 1 (T(a && b)); the
                # oracle's reduce takes them for integers only
                nz = x != 0
                want = dtype(bool(old) and bool(nz.all())) if op == 7 else dtype(bool(old) or bool(nz.any()))
            else:
                want = O.reduce(x, old, op)
            for user_identity in (0, 1):
                out = dev(torch, np.array([old], dtype=dtype))
                xd = dev(torch, x)
                rc = tk.tk_reducer(ATOMIC_TAGS[dtype], op, xd.data_ptr(), n, out.data_ptr(), groups, local,
                                   user_identity, _stream(torch))
                assert rc == 0
                torch.cuda.synchronize()
                got = host(out)[0]
                if np.issubdtype(dtype, np.floating) and op == 1:
                    assert np.isclose(got, want, rtol=1e-3), (dtype, op, n, got, want)
                else:
                    assert got == want, (dtype, op, n, user_identity, got, want)
