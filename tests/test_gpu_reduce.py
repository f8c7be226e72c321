"""GPU parity: reduce / sycl::reduction / transform_reduce against the oracle.

Integer results must be bit-exact.  Floating-point results are compared with
the FP64 oracle with the bound stated in each test.
"""
import math

import numpy as np
import pytest

import opensycl_amd as acpp
import oracle as O
from gpu_util import (ALL_OPS, FP_OPS, INT_TYPES, compensated_bound, dev, fold_bound, host, op_input,
                      rand, reduce_depth)

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 63, 64, 65, 127, 1000, 4099, 65536 + 7, (1 << 20) + 3]


def _out(torch, dtype, value=0):
    return dev(torch, np.array([value], dtype=dtype))


@pytest.mark.parametrize("dtype", INT_TYPES)
def test_reduce_int_all_ops_exact(gpu, q, scratch, dtype):
    torch, _ = gpu
    for op in ALL_OPS:
        for n in SIZES:
            x = op_input(dtype, op, n, seed=n + op)
            init = dtype(3) if op != 4 else dtype(np.iinfo(dtype).max)
            xd = dev(torch, x)
            out = _out(torch, dtype)
            acpp.reduce(q, scratch, xd, out, init, op)
            q.wait()
            assert host(out)[0] == O.reduce(x, init, op), (dtype, op, n)


@pytest.mark.parametrize("offset", [1, 2, 3, 5])
def test_reduce_misaligned_views(gpu, q, scratch, offset):
    torch, _ = gpu
    for dtype in (np.int8, np.int16, np.int32, np.int64):
        x = rand(dtype, 100003 + offset, seed=offset)
        xd = dev(torch, x)[offset:]
        out = _out(torch, dtype)
        acpp.reduce(q, scratch, xd, out, dtype(0), acpp.PLUS)
        q.wait()
        assert host(out)[0] == O.reduce(x[offset:], dtype(0))


@pytest.mark.parametrize("cfg", [(64, 2, 0), (128, 4, 1), (256, 8, 2), (512, 4, 0), (1024, 2, 0),
                                 (256, 4, 1)])
def test_reduce_launch_configs(gpu, q, scratch, cfg):
    torch, _ = gpu
    x = rand(np.int32, (1 << 22) + 77, seed=1)
    xd = dev(torch, x)
    out = _out(torch, np.int32)
    acpp.reduce(q, scratch, xd, out, np.int32(11), acpp.PLUS, config=cfg)
    q.wait()
    assert host(out)[0] == O.reduce(x, np.int32(11))


def test_reduce_empty_leaves_out_untouched(gpu, q, scratch):
    torch, _ = gpu
    out = _out(torch, np.int32, 1234)
    acpp.reduce(q, scratch, torch.empty(0, dtype=torch.int32, device="cuda"), out, 5)
    q.wait()
    assert host(out)[0] == 1234


def test_golden_reference_reductions(gpu, q, scratch):
    # tests/sycl/reduction.cpp:250-285, 323-344; tests/pstl/reduce.cpp:24-68
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))
    torch, _ = gpu
    x = dev(torch, np.arange(1024, dtype=np.int32))
    out = _out(torch, np.int32, 777)
    acpp.reduction(q, scratch, x, out, acpp.PLUS, initialize_to_identity=True)
    q.wait()
    assert host(out)[0] == g["sycl_reduction_accessor"]["expected_sum"]
    acpp.reduction(q, scratch, x, out, acpp.MAX, initialize_to_identity=True)
    q.wait()
    assert host(out)[0] == g["sycl_reduction_accessor"]["expected_max"]
    out = _out(torch, np.int32, 0)
    acpp.reduction(q, scratch, x, out, acpp.PLUS)
    q.wait()
    assert host(out)[0] == g["sycl_reduction_incremental"]["expected_after_first"]
    acpp.reduction(q, scratch, x, out, acpp.PLUS)
    q.wait()
    assert host(out)[0] == g["sycl_reduction_incremental"]["expected_after_second"]
    for n_s, expected in g["sycl_reduction_multiplies_generator"]["expected_products"].items():
        i = np.arange(int(n_s))
        xm = np.where(i > 1500, 1, np.where(i % 100 == 0, 2, 1)).astype(np.int64)
        out = _out(torch, np.int64, -5)
        acpp.reduction(q, scratch, dev(torch, xm), out, acpp.MULTIPLY, initialize_to_identity=True)
        q.wait()
        assert host(out)[0] == expected
    for c in g["pstl_reduce"]["cases"]:
        dt = np.int64 if c.get("dtype") == "int64" else np.int32
        out = _out(torch, dt, -1)
        acpp.reduce(q, scratch, dev(torch, np.arange(c["n"], dtype=dt)), out, dt(c["init"]))
        q.wait()
        expected = c["expected"] if c["n"] > 0 else -1  # empty input: *out untouched
        assert host(out)[0] == expected


@pytest.mark.parametrize("dtype", [np.int8, np.int32, np.uint64, np.float32, np.float64])
def test_reduction_semantics(gpu, q, scratch, dtype):
    torch, _ = gpu
    for op in (FP_OPS if np.issubdtype(dtype, np.floating) else ALL_OPS):
        for n in (0, 1, 20, 128, 1000, 16384):
            x = op_input(dtype, op, n, seed=n)
            start = dtype(5)
            for to_id in (False, True):
                out = _out(torch, dtype, start)
                acpp.reduction(q, scratch, dev(torch, x), out, op, initialize_to_identity=to_id)
                q.wait()
                got = host(out)[0]
                if np.issubdtype(dtype, np.floating) and op in (0, 1):
                    # FP64 reference within the kernel's fold bound
                    eps = 2.0 ** -24 if dtype == np.float32 else 2.0 ** -53
                    depth = reduce_depth(n, acpp.compute_units()) + 1
                    x64 = x.astype(np.float64)
                    if op == 0:
                        exact = math.fsum(x64) + (0.0 if to_id else float(start))
                        bound = depth * eps * (np.abs(x64).sum() + abs(float(start)))
                    else:
                        # every one of the n multiplications contributes a
                        # relative rounding error: |err| <= (n + 1) eps |prod|
                        exact = float(np.prod(x64)) * (1.0 if to_id else float(start))
                        bound = 1.01 * (n + 1) * eps * abs(exact)
                    assert abs(float(got) - exact) <= bound + 1e-300, (op, n, to_id)
                else:
                    assert got == O.reduction(x, start, op, to_id), (op, n, to_id)


def fp_sum_bound(x64_abs_sum, n, eps):
    return fold_bound(x64_abs_sum, reduce_depth(n, acpp.compute_units()), eps)


def comp_bound(exact, abs_sum, n):
    return compensated_bound(exact, abs_sum, n, acpp.compute_units())


@pytest.mark.parametrize("n", [1, 1000, (1 << 20) + 5, 1 << 24])
def test_reduce_f32_plus_vs_fp64(gpu, q, scratch, n):
    torch, _ = gpu
    x = rand(np.float32, n, seed=n)
    out = _out(torch, np.float32)
    acpp.reduce(q, scratch, dev(torch, x), out, np.float32(0.5))
    q.wait()
    exact = O.sum_fp64(x) + 0.5
    bound = fp_sum_bound(O.abs_sum_fp64(x) + 0.5, n, 2.0 ** -24) + abs(exact) * 2 ** -24
    assert abs(float(host(out)[0]) - exact) <= bound
    # deterministic: same launch geometry -> same bits
    out2 = _out(torch, np.float32)
    acpp.reduce(q, scratch, dev(torch, x), out2, np.float32(0.5))
    q.wait()
    assert host(out)[0].tobytes() == host(out2)[0].tobytes()


def test_reduce_fp_minmax_exact(gpu, q, scratch):
    torch, _ = gpu
    for dtype in (np.float32, np.float64):
        x = rand(dtype, 300007, seed=3)
        for op, init in ((acpp.MIN, dtype(np.inf)), (acpp.MAX, dtype(-np.inf))):
            out = _out(torch, dtype)
            acpp.reduce(q, scratch, dev(torch, x), out, init, op)
            q.wait()
            assert host(out)[0] == O.reduce(x, init, op)
    # all-negative maximum: the true maximum, not numeric_limits<float>::min()
    x = -np.arange(1, 1000, dtype=np.float32)
    out = _out(torch, np.float32)
    acpp.reduce(q, scratch, dev(torch, x), out, np.float32(-1e30), acpp.MAX)
    q.wait()
    assert host(out)[0] == -1.0


def test_reduce_f64_plus(gpu, q, scratch):
    torch, _ = gpu
    x = rand(np.float64, 1 << 20, seed=4)
    out = _out(torch, np.float64)
    acpp.reduce(q, scratch, dev(torch, x), out, 0.0)
    q.wait()
    exact = math.fsum(x)
    assert abs(host(out)[0] - exact) <= fp_sum_bound(np.abs(x).sum(), x.size, 2.0 ** -53)


def _stress_f32(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n).astype(np.float32)
    idx = rng.random(n) < 0.01
    x[idx] *= np.float32(1e4)
    return x


def _bf16_bits(x):
    return (x.view(np.uint32) >> 16).astype(np.uint16)


@pytest.mark.parametrize("n", [1, 777, (1 << 20) + 9, 1 << 24])
def test_compensated_sum_f32_and_bf16(gpu, q, scratch, n):
    torch, _ = gpu
    x = _stress_f32(n, n)
    exact = O.sum_fp64(x)
    abs_sum = O.abs_sum_fp64(x)
    out = _out(torch, np.float32)
    acpp.reduce_compensated(q, scratch, dev(torch, x), out, 0.0)
    q.wait()
    err = abs(float(host(out)[0]) - exact)
    # one final rounding to fp32 plus a second-order term
    assert err <= comp_bound(exact, abs_sum, n) + 1e-30
    b = _bf16_bits(x)
    exact_b = O.sum_fp64(b)
    bt = dev(torch, b).view(torch.bfloat16)
    out = _out(torch, np.float32)
    acpp.reduce_compensated(q, scratch, bt, out, 0.0)
    q.wait()
    assert abs(float(host(out)[0]) - exact_b) <= comp_bound(exact_b, abs_sum, n) + 1e-30
    # plain fp32 accumulation of bf16 input
    out = _out(torch, np.float32)
    acpp.reduce(q, scratch, bt, out, 0.0)
    q.wait()
    assert abs(float(host(out)[0]) - exact_b) <= fp_sum_bound(abs_sum, n, 2.0 ** -24)


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.uint32, np.uint64])
def test_dot_int_exact(gpu, q, scratch, dtype):
    torch, _ = gpu
    for n in (1, 100, 65537, 1 << 20):
        a = rand(dtype, n, seed=n)
        b = rand(dtype, n, seed=n + 1)
        out = _out(torch, dtype)
        acpp.transform_reduce(q, scratch, dev(torch, a), dev(torch, b), out, dtype(9))
        q.wait()
        assert host(out)[0] == O.dot(a, b, dtype(9))


def test_golden_babelstream_dot_partials(gpu, q, scratch):
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "reference_vectors.json")))["babelstream_dot"]
    torch, _ = gpu
    a = np.arange(1024, dtype=np.int32)
    gsz = g["num_groups"] * g["local_size"]
    got = []
    for grp in range(g["num_groups"]):
        idx = np.array([j for j in range(1024) if (j % gsz) // g["local_size"] == grp])
        out = _out(torch, np.int32)
        acpp.transform_reduce(q, scratch, dev(torch, a[idx]), dev(torch, a[idx]), out, 0)
        q.wait()
        got.append(int(host(out)[0]))
    assert got == g["expected_group_partials"]


@pytest.mark.parametrize("n", [3, 4097, (1 << 20) + 1, 1 << 24])
def test_dot_fp(gpu, q, scratch, n):
    torch, _ = gpu
    a = _stress_f32(n, 2 * n)
    b = rand(np.float32, n, seed=n)
    exact = O.dot_fp64(a, b)
    abs_sum = float(np.abs(a.astype(np.float64) * b.astype(np.float64)).sum())
    ad, bd = dev(torch, a), dev(torch, b)
    out = _out(torch, np.float32)
    acpp.transform_reduce(q, scratch, ad, bd, out, 0.0)
    q.wait()
    assert abs(float(host(out)[0]) - exact) <= fp_sum_bound(abs_sum, n, 2.0 ** -24)
    out = _out(torch, np.float32)
    acpp.transform_reduce(q, scratch, ad, bd, out, 0.0, compensated=True)
    q.wait()
    assert abs(float(host(out)[0]) - exact) <= comp_bound(exact, abs_sum, n)
    ab, bb = _bf16_bits(a), _bf16_bits(b)
    exact_b = O.dot_fp64(ab, bb)
    abt = dev(torch, ab).view(torch.bfloat16)
    bbt = dev(torch, bb).view(torch.bfloat16)
    out = _out(torch, np.float32)
    acpp.transform_reduce(q, scratch, abt, bbt, out, 0.0, compensated=True)
    q.wait()
    assert abs(float(host(out)[0]) - exact_b) <= comp_bound(exact_b, abs_sum, n)
    out = _out(torch, np.float32)
    acpp.transform_reduce(q, scratch, abt, bbt, out, 0.0)
    q.wait()
    assert abs(float(host(out)[0]) - exact_b) <= fp_sum_bound(abs_sum, n, 2.0 ** -24)


def test_dot_misaligned_operands(gpu, q, scratch):
    torch, _ = gpu
    a = rand(np.int32, 50001, seed=1)
    b = rand(np.int32, 50002, seed=2)
    ad = dev(torch, a)
    bd = dev(torch, b)[1:]
    out = _out(torch, np.int32)
    acpp.transform_reduce(q, scratch, ad, bd, out, 0)
    q.wait()
    assert host(out)[0] == O.dot(a, b[1:], np.int32(0))


def test_reduce_large_int32_exact(gpu, q, scratch):
    # 2^28 int32 uniform in [-2^15, 2^15]: the bandwidth-scale workload
    torch, _ = gpu
    n = 1 << 28
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    xd = torch.randint(-(1 << 15), (1 << 15) + 1, (n,), dtype=torch.int32, device="cuda",
                       generator=g)
    out = _out(torch, np.int32)
    acpp.reduce(q, scratch, xd, out, np.int32(0))
    q.wait()
    assert host(out)[0] == O.reduce(host(xd), np.int32(0))
