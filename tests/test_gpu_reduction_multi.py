"""GPU parity: several sycl::reduction objects in one kernel
(acpp_reduction_multi_<t>), after tests/sycl/reduction.cpp:116-208
(test_two_reductions: a plus and a multiplies reduction over 128*128
elements in one parallel_for, initialize_to_identity), plus aliasing
inputs, accumulate mode, empty ranges and unaligned inputs; checked against
the oracle's sycl::reduction fold."""
import json
import os

import numpy as np
import pytest

import opensycl_amd as acpp
import oracle as O
from gpu_util import dev, host, fold_bound, reduce_depth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))


def _mul_gen(n):
    i = np.arange(n)
    return np.where(i > 1500, 1, np.where(i % 100 == 0, 2, 1))


@pytest.mark.parametrize("dtype", [np.uint32, np.int32, np.int64, np.float32, np.float64])
def test_two_reductions_reference_case(gpu, q, scratch, dtype):
    torch, _ = gpu
    n = 128 * 128
    a = np.arange(n).astype(dtype)            # input_generator<T, plus>
    b = _mul_gen(n).astype(dtype)             # input_generator<T, multiplies>
    ad, bd = dev(torch, a), dev(torch, b)
    o0 = dev(torch, np.array([dtype(123)]))
    o1 = dev(torch, np.array([dtype(45)]))
    acpp.reduction_multi(q, scratch, [ad, bd], [o0, o1], [acpp.PLUS, acpp.MULTIPLY],
                         initialize_to_identity=True)
    q.wait()
    exp_sum = O.reduction(a, dtype(0), O.PLUS, True)
    exp_mul = O.reduction(b, dtype(0), O.MULTIPLY, True)
    assert exp_mul == GOLDEN["sycl_reduction_multiplies_generator"]["expected_products"]["16384"]
    if np.issubdtype(dtype, np.floating):
        assert abs(host(o0)[0] - exp_sum) <= 1e-3 * abs(exp_sum)   # the reference's tolerance
        assert host(o1)[0] == exp_mul
    else:
        assert host(o0)[0] == exp_sum and host(o1)[0] == exp_mul


@pytest.mark.parametrize("n", [0, 1, 1000, (1 << 22) + 7])
def test_four_reductions_aliasing_and_accumulate(gpu, q, scratch, n):
    torch, _ = gpu
    rng = np.random.default_rng(n)
    x = rng.integers(-(1 << 40), 1 << 40, n, dtype=np.int64)
    xd = dev(torch, x)
    ops = [acpp.PLUS, acpp.MIN, acpp.MAX, acpp.BIT_XOR]
    starts = [np.int64(5), np.int64(-7), np.int64(9), np.int64(0x33)]
    for init_id in (True, False):
        outs = [dev(torch, np.array([s])) for s in starts]
        acpp.reduction_multi(q, scratch, [xd, xd, xd, xd], outs, ops, initialize_to_identity=init_id)
        q.wait()
        for o, op, s in zip(outs, ops, starts):
            assert host(o)[0] == O.reduction(x, s, op, init_id), (n, op, init_id)


# 1- and 2-byte elements (16 or 8 per 16-byte load; sums and products wrap),
# every operator, aligned views and views 1 or 3 elements off (scalar head).
@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.uint16])
def test_narrow_types(gpu, q, scratch, dtype):
    torch, _ = gpu
    info = np.iinfo(dtype)
    rng = np.random.default_rng(3)
    for n in (1, 1000, (1 << 22) + 7):
        x = rng.integers(info.min, info.max + 1, n + 3).astype(dtype)
        xd_full = dev(torch, x)
        for off in (0, 1, 3):
            xs, xd = x[off:off + n], xd_full[off:off + n]
            for ops in ([acpp.PLUS, acpp.MIN, acpp.MAX, acpp.BIT_XOR],
                        [acpp.MULTIPLY, acpp.BIT_AND, acpp.BIT_OR, acpp.LOGICAL_OR]):
                starts = [dtype(5), dtype(7), dtype(9), dtype(1)]
                for init_id in (True, False):
                    outs = [dev(torch, np.array([s])) for s in starts]
                    acpp.reduction_multi(q, scratch, [xd] * 4, outs, ops, initialize_to_identity=init_id)
                    q.wait()
                    for o, op, s in zip(outs, ops, starts):
                        assert host(o)[0] == O.reduction(xs, s, op, init_id), (dtype, n, off, op, init_id)


@pytest.mark.parametrize("offset", [0, 1, 3])
def test_unaligned_float_inputs(gpu, q, scratch, offset):
    torch, _ = gpu
    rng = np.random.default_rng(offset)
    n = 3 * 1024 * 1024 + 5
    a = rng.standard_normal(n + offset).astype(np.float32)
    ad = dev(torch, a)[offset:]
    o = [torch.zeros(1, dtype=torch.float32, device="cuda") for _ in range(3)]
    acpp.reduction_multi(q, scratch, [ad, ad, ad], o, [acpp.PLUS, acpp.MIN, acpp.MAX], True)
    q.wait()
    x = a[offset:]
    s = O.sum_fp64(x)
    # 512 work-items x 4 accumulators per CU, like the single reduction
    bound = fold_bound(np.abs(x.astype(np.float64)).sum(), reduce_depth(n, acpp.compute_units()),
                       2.0 ** -24)
    assert abs(host(o[0])[0] - s) <= bound
    assert host(o[1])[0] == x.min() and host(o[2])[0] == x.max()


def test_rejects_bad_arguments(gpu, q, scratch):
    torch, _ = gpu
    x = torch.zeros(10, dtype=torch.float32, device="cuda")
    o = torch.zeros(1, dtype=torch.float32, device="cuda")
    with pytest.raises(acpp.AcppError):
        acpp.reduction_multi(q, scratch, [x], [o], [acpp.BIT_AND])
    with pytest.raises(ValueError):
        acpp.reduction_multi(q, scratch, [x] * 5, [o] * 5, [acpp.PLUS] * 5)
