"""GPU parity at the full workload sizes (2^28 elements; 65,536 segments).

The smaller parity tests cover every type, operator and edge case; these run
the bandwidth-scale workloads themselves through the same entry points and
check them against the oracle: bit-exact for integer work, within the stated
bound against an FP64 oracle for floating point.  Inputs are generated on the
GPU (seeded torch generators) and copied back for the oracle.
"""
import math

import numpy as np
import pytest

import opensycl_amd as acpp
import oracle as O
from gpu_util import host, lognormal_offsets

pytestmark = pytest.mark.gpu

N = 1 << 28


def _gen(torch, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return g


def _fp_bound(abs_sum, n, eps):
    # a fold of depth at most d roundings: |err| <= d * eps * sum|x|
    return (64 + 2 * int(math.log2(max(n, 2)))) * eps * abs_sum


def _stress(torch, seed):
    # N(0,1) with 1 % of the entries scaled by 1e4 (cancellation stress)
    g = _gen(torch, seed)
    x = torch.randn(N, generator=g, device="cuda", dtype=torch.float32)
    x[torch.rand(N, generator=g, device="cuda") < 0.01] *= 1e4
    return x


def test_dot_f32_2e28(gpu, q, scratch):
    # transform_reduce dot product of two 2^28 fp32 N(0,1) vectors
    torch, _ = gpu
    a = torch.randn(N, generator=_gen(torch, 1), device="cuda", dtype=torch.float32)
    b = torch.randn(N, generator=_gen(torch, 2), device="cuda", dtype=torch.float32)
    out = torch.zeros(1, dtype=torch.float32, device="cuda")
    acpp.transform_reduce(q, scratch, a, b, out, 0.0)
    comp = torch.zeros(1, dtype=torch.float32, device="cuda")
    acpp.transform_reduce(q, scratch, a, b, comp, 0.0, compensated=True)
    q.wait()
    ah, bh = host(a), host(b)
    exact = O.dot_fp64(ah, bh)
    abs_sum = float(np.abs(ah.astype(np.float64) * bh.astype(np.float64)).sum())
    assert abs(float(out.item()) - exact) <= _fp_bound(abs_sum, N, 2.0 ** -24)
    assert abs(float(comp.item()) - exact) <= 2.0 ** -24 * abs(exact) + 2.0 ** -40 * abs_sum


def test_precision_sums_2e28(gpu, q, scratch):
    # fp32 and bf16 sums of 2^28 stressed values: plain within the fold
    # bound, compensated within one fp32 rounding of the FP64 sum
    torch, _ = gpu
    x = _stress(torch, 3)
    xh = host(x)
    abs_sum = O.abs_sum_fp64(xh)
    exact = O.sum_fp64(xh)
    for compensated in (False, True):
        out = torch.zeros(1, dtype=torch.float32, device="cuda")
        if compensated:
            acpp.reduce_compensated(q, scratch, x, out, 0.0)
        else:
            acpp.reduce(q, scratch, x, out, 0.0)
        q.wait()
        err = abs(float(out.item()) - exact)
        if compensated:
            assert err <= 2.0 ** -24 * abs(exact) + 2.0 ** -40 * abs_sum
        else:
            assert err <= _fp_bound(abs_sum, N, 2.0 ** -24)
    xb = x.to(torch.bfloat16)
    exact_b = O.sum_fp64((host(xb.view(torch.int16)).view(np.uint16)))
    out = torch.zeros(1, dtype=torch.float32, device="cuda")
    acpp.reduce_compensated(q, scratch, xb, out, 0.0)
    q.wait()
    # bf16 rounding changes each |x| by at most 2^-8 relative
    assert abs(float(out.item()) - exact_b) <= 2.0 ** -24 * abs(exact_b) + 2.0 ** -40 * 1.01 * abs_sum


@pytest.mark.parametrize("num_bins", [4096, 65536])
def test_histogram_zipf_2e28(gpu, q, scratch, num_bins):
    # 2^28 uint32 keys drawn Zipf(1.1) on {0..num_bins-1} by inverse CDF
    torch, _ = gpu
    p = torch.arange(1, num_bins + 1, dtype=torch.float64, device="cuda") ** -1.1
    cdf = torch.cumsum(p, 0)
    cdf /= cdf[-1].clone()
    u = torch.rand(N, generator=_gen(torch, 4), device="cuda", dtype=torch.float64)
    keys = torch.clamp(torch.searchsorted(cdf, u, right=True), max=num_bins - 1).to(torch.int32)
    del u
    keys = keys.view(torch.uint32)
    for strategy in (acpp.ATOMICS_PRIVATIZED, acpp.ATOMICS_AUTO):
        bins = torch.zeros(num_bins, dtype=torch.uint32, device="cuda")
        acpp.histogram(q, scratch, keys, bins, strategy)
        q.wait()
        got = host(bins)
        assert int(got.astype(np.uint64).sum()) == N
        if strategy == acpp.ATOMICS_PRIVATIZED:
            expected = O.histogram(host(keys), num_bins)
        assert np.array_equal(got, expected), strategy


def test_fetch_max_2e28(gpu, q, scratch):
    # 2^28 uniform int32 folded by fetch_max into 1024 slots (slot = i mod 1024)
    torch, _ = gpu
    v = torch.randint(-(1 << 31), 1 << 31, (N,), generator=_gen(torch, 5), device="cuda",
                      dtype=torch.int32)
    t = torch.full((1024,), np.iinfo(np.int32).min, dtype=torch.int32, device="cuda")
    acpp.atomic_fetch_strided(q, scratch, "max", v, t, acpp.ATOMICS_PRIVATIZED)
    q.wait()
    got = host(t)
    assert np.array_equal(got, O.fetch_max_strided(host(v), 1024))
    assert np.array_equal(got, host(v.view(-1, 1024).max(dim=0).values))


def test_segmented_reduce_lognormal_workload(gpu, q, scratch):
    # 65,536 segments, log-normal lengths (median 512, sigma 2, clipped to
    # [1, 2^20]): fp32 segmented sum against the FP64 per-segment sums
    torch, _ = gpu
    off = lognormal_offsets(65536, seed=0)
    n = int(off[-1])
    S = len(off) - 1
    x = torch.randn(n, generator=_gen(torch, 6), device="cuda", dtype=torch.float32)
    od = torch.from_numpy(off).cuda()
    out = torch.empty(S, dtype=torch.float32, device="cuda")
    acpp.segmented_reduce(q, scratch, x, od, out, 0.0)
    q.wait()
    got = host(out).astype(np.float64)
    xh = host(x)
    exact = O.segmented_reduce_f32_fp64(xh, off)
    lens = np.diff(off.astype(np.int64))
    assert lens.min() >= 1  # reduceat below needs non-empty segments
    abs_sums = np.add.reduceat(np.abs(xh.astype(np.float64)), off[:-1].astype(np.int64))
    bound = (64 + 2 * np.log2(lens + 2)) * 2.0 ** -24 * abs_sums * 1.0001
    assert np.all(np.abs(got - exact) <= bound)
