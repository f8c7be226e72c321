"""GPU parity at the full workload sizes (2^28 elements; 65,536 segments).

The smaller parity tests cover every type, operator and edge case; these run
the bandwidth-scale workloads themselves through the same entry points and
check them against the oracle: bit-exact for integer work, within the stated
bound against an FP64 oracle for floating point.  Inputs are generated on the
GPU (seeded torch generators) and copied back for the oracle.
"""
import numpy as np
import pytest

import opensycl_amd as acpp
import oracle as O
from gpu_util import (compensated_bound, exact_ints, fold_bound, host, lognormal_offsets,
                      reduce_depth, scan_depth)

pytestmark = pytest.mark.gpu

N = 1 << 28
EPS32 = 2.0 ** -24


def _gen(torch, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return g


def _fp_bound(abs_sum, n):
    return fold_bound(abs_sum, reduce_depth(n, acpp.compute_units()), EPS32)


def _comp_bound(exact, abs_sum, n):
    return compensated_bound(exact, abs_sum, n, acpp.compute_units())


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
    assert abs(float(out.item()) - exact) <= _fp_bound(abs_sum, N)
    assert abs(float(comp.item()) - exact) <= _comp_bound(exact, abs_sum, N)


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
            assert err <= _comp_bound(exact, abs_sum, N)
        else:
            assert err <= _fp_bound(abs_sum, N)
    xb = x.to(torch.bfloat16)
    bits = host(xb.view(torch.int16)).view(np.uint16)
    exact_b = O.sum_fp64(bits)
    abs_b = O.sum_fp64(bits & np.uint16(0x7FFF))
    for compensated in (False, True):
        out = torch.zeros(1, dtype=torch.float32, device="cuda")
        if compensated:
            acpp.reduce_compensated(q, scratch, xb, out, 0.0)
        else:
            acpp.reduce(q, scratch, xb, out, 0.0)
        q.wait()
        err = abs(float(out.item()) - exact_b)
        if compensated:
            assert err <= _comp_bound(exact_b, abs_b, N)
        else:
            assert err <= _fp_bound(abs_b, N)


def test_precision_dots_2e28(gpu, q, scratch):
    # fp32 and bf16 dot products of 2^28 stressed values with 2^28 N(0,1):
    # plain within the fold bound, compensated (Dot2) within one fp32
    # rounding of the FP64 dot plus the second-order term
    torch, _ = gpu
    a = _stress(torch, 7)
    b = torch.randn(N, generator=_gen(torch, 8), device="cuda", dtype=torch.float32)
    ah, bh = host(a), host(b)
    cases = [(a, b, O.dot_fp64(ah, bh), O.dot_fp64(np.abs(ah), np.abs(bh)))]
    ab, bb = a.to(torch.bfloat16), b.to(torch.bfloat16)
    abits = host(ab.view(torch.int16)).view(np.uint16)
    bbits = host(bb.view(torch.int16)).view(np.uint16)
    cases.append((ab, bb, O.dot_fp64(abits, bbits),
                  O.dot_fp64(abits & np.uint16(0x7FFF), bbits & np.uint16(0x7FFF))))
    for x1, x2, exact, abs_sum in cases:
        for compensated in (False, True):
            out = torch.zeros(1, dtype=torch.float32, device="cuda")
            acpp.transform_reduce(q, scratch, x1, x2, out, 0.0, compensated=compensated)
            q.wait()
            err = abs(float(out.item()) - exact)
            bound = _comp_bound(exact, abs_sum, N) if compensated else _fp_bound(abs_sum, N)
            assert err <= bound, (x1.dtype, compensated, err, bound)


def test_exact_sums_and_dots_2e28(gpu, q, scratch):
    # Exactly summable inputs: non-zero integers in +-[1, 8] (exact in fp32
    # and bf16).  Every sum and dot product, plain or compensated, must hit
    # the integer result exactly, so one lost element (or one lost 16-byte
    # vector, unless its elements cancel) fails the test.
    torch, _ = gpu
    xi = exact_ints(torch, N, 9)
    yi = exact_ints(torch, N, 10)
    xh = host(xi).astype(np.int64)
    yh = host(yi).astype(np.int64)
    want_sum = int(O.reduce(xh, np.int64(0)))
    want_dot = int(O.dot(xh, yh, np.int64(0)))
    del xh, yh
    for dt in (torch.float32, torch.bfloat16):
        x, y = xi.to(dt), yi.to(dt)
        for compensated in (False, True):
            out = torch.zeros(1, dtype=torch.float32, device="cuda")
            if compensated:
                acpp.reduce_compensated(q, scratch, x, out, 0.0)
            else:
                acpp.reduce(q, scratch, x, out, 0.0)
            dot = torch.zeros(1, dtype=torch.float32, device="cuda")
            acpp.transform_reduce(q, scratch, x, y, dot, 0.0, compensated=compensated)
            q.wait()
            assert float(out.item()) == want_sum, (dt, compensated)
            assert float(dot.item()) == want_dot, (dt, compensated)
        # a view that starts one element in, so the scalar head is used too
        out = torch.zeros(1, dtype=torch.float32, device="cuda")
        acpp.reduce(q, scratch, x[1:], out, 0.0)
        q.wait()
        assert float(out.item()) == want_sum - int(xi[0].item()), dt
        del x, y


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
    # segment partials of whole tiles chained by the fix-up: one level per
    # tile the segment spans, on top of the in-tile depth
    bound = fold_bound(abs_sums, scan_depth(lens, 1024), EPS32)
    assert np.all(np.abs(got - exact) <= bound)
    # exactly summable values: every segment sum must be exact
    xi = exact_ints(torch, n, 11)
    xe = xi.to(torch.float32)
    acpp.segmented_reduce(q, scratch, xe, od, out, 0.0)
    q.wait()
    want = O.segmented_reduce(host(xi).astype(np.int64), off, np.int64(0))
    assert np.array_equal(host(out).astype(np.int64), want)
