"""GPU parity: segmented inclusive scans and reductions, direct and through
hipGraph plans, against the oracle's per-segment serial scan / fold."""
import math

import numpy as np
import pytest

import opensycl_amd as acpp
import oracle as O
from gpu_util import dev, exact_ints, fold_bound, host, lognormal_offsets, rand, scan_depth

# segmented tiles hold at least 128 work-items x 2 rows x 16 bytes
MIN_TILE = 512

pytestmark = pytest.mark.gpu


def layouts():
    rng = np.random.default_rng(5)
    yield "lognormal", lognormal_offsets(3000, seed=1, max_len=1 << 18)
    yield "ones", np.arange(0, 20001, dtype=np.uint64)
    lens = rng.integers(0, 5, 5000).astype(np.uint64)  # many empty segments
    yield "with_empty", np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    yield "one_huge", np.array([0, (1 << 22) + 77], dtype=np.uint64)
    lens = np.array([0, 0, 3, 0, 70000, 0, 1, 123456, 0], dtype=np.uint64)
    yield "mixed_edges", np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)


def _values(dtype, n, seed):
    if np.issubdtype(dtype, np.floating):
        return rand(dtype, n, seed)
    return rand(dtype, n, seed, lo=-(1 << 20), hi=1 << 20)


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.float32, np.float64])
def test_segmented_scan(gpu, q, scratch, dtype):
    torch, _ = gpu
    for name, off in layouts():
        n = int(off[-1])
        x = _values(dtype, n, seed=n)
        xd, od = dev(torch, x), dev(torch, off)
        for op in (acpp.PLUS, acpp.MAX):
            yd = torch.zeros_like(xd)
            acpp.segmented_inclusive_scan(q, scratch, xd, yd, od, op)
            q.wait()
            got = host(yd)
            if np.issubdtype(dtype, np.floating) and op == acpp.PLUS:
                # FP64 per-segment prefix sums; fold depth bound per element
                eps = np.finfo(dtype).eps
                ok = True
                x64 = x.astype(np.float64)
                for s in range(len(off) - 1):
                    b, e = int(off[s]), int(off[s + 1])
                    if e > b:
                        ex = np.cumsum(x64[b:e])
                        bound = fold_bound(np.cumsum(np.abs(x64[b:e])),
                                           scan_depth(np.arange(e - b), MIN_TILE), eps)
                        ok &= bool(np.all(np.abs(got[b:e].astype(np.float64) - ex) <= bound + 1e-300))
                assert ok, (name, op)
            else:
                assert np.array_equal(got, O.segmented_inclusive_scan(x, off, op)), (name, op)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_segmented_exact_fp(gpu, q, scratch, dtype):
    # integer-valued floats: per-segment sums and prefix sums are exact, so
    # the floating-point kernels must agree bit for bit with the oracle's
    # integer scan and fold; a lost or duplicated element fails here
    torch, _ = gpu
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    for name, off in layouts():
        n = int(off[-1])
        xi = exact_ints(torch, n, n % 1000)
        xh = host(xi).astype(np.int64)
        xd, od = xi.to(tdt), dev(torch, off)
        yd = torch.zeros_like(xd)
        acpp.segmented_inclusive_scan(q, scratch, xd, yd, od)
        out = torch.zeros(len(off) - 1, dtype=tdt, device="cuda")
        acpp.segmented_reduce(q, scratch, xd, od, out, 0.0)
        q.wait()
        assert np.array_equal(host(yd).astype(np.int64), O.segmented_inclusive_scan(xh, off)), name
        assert np.array_equal(host(out).astype(np.int64), O.segmented_reduce(xh, off, np.int64(0))), name


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.float32, np.float64])
def test_segmented_reduce(gpu, q, scratch, dtype):
    torch, _ = gpu
    for name, off in layouts():
        n = int(off[-1])
        S = len(off) - 1
        x = _values(dtype, n, seed=n + 1)
        xd, od = dev(torch, x), dev(torch, off)
        for op, init in ((acpp.PLUS, dtype(7)), (acpp.MIN, dtype(1 << 20)), (acpp.MAX, dtype(-5))):
            out = torch.full((S,), -99, dtype=xd.dtype, device="cuda")
            acpp.segmented_reduce(q, scratch, xd, od, out, init, op)
            q.wait()
            got = host(out)
            if np.issubdtype(dtype, np.floating) and op == acpp.PLUS:
                exact = O.segmented_reduce_f32_fp64(x.astype(np.float32), off) if dtype == np.float32 \
                    else np.array([math.fsum(x[int(off[s]):int(off[s + 1])]) for s in range(S)])
                lens = np.diff(off.astype(np.int64))
                absx = np.abs(x.astype(np.float64))
                abs_sums = np.array([absx[int(off[s]):int(off[s + 1])].sum() for s in range(S)])
                eps = np.finfo(dtype).eps
                bound = fold_bound(abs_sums + 7, scan_depth(lens, MIN_TILE) + 1, eps) + 1e-300
                assert np.all(np.abs(got.astype(np.float64) - (exact + 7)) <= bound), name
            else:
                assert np.array_equal(got, O.segmented_reduce(x, off, init, op)), (name, op)


# flags: 1 = work-group-wide look-back window, 4 = one work-group per tile.
# (64, 2) gives 256-element int64 tiles: thousands of single-wave tiles in
# flight, the heaviest look-back contention (a regression case for the
# look-back backoff; without it the look-backs ran out of time).
@pytest.mark.parametrize("cfg", [(64, 2), (128, 4), (256, 2), (512, 4), (1024, 2), (1024, 4),
                                 (256, 4, 0, 1), (256, 4, 0, 4), (512, 8, 0, 5), (64, 2, 0, 4),
                                 (1024, 8, 0, 0), (64, 2, 0, 1)])
def test_segmented_configs(gpu, q, scratch, cfg):
    torch, _ = gpu
    off = lognormal_offsets(2000, seed=3, max_len=1 << 17)
    n = int(off[-1])
    x = rand(np.int64, n, seed=9, lo=-1000, hi=1000)
    xd, od = dev(torch, x), dev(torch, off)
    yd = torch.zeros_like(xd)
    acpp.segmented_inclusive_scan(q, scratch, xd, yd, od, acpp.PLUS, config=cfg)
    out = torch.zeros(len(off) - 1, dtype=torch.int64, device="cuda")
    acpp.segmented_reduce(q, scratch, xd, od, out, np.int64(0), acpp.PLUS, config=cfg)
    q.wait()
    assert scratch.watchdog_trips() == 0
    assert np.array_equal(host(yd), O.segmented_inclusive_scan(x, off))
    assert np.array_equal(host(out), O.segmented_reduce(x, off, np.int64(0)))


def test_segmented_plan_replays(gpu, q, scratch):
    # hipGraph replays, each checked, with eager work between replays
    torch, _ = gpu
    off = lognormal_offsets(4000, seed=4, max_len=1 << 18)
    n = int(off[-1])
    x = rand(np.int64, n, seed=2, lo=-(1 << 20), hi=1 << 20)
    xd, od = dev(torch, x), dev(torch, off)
    yd = torch.zeros_like(xd)
    f = rand(np.float32, n, seed=3)
    fd = dev(torch, f)
    red = torch.zeros(len(off) - 1, dtype=torch.float32, device="cuda")
    scan_plan = acpp.SegmentedPlan.inclusive_scan(scratch, xd, yd, od)
    red_plan = acpp.SegmentedPlan.reduce(scratch, fd, od, red, 0.0)
    exp_scan = O.segmented_inclusive_scan(x, off)
    exact = O.segmented_reduce_f32_fp64(f, off)
    lens = np.diff(off.astype(np.int64))
    abs_sums = O.segmented_reduce_f32_fp64(np.abs(f), off)
    bound = fold_bound(abs_sums, scan_depth(lens, MIN_TILE), np.finfo(np.float32).eps) + 1e-30
    for rep in range(4):
        yd.zero_()
        red.zero_()
        scan_plan.launch(q)
        red_plan.launch(q)
        q.wait()
        assert np.array_equal(host(yd), exp_scan), rep
        got = host(red).astype(np.float64)
        assert np.all(np.abs(got - exact) <= bound), rep
        # eager work on the same stream between replays
        tmp = torch.zeros_like(xd)
        acpp.inclusive_scan(q, scratch, xd, tmp)
        q.wait()
        # change the input: the next replay must see the new values
        x = x + 1
        xd.copy_(dev(torch, x))
        exp_scan = O.segmented_inclusive_scan(x, off)
    scan_plan.close()
    red_plan.close()


def test_segmented_lognormal_workload_shape(gpu, q, scratch):
    # 65,536 segments, log-normal lengths (median 512, sigma 2, clipped to
    # [1, 2^20]): int64 segmented inclusive scan, exact
    torch, _ = gpu
    off = lognormal_offsets(65536, seed=0)
    n = int(off[-1])
    x = rand(np.int64, n, seed=7, lo=-(1 << 20), hi=1 << 20)
    xd, od = dev(torch, x), dev(torch, off)
    yd = torch.empty_like(xd)
    acpp.segmented_inclusive_scan(q, scratch, xd, yd, od)
    q.wait()
    y = host(yd)
    del yd
    assert np.array_equal(y, O.segmented_inclusive_scan(x, off))


def test_segmented_plan_first_capture_allocates(gpu, q):
    # A plan made with a fresh cache has to allocate its scratch while the
    # launch is being captured; the plan must still work, and later torch
    # work must not see a stale HIP error.
    torch, _ = gpu
    off = lognormal_offsets(500, seed=8, max_len=1 << 15)
    n = int(off[-1])
    x = rand(np.int64, n, seed=4, lo=-1000, hi=1000)
    xd, od = dev(torch, x), dev(torch, off)
    yd = torch.zeros_like(xd)
    cache = acpp.AllocationCache()
    with acpp.AllocationGroup(cache, 0) as g:
        plan = acpp.SegmentedPlan.inclusive_scan(g, xd, yd, od)
        plan.launch(q)
        q.wait()
        assert np.array_equal(host(yd), O.segmented_inclusive_scan(x, off))
        z = torch.ones(1000, device="cuda").cumsum(0)  # torch work after the capture
        assert z[-1].item() == 1000.0
        plan.close()
    cache.close()


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_segmented_views_one_element_off(gpu, q, scratch, dtype):
    # input and output views one element into their buffers (not 16-byte
    # aligned): 16-byte accesses at element alignment
    torch, _ = gpu
    off = lognormal_offsets(3000, seed=11, max_len=1 << 16)
    n = int(off[-1])
    x = rand(dtype, n + 1, seed=12, lo=-1000, hi=1000)
    xd = dev(torch, x)[1:]
    od = dev(torch, off)
    yb = torch.zeros(n + 1, dtype=xd.dtype, device="cuda")
    acpp.segmented_inclusive_scan(q, scratch, xd, yb[1:], od, acpp.PLUS)
    out = torch.zeros(len(off) - 1, dtype=xd.dtype, device="cuda")
    acpp.segmented_reduce(q, scratch, xd, od, out, dtype(0), acpp.PLUS)
    q.wait()
    assert np.array_equal(host(yb)[1:], O.segmented_inclusive_scan(x[1:], off))
    assert np.array_equal(host(out), O.segmented_reduce(x[1:], off, dtype(0), acpp.PLUS))


def test_plan_keeps_its_group_and_buffers_alive(gpu, q):
    # A plan built from a temporary group and temporary tensors: after the
    # builder's references are gone and the allocator and cache have been
    # churned, a replay still scans the right data into the right output.
    import gc
    torch, cache = gpu
    off = lognormal_offsets(500, seed=9, max_len=1 << 14)
    n = int(off[-1])
    x = rand(np.int64, n, seed=9, lo=-1000, hi=1000)

    def make():
        g = acpp.AllocationGroup(cache, 0)
        xd, od = dev(torch, x), dev(torch, off)
        yd = torch.zeros_like(xd)
        return acpp.SegmentedPlan.inclusive_scan(g, xd, yd, od), yd

    plan, yd = make()
    gc.collect()
    junk = [torch.full((n,), 7, dtype=torch.int64, device="cuda") for _ in range(3)]
    with acpp.AllocationGroup(cache, 0) as other:
        acpp.inclusive_scan(q, other, junk[0], junk[1])
        plan.launch(q)
        acpp.inclusive_scan(q, other, junk[1], junk[2])
        q.wait()
    assert np.array_equal(host(yd), O.segmented_inclusive_scan(x, off))
    assert int(host(junk[2])[-1]) == 7 * n * (n + 1) // 2  # n < 2^23: no wrap
    plan.close()
