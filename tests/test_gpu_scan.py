"""GPU parity: inclusive_scan / exclusive_scan against the oracle.

Integer scans must be bit-exact for every operator, size, alignment and
launch configuration; floating-point sums are compared with an FP64 prefix
sum within the bound stated in the test.
"""

import numpy as np
import pytest

import opensycl_amd as acpp
import oracle as O
from gpu_util import (ALL_OPS, INT_TYPES, dev, fold_bound, host, lognormal_offsets, op_input, rand,
                      scan_depth)

pytestmark = pytest.mark.gpu

# size ladder of tests/pstl/inclusive_scan.cpp:158-176 plus tile boundaries
# (a tile is 256 work-items x 8 rows x 16 bytes = 8192 int32)
SIZES = [1, 2, 127, 1000, 8191, 8192, 8193, 100003, 1 << 20]


def _scan_case(torch, q, scratch, x, op, kind, init=None, config=None):
    xd = dev(torch, x)
    yd = torch.empty_like(xd)
    if kind == "inclusive":
        acpp.inclusive_scan(q, scratch, xd, yd, op, init, config=config)
    else:
        acpp.exclusive_scan(q, scratch, xd, yd, init, op, config=config)
    q.wait()
    return host(yd)


@pytest.mark.parametrize("dtype", INT_TYPES)
def test_scan_int_all_ops_exact(gpu, q, scratch, dtype):
    torch, _ = gpu
    for op in ALL_OPS:
        for n in SIZES:
            x = op_input(dtype, op, n, seed=n * 7 + op)
            init = dtype(3)
            got = _scan_case(torch, q, scratch, x, op, "inclusive")
            assert np.array_equal(got, O.inclusive_scan(x, op)), (dtype, op, n, "incl")
            got = _scan_case(torch, q, scratch, x, op, "inclusive", init)
            assert np.array_equal(got, O.inclusive_scan(x, op, init)), (dtype, op, n, "incl+init")
            got = _scan_case(torch, q, scratch, x, op, "exclusive", init)
            assert np.array_equal(got, O.exclusive_scan(x, init, op)), (dtype, op, n, "excl")


def test_pstl_ladder_closed_forms(gpu, q, scratch):
    # tests/pstl/inclusive_scan.cpp, exclusive_scan.cpp: iota input, plus with
    # init 3 and multiplication with init 3.
    torch, _ = gpu
    for n in (1, 127, 1000, 1 << 20):
        i = np.arange(n, dtype=np.int64)
        x = i.astype(np.int32)
        inc = _scan_case(torch, q, scratch, x, acpp.PLUS, "inclusive", np.int32(3))
        exc = _scan_case(torch, q, scratch, x, acpp.PLUS, "exclusive", np.int32(3))
        assert np.array_equal(inc, ((i * (i + 1) // 2 + 3) & 0xffffffff).astype(np.uint32).view(np.int32))
        assert np.array_equal(exc, ((i * (i - 1) // 2 + 3) & 0xffffffff).astype(np.uint32).view(np.int32))
        m = _scan_case(torch, q, scratch, i.astype(np.uint64), acpp.MULTIPLY, "inclusive", np.uint64(3))
        assert np.all(m == 0)
        me = _scan_case(torch, q, scratch, i.astype(np.uint64), acpp.MULTIPLY, "exclusive", np.uint64(3))
        assert me[0] == 3 and np.all(me[1:] == 0)


# 4th field 1 = work-group-wide look-back window instead of one wave
@pytest.mark.parametrize("cfg", [(64, 2), (128, 4), (256, 2), (256, 4), (512, 8), (1024, 8),
                                 (1024, 2), (64, 2, 0, 1), (256, 8, 0, 1), (1024, 8, 0, 1)])
def test_scan_launch_configs(gpu, q, scratch, cfg):
    torch, _ = gpu
    for dtype in (np.int32, np.int64, np.float32):
        x = rand(dtype, 3 * 1024 * 1024 + 17, seed=cfg[0] + cfg[1])
        if dtype == np.float32:
            x = np.round(x * 100).astype(np.float32)  # small integers: fp sums are exact
        got = _scan_case(torch, q, scratch, x, acpp.PLUS, "exclusive", dtype(1), config=cfg)
        assert np.array_equal(got, O.exclusive_scan(x, dtype(1))), (dtype, cfg)
    assert scratch.watchdog_trips() == 0


# 1-byte elements: 4-row tiles widened in registers, at several group
# sizes, every operator, with partial last tiles.
@pytest.mark.parametrize("cfg", [None, (1024, 4), (512, 4), (256, 4), (64, 4)])
def test_byte_scan_kernels(gpu, q, scratch, cfg):
    torch, _ = gpu
    for dtype in (np.int8, np.uint8):
        for op in ALL_OPS:
            for n in (1000, 3 * 131072 + 77):
                x = op_input(dtype, op, n, seed=n + op)
                init = dtype(3)
                got = _scan_case(torch, q, scratch, x, op, "inclusive", config=cfg)
                assert np.array_equal(got, O.inclusive_scan(x, op)), (dtype, op, n, cfg, "incl")
                got = _scan_case(torch, q, scratch, x, op, "inclusive", init, config=cfg)
                assert np.array_equal(got, O.inclusive_scan(x, op, init)), (dtype, op, n, cfg, "incl+init")
                got = _scan_case(torch, q, scratch, x, op, "exclusive", init, config=cfg)
                assert np.array_equal(got, O.exclusive_scan(x, init, op)), (dtype, op, n, cfg, "excl")
    assert scratch.watchdog_trips() == 0


# 8-bit scans on packed 16-bit pairs (element j and j + 8 in one register):
# 8-row tiles at several group sizes, with partial last tiles, full-range
# random bytes, every operator the packed kernel covers; 4 rows and the
# work-group look-back take the generic kernel.
PACKED_OPS = [acpp.PLUS, acpp.MULTIPLY, acpp.MIN, acpp.MAX, acpp.BIT_AND, acpp.BIT_OR, acpp.BIT_XOR]


@pytest.mark.parametrize("cfg", [None, (1024, 8), (512, 8), (256, 8), (64, 8), (1024, 4), (256, 0, 0, 1)])
def test_packed_byte_scans(gpu, q, scratch, cfg):
    torch, _ = gpu
    for dtype in (np.int8, np.uint8):
        for op in PACKED_OPS:
            for n in (1, 15, 16, 17, 1000, 3 * 131072 + 77, (1 << 21) + 5):
                rng = np.random.default_rng(n + 31 * op)
                if op == acpp.MULTIPLY:
                    x = rng.choice(np.array([1, 3, 5, 7, 255 if dtype == np.uint8 else -1], dtype=dtype), n)
                elif op == acpp.BIT_AND:
                    x = (np.iinfo(dtype).max if dtype == np.uint8 else -1) ^ (rng.random(n) < 0.001).astype(dtype)
                    x = x.astype(dtype)
                else:
                    x = rng.integers(np.iinfo(dtype).min, np.iinfo(dtype).max + 1, n).astype(dtype)
                init = dtype(-3 if dtype == np.int8 else 250)
                got = _scan_case(torch, q, scratch, x, op, "inclusive", config=cfg)
                assert np.array_equal(got, O.inclusive_scan(x, op)), (dtype, op, n, cfg, "incl")
                got = _scan_case(torch, q, scratch, x, op, "inclusive", init, config=cfg)
                assert np.array_equal(got, O.inclusive_scan(x, op, init)), (dtype, op, n, cfg, "incl+init")
                got = _scan_case(torch, q, scratch, x, op, "exclusive", init, config=cfg)
                assert np.array_equal(got, O.exclusive_scan(x, init, op)), (dtype, op, n, cfg, "excl")
    assert scratch.watchdog_trips() == 0


def test_packed_byte_plus_scan_2e28(gpu, q, scratch):
    # full size, checked against the int64 running sum modulo 256 (exact)
    torch, _ = gpu
    n = 1 << 28
    x = torch.randint(-128, 128, (n,), dtype=torch.int8, device="cuda")
    y = torch.empty_like(x)
    acpp.inclusive_scan(q, scratch, x, y)
    q.wait()
    want = (torch.cumsum(x.to(torch.int64), 0) & 0xff).to(torch.uint8).view(torch.int8)
    assert torch.equal(y, want)
    del want
    acpp.exclusive_scan(q, scratch, x, y, np.int8(7))
    q.wait()
    want = ((torch.cumsum(x.to(torch.int64), 0) - x.to(torch.int64) + 7) & 0xff).to(torch.uint8).view(torch.int8)
    assert torch.equal(y, want)
    assert scratch.watchdog_trips() == 0


def test_byte_scan_rows_rejected(gpu, q, scratch):
    # the logical operators have no packed 8-row kernel and take 4 rows;
    # the work-group look-back takes the generic kernel (4 rows)
    torch, _ = gpu
    x = torch.zeros(1000, dtype=torch.int8, device="cuda")
    y = torch.empty_like(x)
    with pytest.raises(acpp.AcppError):
        acpp.inclusive_scan(q, scratch, x, y, acpp.LOGICAL_AND, config=(256, 8))
    with pytest.raises(acpp.AcppError):
        acpp.inclusive_scan(q, scratch, x, y, acpp.PLUS, config=(256, 8, 0, 1))


@pytest.mark.parametrize("offset", [1, 2, 3])
def test_scan_misaligned_and_in_place(gpu, q, scratch, offset):
    torch, _ = gpu
    for dtype in (np.int8, np.int16, np.int32, np.int64):
        x = rand(dtype, 100000 + offset, seed=offset)
        xd = dev(torch, x)[offset:]
        yd = torch.empty(100000, dtype=xd.dtype, device="cuda")
        acpp.inclusive_scan(q, scratch, xd, yd)
        q.wait()
        assert np.array_equal(host(yd), O.inclusive_scan(x[offset:]))
        # in place (d_first == first)
        zd = dev(torch, x)
        acpp.exclusive_scan(q, scratch, zd, zd, dtype(4))
        q.wait()
        assert np.array_equal(host(zd), O.exclusive_scan(x, dtype(4)))


def test_scan_repeated_launches_share_scratch(gpu, q, scratch):
    # the sync region is reused launch after launch (epochs advance); sizes
    # alternate so stale words from a larger launch are present.
    torch, _ = gpu
    for rep in range(12):
        n = [5000000, 70000, 1 << 20, 123][rep % 4]
        x = rand(np.int32, n, seed=rep)
        got = _scan_case(torch, q, scratch, x, acpp.PLUS, "inclusive")
        assert np.array_equal(got, O.inclusive_scan(x)), rep


@pytest.mark.parametrize("dtype,eps", [(np.float32, 2.0 ** -24), (np.float64, 2.0 ** -53)])
def test_scan_fp_sum_vs_fp64(gpu, q, scratch, dtype, eps):
    torch, _ = gpu
    for n in (1, 1000, 100003, 1 << 21):
        x = rand(dtype, n, seed=n)
        got = _scan_case(torch, q, scratch, x, acpp.PLUS, "inclusive").astype(np.float64)
        x64 = x.astype(np.float64)
        exact = np.cumsum(x64)
        # tiles hold at least 256 x 8 x 16 bytes (4096 doubles)
        bound = fold_bound(np.cumsum(np.abs(x64)), scan_depth(np.arange(n), 4096), eps) + 1e-300
        # np.cumsum in FP64 differs from the exact prefix by far less than eps(f32)
        assert np.all(np.abs(got - exact) <= bound)


def test_scan_fp_minmax_exact(gpu, q, scratch):
    torch, _ = gpu
    for dtype in (np.float32, np.float64):
        x = rand(dtype, 200001, seed=5)
        for op in (acpp.MIN, acpp.MAX):
            got = _scan_case(torch, q, scratch, x, op, "inclusive")
            assert np.array_equal(got, O.inclusive_scan(x, op))


def test_scan_large_int32_exact(gpu, q, scratch):
    # 2^28 int32 in [-2^15, 2^15] (the bandwidth-scale workload), exclusive, init 0
    torch, _ = gpu
    n = 1 << 28
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    xd = torch.randint(-(1 << 15), (1 << 15) + 1, (n,), dtype=torch.int32, device="cuda",
                       generator=g)
    yd = torch.empty_like(xd)
    acpp.exclusive_scan(q, scratch, xd, yd, np.int32(0))
    q.wait()
    x = host(xd)
    del xd
    y = host(yd)
    del yd
    assert np.array_equal(y, O.exclusive_scan(x, np.int32(0)))


def test_look_back_kernels_share_scratch_regions(gpu, q):
    # Scans, copy_if and segmented scans draw their status regions from one
    # cache and continue each region's epoch count.  Stale words of one kernel
    # must never read as live to another (all use one tag layout), and no
    # look-back may give up waiting.  Small tiles: many statuses per region.
    torch, cache = gpu
    off = lognormal_offsets(2000, seed=3, max_len=1 << 17)
    ns = int(off[-1])
    xs = rand(np.int64, ns, seed=9, lo=-1000, hi=1000)
    xsd, od = dev(torch, xs), dev(torch, off)
    exp_seg = O.segmented_inclusive_scan(xs, off)
    x32 = rand(np.int32, 3 * 1024 * 1024 + 17, seed=4)
    x32d = dev(torch, x32)
    exp_scan = O.exclusive_scan(x32, np.int32(1))
    exp_sel, exp_k = O.copy_if(x32, O.PRED_GT, 0)
    for rep in range(4):
        for cfg_seg, cfg_scan in (((64, 2), (128, 4)), ((128, 2), (64, 2)), ((256, 4, 0, 1), (128, 4, 0, 1))):
            g = acpp.AllocationGroup(cache, 0)
            ys = torch.zeros_like(xsd)
            acpp.segmented_inclusive_scan(q, g, xsd, ys, od, acpp.PLUS, config=cfg_seg)
            y32 = torch.empty_like(x32d)
            acpp.exclusive_scan(q, g, x32d, y32, np.int32(1), acpp.PLUS, config=cfg_scan)
            sel = torch.empty_like(x32d)
            cnt = torch.zeros(1, dtype=torch.uint64, device="cuda")
            acpp.copy_if(q, g, x32d, sel, acpp.PRED_GT, 0, num_copied=cnt)
            q.wait()
            assert g.watchdog_trips() == 0, (rep, cfg_seg, cfg_scan)
            assert np.array_equal(host(ys), exp_seg), (rep, cfg_seg)
            assert np.array_equal(host(y32), exp_scan), (rep, cfg_scan)
            k = int(host(cnt)[0])
            assert k == exp_k and np.array_equal(host(sel)[:k], exp_sel), rep
            g.close()
