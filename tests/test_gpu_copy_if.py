"""GPU parity: copy_if (stream compaction) against the oracle's restatement
of algorithm.hpp:178-243, with the reference's own test cases
(tests/pstl/copy_if.cpp) and every predicate kind, type and edge size."""
import numpy as np
import pytest

import opensycl_amd as acpp
import oracle as O
from gpu_util import dev, host

pytestmark = pytest.mark.gpu

TYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64]


def _run(torch, q, scratch, x, pred, operand=0, mask=0, stencil=None, config=None, offset=0):
    xd_full = dev(torch, np.concatenate([np.zeros(offset, dtype=x.dtype), x]))
    xd = xd_full[offset:]
    d = torch.empty((max(x.size, 1),), dtype=xd.dtype, device="cuda")
    cnt64 = torch.full((1,), 12345, dtype=torch.int64, device="cuda")
    cnt = cnt64.view(torch.uint64)
    if stencil is None:
        acpp.copy_if(q, scratch, xd, d, pred, operand, mask, cnt, config=config)
    else:
        acpp.copy_if_stencil(q, scratch, xd, dev(torch, stencil), d, cnt, config=config)
    q.wait()
    c = int(cnt64.item())
    return host(d)[:c], c


@pytest.mark.parametrize("n,gen,count", [
    (0, lambda i: i, 0), (1, lambda i: i + 3, 0), (1000, lambda i: 1, 0),
    (1000, lambda i: 2 * i, 1000), (1000, lambda i: i, 500)])
def test_reference_cases(gpu, q, scratch, n, gen, count):
    # tests/pstl/copy_if.cpp:25-64 (predicate x % 2 == 0)
    torch, _ = gpu
    x = np.array([gen(i) for i in range(n)], dtype=np.int32)
    out, c = _run(torch, q, scratch, x, acpp.PRED_MASK_EQ, 0, 1)
    assert c == count
    assert np.array_equal(out, O.copy_if(x, O.PRED_MASK_EQ, 0, 1)[0])


@pytest.mark.parametrize("dtype", TYPES)
def test_predicates_and_sizes(gpu, q, scratch, dtype):
    torch, _ = gpu
    rng = np.random.default_rng(1)
    for n in (1, 63, 64, 4097, 131072 + 5, (1 << 22) + 17):
        if np.issubdtype(dtype, np.integer):
            x = rng.integers(0, 100, n).astype(dtype)
            operand, mask = dtype(37), dtype(6)
        else:
            x = np.round(rng.standard_normal(n) * 4).astype(dtype)
            operand, mask = dtype(0.0), dtype(0)
        preds = range(8) if np.issubdtype(dtype, np.integer) else range(7)
        for pred in preds:
            out, c = _run(torch, q, scratch, x, pred, operand, mask)
            ref, rc = O.copy_if(x, pred, operand, mask)
            assert c == rc, (n, pred)
            assert np.array_equal(out, ref), (n, pred)


@pytest.mark.parametrize("dtype", TYPES)
def test_stencil_and_misaligned(gpu, q, scratch, dtype):
    torch, _ = gpu
    rng = np.random.default_rng(2)
    n = 3 * 1024 * 1024 + 11
    x = rng.integers(-1000, 1000, n).astype(dtype)  # narrow types wrap
    st = (rng.random(n) < 0.3).astype(np.uint8)
    out, c = _run(torch, q, scratch, x, None, stencil=st)
    ref, rc = O.copy_if(x, O.PRED_STENCIL, stencil=st)
    assert c == rc and np.array_equal(out, ref)
    ref_st, rc_st = ref, rc
    for offset in (1, 3):  # input not 4-byte aligned: scalar loads
        out, c = _run(torch, q, scratch, x, acpp.PRED_GT, dtype(0), offset=offset)
        ref, rc = O.copy_if(x, O.PRED_GT, dtype(0))
        assert c == rc and np.array_equal(out, ref), offset
    # stencil views: 1 byte off (byte-wise stencil reads) and 16 bytes off
    # (one load per data vector)
    for soff in (1, 16):
        sd_full = dev(torch, np.concatenate([np.full(soff, 7, dtype=np.uint8), st]))
        d = torch.empty((n,), dtype=torch.from_numpy(x[:0]).dtype, device="cuda")
        cnt64 = torch.zeros(1, dtype=torch.int64, device="cuda")
        acpp.copy_if_stencil(q, scratch, dev(torch, x), sd_full[soff:], d, cnt64.view(torch.uint64))
        q.wait()
        c = int(cnt64.item())
        assert c == rc_st and np.array_equal(host(d)[:c], ref_st), soff


@pytest.mark.parametrize("dtype", [np.int8, np.int16, np.int32, np.int64])
@pytest.mark.parametrize("kind", ["all", "none", "runs"])
def test_full_empty_and_run_selections(gpu, q, scratch, dtype, kind):
    # Every element of every wave row kept (each row's LDS segment filled to
    # its 64 x VEC capacity), none kept, and long alternating runs that
    # start and end inside rows, waves and tiles; several full tiles plus a
    # partial one.
    torch, _ = gpu
    n = (1 << 22) + 4099
    i = np.arange(n)
    if kind == "all":
        x = (i % 50).astype(dtype)
    elif kind == "none":
        x = (i % 50 + 100).astype(dtype)
    else:
        x = np.where((i // 777) % 2 == 0, i % 50, 100).astype(dtype)
    out, c = _run(torch, q, scratch, x, acpp.PRED_LT, dtype(60))
    ref, rc = O.copy_if(x, O.PRED_LT, dtype(60))
    assert c == rc and np.array_equal(out, ref)


@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.int32, np.int64])
def test_unaligned_output(gpu, q, scratch, dtype):
    # outputs 1, 2 and 3 elements past an aligned address (1-byte elements
    # are stored four to a 4-byte store between byte-wise segment ends);
    # the elements around the output stay untouched
    torch, _ = gpu
    x = (np.arange(1_000_003) % 97).astype(dtype)
    xd = dev(torch, x)
    ref, rc = O.copy_if(x, O.PRED_LT, dtype(40))
    for off in (1, 2, 3):
        d_full = torch.full((x.size + 8,), 99, dtype=xd.dtype, device="cuda")
        cnt64 = torch.zeros(1, dtype=torch.int64, device="cuda")
        acpp.copy_if(q, scratch, xd, d_full[off:], acpp.PRED_LT, dtype(40), 0, cnt64.view(torch.uint64))
        q.wait()
        assert int(cnt64.item()) == rc, off
        got = host(d_full)
        assert np.array_equal(got[off:off + rc], ref), off
        assert np.all(got[:off] == 99) and np.all(got[off + rc:] == 99), off


@pytest.mark.parametrize("cfg", [(64, 0), (256, 0), (512, 8), (1024, 8)])
def test_work_group_sizes(gpu, q, scratch, cfg):
    torch, _ = gpu
    x = np.arange(2_000_003, dtype=np.int64) * 7 % 1009
    out, c = _run(torch, q, scratch, x, acpp.PRED_LT, np.int64(500), config=cfg)
    ref, rc = O.copy_if(x, O.PRED_LT, np.int64(500))
    assert c == rc and np.array_equal(out, ref)


def test_rejects_bad_arguments(gpu, q, scratch):
    torch, _ = gpu
    x = torch.zeros(10, dtype=torch.float32, device="cuda")
    with pytest.raises(acpp.AcppError):
        acpp.copy_if(q, scratch, x, x, acpp.PRED_MASK_EQ, 0.0, 1.0)   # mask on floats
    with pytest.raises(acpp.AcppError):
        acpp.copy_if(q, scratch, x, x, 42)
