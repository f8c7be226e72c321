"""GPU parity: bulk atomics (histogram by fetch_add, strided fetch ops).

Final contents must be bit-exact against the oracle's serial fold
(order-independent operations) for both the privatized and the global
strategy.
"""
import numpy as np
import pytest

import opensycl_amd as acpp
import oracle as O
from gpu_util import dev, host, rand, zipf_keys

pytestmark = pytest.mark.gpu

STRATS = [acpp.ATOMICS_PRIVATIZED, acpp.ATOMICS_GLOBAL, acpp.ATOMICS_AUTO]


# 4096 bins keep 8 LDS copies of each counter, 8192 bins 4, 16384 bins 2,
# 32768 one; above that the counters are packed 16-bit
@pytest.mark.parametrize("num_bins", [1, 100, 4096, 8192, 16384, 32768, 32769, 65536])
@pytest.mark.parametrize("strategy", STRATS)
def test_histogram_zipf(gpu, q, scratch, num_bins, strategy):
    torch, _ = gpu
    for n in (1, 1001, (1 << 22) + 3):
        keys = zipf_keys(n, num_bins + 10, seed=n)  # some keys out of range
        start = rand(np.uint32, num_bins, seed=1, lo=0, hi=1000)
        bins = dev(torch, start)
        acpp.histogram(q, scratch, dev(torch, keys), bins, strategy)
        q.wait()
        assert np.array_equal(host(bins), O.histogram(keys, num_bins, start)), (n, strategy)


@pytest.mark.parametrize("num_bins", [4096, 65536])
def test_histogram_hot_bins_drain(gpu, q, scratch, num_bins):
    # every key in bin 0 or bin 1: the 16-bit LDS fields cross 0x8000 many
    # times per work-group and must drain without carrying into a neighbour
    torch, _ = gpu
    n = 1 << 24
    for hot in (0, 1, num_bins - 1):
        keys = np.full(n, hot, dtype=np.uint32)
        keys[::7] = (hot + 1) % num_bins
        bins = torch.zeros(num_bins, dtype=torch.uint32, device="cuda")
        acpp.histogram(q, scratch, dev(torch, keys), bins, acpp.ATOMICS_PRIVATIZED)
        q.wait()
        assert np.array_equal(host(bins), O.histogram(keys, num_bins)), hot


def test_histogram_misaligned_keys(gpu, q, scratch):
    torch, _ = gpu
    keys = zipf_keys(1000003, 4096, seed=3)
    for off in (1, 2, 3):
        kd = dev(torch, keys)[off:]
        for strat in (acpp.ATOMICS_PRIVATIZED, acpp.ATOMICS_GLOBAL):
            bins = torch.zeros(4096, dtype=torch.uint32, device="cuda")
            acpp.histogram(q, scratch, kd, bins, strat)
            q.wait()
            assert np.array_equal(host(bins), O.histogram(keys[off:], 4096))


@pytest.mark.parametrize("dtype", [np.int32, np.uint32, np.int64, np.uint64, np.float32,
                                   np.float64])
@pytest.mark.parametrize("num_slots", [1, 64, 1000, 1024, 4096, 70000])
def test_strided_fetch_ops(gpu, q, scratch, dtype, num_slots):
    torch, _ = gpu
    n = (1 << 20) + 5
    if np.issubdtype(dtype, np.floating):
        # small integers: every partial sum stays below 2^24, so float sums
        # are exact in any order
        v = np.round(rand(np.float64, n, seed=num_slots) * 4).astype(dtype)
    else:
        v = rand(dtype, n, seed=num_slots, lo=-(1 << 20) if np.iinfo(dtype).min < 0 else 0,
                 hi=1 << 20)
    idx = (np.arange(n) % num_slots).astype(np.uint32)
    for op in ("max", "min", "add"):
        for strat in STRATS:
            if strat == acpp.ATOMICS_PRIVATIZED and num_slots * v.itemsize > 128 * 1024:
                continue
            start = np.zeros(num_slots, dtype=dtype) + dtype(3)
            t = dev(torch, start)
            acpp.atomic_fetch_strided(q, scratch, op, dev(torch, v), t, strat)
            q.wait()
            exp = start.copy()
            if op == "max":
                np.maximum.at(exp, idx, v)
            elif op == "min":
                np.minimum.at(exp, idx, v)
            else:
                np.add.at(exp, idx, v)
            assert np.array_equal(host(t), exp), (op, strat)


def test_fetch_max_i32_oracle(gpu, q, scratch):
    # The atomics workload shape: uniform int32 into 1024 slots, slot = i mod 1024
    torch, _ = gpu
    n = 1 << 24
    v = rand(np.int32, n, seed=11)
    for strat in STRATS:
        t = torch.full((1024,), np.iinfo(np.int32).min, dtype=torch.int32, device="cuda")
        acpp.atomic_fetch_strided(q, scratch, "max", dev(torch, v), t, strat)
        q.wait()
        assert np.array_equal(host(t), O.fetch_max_strided(v, 1024)), strat
    # misaligned view -> head elements through the scalar path
    t = torch.full((1024,), np.iinfo(np.int32).min, dtype=torch.int32, device="cuda")
    acpp.atomic_fetch_strided(q, scratch, "max", dev(torch, v)[3:], t, acpp.ATOMICS_PRIVATIZED)
    q.wait()
    assert np.array_equal(host(t), O.fetch_max_strided(v[3:], 1024))


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.float32])
def test_strided_lds_form_misaligned_view(gpu, q, scratch, dtype):
    # 1000 slots do not divide the work-group stride: the LDS form, with a
    # view 3 elements into the buffer (scalar head, then 16-byte loads)
    torch, _ = gpu
    n = (1 << 21) + 9
    v = np.round(rand(np.float64, n + 3, seed=5) * 4).astype(dtype)
    vd = dev(torch, v)[3:]
    idx = (np.arange(n) % 1000).astype(np.uint32)
    for op in ("max", "min", "add"):
        start = np.zeros(1000, dtype=dtype) + dtype(3)
        t = dev(torch, start)
        acpp.atomic_fetch_strided(q, scratch, op, vd, t, acpp.ATOMICS_PRIVATIZED)
        q.wait()
        exp = start.copy()
        {"max": np.maximum, "min": np.minimum, "add": np.add}[op].at(exp, idx, v[3:])
        assert np.array_equal(host(t), exp), op
