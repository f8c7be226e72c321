"""The error path of the waits between work-groups.

A look-back (scans, copy_if, segmented scans) or a reduction's wait for the
other groups' partials gives up after its watchdog time; its results are
then wrong, and the library must say so.  These tests force the give-up
from the host (launch-config flag ACPP_VARIANT_FORCE_GIVE_UP: every such
wait gives up at once, and every kernel still ends normally) and check that
the give-up is reported as ERROR_WATCHDOG at the next synchronisation point
(Queue.wait, acpp_check_asynchronous_errors, group release), once, and that
work after it runs and reports normally.

Each test uses a cache of its own: a forced give-up leaves its count in the
scratch regions, which must not reach the other tests' zero-trip checks.
"""
import numpy as np
import pytest

import opensycl_amd as acpp
import oracle as O
from opensycl_amd._abi import LaunchConfig
from gpu_util import dev, host, rand

pytestmark = pytest.mark.gpu

FORCE = LaunchConfig(0, 0, 0, acpp.VARIANT_FORCE_GIVE_UP)


@pytest.fixture()
def own(gpu):
    torch, _ = gpu
    cache = acpp.AllocationCache()
    group = acpp.AllocationGroup(cache, 0)
    yield torch, group
    torch.cuda.synchronize()
    try:
        group.close()
    except acpp.AcppError:
        pass
    cache.close()
    acpp.check_asynchronous_errors(0)  # leave no raised flag behind


def _expect_watchdog(q):
    with pytest.raises(acpp.AcppError) as info:
        q.wait()
    assert info.value.status == acpp.ERROR_WATCHDOG
    q.wait()  # reported once


def test_scan_give_up_is_reported(own, q):
    torch, g = own
    x = np.ones(3 * 8192, dtype=np.int32)  # three 256 x 8 x 4 tiles
    xd = dev(torch, x)
    yd = torch.empty_like(xd)
    acpp.inclusive_scan(q, g, xd, yd, config=FORCE)
    _expect_watchdog(q)
    assert g.watchdog_trips() >= 2  # tiles 1 and 2 gave up
    # the same group and queue work normally afterwards
    acpp.exclusive_scan(q, g, xd, yd, 5)
    q.wait()
    assert np.array_equal(host(yd), O.exclusive_scan(x, np.int32(5)))


def test_reduce_give_up_is_reported(own, q):
    torch, g = own
    x = rand(np.int64, 1 << 20, seed=3)
    xd = dev(torch, x)
    out = torch.zeros(1, dtype=torch.int64, device="cuda")
    acpp.reduce(q, g, xd, out, 0, config=FORCE)
    _expect_watchdog(q)
    acpp.reduce(q, g, xd, out, 0)
    q.wait()
    assert host(out)[0] == O.reduce(x, np.int64(0))
    outs = [torch.zeros(1, dtype=torch.int64, device="cuda") for _ in range(2)]
    acpp.reduction_multi(q, g, [xd, xd], outs, [acpp.PLUS, acpp.MAX], True, config=FORCE)
    _expect_watchdog(q)


def test_copy_if_and_segmented_give_up_are_reported(own, q):
    torch, g = own
    x = rand(np.int32, 5 * 8192 * 4, seed=4)
    xd = dev(torch, x)
    d = torch.empty_like(xd)
    acpp.copy_if(q, g, xd, d, acpp.PRED_GT, 0, config=FORCE)
    _expect_watchdog(q)
    off = np.array([0, 100, 70000, x.size], dtype=np.uint64)
    y = torch.empty_like(xd)
    acpp.segmented_inclusive_scan(q, g, xd, y, dev(torch, off), config=FORCE)
    _expect_watchdog(q)
    acpp.segmented_inclusive_scan(q, g, xd, y, dev(torch, off))
    q.wait()
    assert np.array_equal(host(y), O.segmented_inclusive_scan(x, off))


def test_group_release_reports_an_unchecked_give_up(gpu):
    torch, _ = gpu
    cache = acpp.AllocationCache()
    q = acpp.Queue()
    g = acpp.AllocationGroup(cache, 0)
    xd = torch.ones(3 * 8192, dtype=torch.int32, device="cuda")
    yd = torch.empty_like(xd)
    acpp.inclusive_scan(q, g, xd, yd, config=FORCE)
    torch.cuda.synchronize()  # no check on the way
    with pytest.raises(acpp.AcppError) as info:
        g.close()
    assert info.value.status == acpp.ERROR_WATCHDOG
    # release reports without clearing; the check takes the flag
    with pytest.raises(acpp.AcppError):
        acpp.check_asynchronous_errors(0)
    acpp.check_asynchronous_errors(0)
    cache.close()
