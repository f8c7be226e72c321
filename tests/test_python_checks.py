"""Argument checks of the Python mirror (CPU only: nothing is launched).

The C-ABI receives raw pointers and element counts, so an output that is
too short or of a narrower type would be written past its end on the GPU.
The Python functions must reject such outputs before any library call; the
tensors here live on the CPU, so reaching the library would fail
differently (a ValueError about device memory is raised only after these
checks)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import opensycl_amd as acpp  # noqa: E402


def t(n, dtype=torch.int32):
    return torch.zeros(n, dtype=dtype)


CASES = [
    # (call, expected exception)
    (lambda: acpp.inclusive_scan(None, None, t(10), t(9)), ValueError),
    (lambda: acpp.inclusive_scan(None, None, t(10), t(10, torch.int64)), TypeError),
    (lambda: acpp.exclusive_scan(None, None, t(10, torch.int64), t(10, torch.int32)), TypeError),
    (lambda: acpp.exclusive_scan(None, None, t(10), t(3)), ValueError),
    (lambda: acpp.transform_inclusive_scan(None, None, t(10), t(9)), ValueError),
    (lambda: acpp.transform_exclusive_scan(None, None, t(10), t(10, torch.int16)), TypeError),
    (lambda: acpp.reduce(None, None, t(10), t(0)), ValueError),
    (lambda: acpp.reduce(None, None, t(10, torch.int64), t(1, torch.int32)), TypeError),
    (lambda: acpp.reduce(None, None, t(10, torch.bfloat16), t(1, torch.bfloat16)), TypeError),
    (lambda: acpp.reduce_compensated(None, None, t(10, torch.float32), t(1, torch.float64)), TypeError),
    (lambda: acpp.transform_reduce(None, None, t(10, torch.bfloat16), t(10, torch.bfloat16),
                                   t(1, torch.bfloat16)), TypeError),
    (lambda: acpp.transform_reduce(None, None, t(10, torch.float32), t(10, torch.float32),
                                   t(0, torch.float32), compensated=True), ValueError),
    (lambda: acpp.transform_reduce(None, None, t(10), t(9), t(1)), ValueError),
    (lambda: acpp.transform_reduce_unary(None, None, t(10), t(1, torch.int8)), TypeError),
    (lambda: acpp.reduction(None, None, t(10), t(0)), ValueError),
    (lambda: acpp.reduction_multi(None, None, [t(10)], [t(0)], [acpp.PLUS]), ValueError),
    (lambda: acpp.copy_if(None, None, t(10), t(9), acpp.PRED_NONZERO), ValueError),
    (lambda: acpp.copy_if(None, None, t(10), t(10), acpp.PRED_NONZERO,
                          num_copied=t(1, torch.int64)), TypeError),
    (lambda: acpp.copy_if_stencil(None, None, t(10), t(10, torch.uint8), t(10, torch.int64)), TypeError),
    (lambda: acpp.segmented_inclusive_scan(None, None, t(10), t(9), t(3, torch.uint64)), ValueError),
    (lambda: acpp.segmented_reduce(None, None, t(10), t(5, torch.uint64), t(3)), ValueError),
    (lambda: acpp.segmented_reduce(None, None, t(10), t(0, torch.uint64), t(3)), ValueError),
    (lambda: acpp.atomic_fetch_strided(None, None, "max", t(10), t(0)), ValueError),
    (lambda: acpp.SegmentedPlan.inclusive_scan(None, t(10), t(10, torch.int64), t(3, torch.uint64)),
     TypeError),
    (lambda: acpp.SegmentedPlan.reduce(None, t(10), t(5, torch.uint64), t(2)), ValueError),
]


@pytest.mark.parametrize("i", range(len(CASES)))
def test_bad_outputs_are_rejected_before_launch(i):
    call, exc = CASES[i]
    with pytest.raises(exc) as info:
        call()
    assert "device memory" not in str(info.value)


def test_well_formed_call_reaches_the_device_check():
    # the same call with a good output gets past the checks and stops only
    # at the device-memory test of the pointer conversion
    with pytest.raises(ValueError, match="device memory"):
        h = type("H", (), {"handle": None})()
        acpp.inclusive_scan(h, h, t(10), t(10))


def test_group_release_waits_for_plans():
    # bookkeeping only: a group with live plans defers its release
    g = acpp.AllocationGroup.__new__(acpp.AllocationGroup)
    g._h = None
    g._plans = 1
    g._close_pending = False
    g.close()
    assert g._close_pending
    p = acpp.SegmentedPlan.__new__(acpp.SegmentedPlan)
    p._h = None
    p._keep = ()
    p._scratch = g
    p.close()
    assert g._plans == 0
    np.testing.assert_equal(p._scratch, None)


def test_release_refused_while_a_plan_is_alive():
    # a captured graph keeps using the group's regions: releasing them to the
    # cache before the plan is closed would hand them to other calls
    g = acpp.AllocationGroup.__new__(acpp.AllocationGroup)
    g._h = None
    g._plans = 1
    g._close_pending = False
    with pytest.raises(RuntimeError, match="SegmentedPlan"):
        g.release()
    g.close()  # deferred while the plan lives
    assert g._close_pending
