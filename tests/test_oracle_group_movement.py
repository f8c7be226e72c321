"""Pins the oracle's group data movement (oracle_group_move_<t>) against the
expected values of the reference's tests/sycl/group_functions/
group_functions_misc.cpp: data x[i] = i, and
  group_broadcast (:71-170): every member gets member 0 / member 10 of its group;
  sub_group_broadcast (:174-270): member 0 / 7 of its 64-wide sub-group;
  group shift_group_left / right by 1, permute_group_by_xor 1 and
  select_from_group(id 0) (:281-425): x[i+1], x[i-1], x[i^1], group member 0,
  where the source lies in the group;
  subgroup shift_group_left (:427-):  x[i+1] except the last member of a
  sub-group or group.
The local sizes are those of the reference's harness, plus partial groups.
A planted fault (shift right by param - 1) must fail.  What the reference
leaves undefined (sources outside the group) is the library's documented
convention and is not checked here."""
import numpy as np
import pytest

import oracle as O

LOCAL_SIZES = [25, 64, 144, 256, 1024]


@pytest.fixture(autouse=True)
def _no_fault():
    O.set_fault(0)
    yield
    O.set_fault(0)


def cases_hold(dtype=np.int32):
    ok = True
    for L in LOCAL_SIZES:
        n = 4 * L + L // 2  # a partial last group too
        x = np.arange(n).astype(dtype)
        i = np.arange(n)
        base = (i // L) * L
        lid = i % L
        gsize = np.minimum(L, n - base)
        # group_broadcast: member 0 and member 10 (where it exists)
        ok &= np.array_equal(O.group_move(1, O.MOVE_BROADCAST, 0, x, None, L), base.astype(dtype))
        if L > 10:
            ok &= np.array_equal(O.group_move(1, O.MOVE_BROADCAST, 10, x, None, L), (base + 10).astype(dtype))
        # sub_group_broadcast of member 0 and 7 of the 64-wide sub-group
        sb = base + (lid // 64) * 64
        ok &= np.array_equal(O.group_move(0, O.MOVE_BROADCAST, 0, x, None, L), sb.astype(dtype))
        has7 = np.minimum(64, gsize - (lid // 64) * 64) > 7
        got = O.group_move(0, O.MOVE_BROADCAST, 7, x, None, L)
        ok &= np.array_equal(got[has7], (sb + 7)[has7].astype(dtype))
        # group shifts, permute, select where the source is in the group
        left = O.group_move(1, O.MOVE_SHIFT_LEFT, 1, x, None, L)
        m = lid + 1 < gsize
        ok &= np.array_equal(left[m], (i + 1)[m].astype(dtype))
        right = O.group_move(1, O.MOVE_SHIFT_RIGHT, 1, x, None, L)
        m = lid >= 1
        ok &= np.array_equal(right[m], (i - 1)[m].astype(dtype))
        perm = O.group_move(1, O.MOVE_PERMUTE_XOR, 1, x, None, L)
        m = (lid ^ 1) < gsize  # x[i ^ 1] in the reference's first group: base + (lid ^ 1)
        ok &= np.array_equal(perm[m], (base + (lid ^ 1))[m].astype(dtype))
        sel = O.group_move(1, O.MOVE_SELECT, 0, x, np.zeros(n, dtype=np.uint32), L)
        ok &= np.array_equal(sel, base.astype(dtype))
        # sub-group shift left
        sl = O.group_move(0, O.MOVE_SHIFT_LEFT, 1, x, None, L)
        sub_end = np.minimum(sb + 64, base + gsize)
        m = i + 1 < sub_end
        ok &= np.array_equal(sl[m], (i + 1)[m].astype(dtype))
        # any / all / none over x > param, per group
        for thr in (-1, L // 2, n):
            grp = [x[b:b + L] > thr for b in range(0, n, L)]
            anyv = np.concatenate([np.full(g.size, g.any()) for g in grp]).astype(dtype)
            allv = np.concatenate([np.full(g.size, g.all()) for g in grp]).astype(dtype)
            ok &= np.array_equal(O.group_move(1, O.MOVE_ANY, thr, x, None, L), anyv)
            ok &= np.array_equal(O.group_move(1, O.MOVE_ALL, thr, x, None, L), allv)
            ok &= np.array_equal(O.group_move(1, O.MOVE_NONE, thr, x, None, L), (1 - anyv).astype(dtype))
    return bool(ok)


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.float32, np.float64])
def test_reference_cases(dtype):
    assert cases_hold(dtype)


def test_planted_fault_is_caught():
    O.set_fault(16)
    try:
        assert not cases_hold()
    finally:
        O.set_fault(0)
