"""GPU parity: group_broadcast, any/all/none_of_group, joint_*_of,
shift_group_left/right, permute_group_by_xor and select_from_group from
include/acpp/device/group_functions.hpp, for sub-groups (wave64) and
work-groups, against the oracle's restatement of the SYCL 2020 semantics
(oracle_group_move_<t>; include/hipSYCL/sycl/libkernel/group_functions.hpp:
57-140, 240-276), itself pinned by the cases of the reference's
tests/sycl/group_functions/group_functions_misc.cpp
(tests/test_oracle_group_movement.py).

Out-of-range sources: a sub-group lane keeps its own value (the HIP
shuffles); a work-group reads work-item 0 (the reference's LDS fallback,
generic/hiplike/group_functions.hpp:546-640)."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from gpu_util import dev, host

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TAGS = {np.int32: 1, np.int64: 2, np.float32: 4, np.float64: 5}
WHICH = {"broadcast": 0, "shift_left": 1, "shift_right": 2, "permute_xor": 3, "select": 4,
         "any": 5, "all": 6, "none": 7, "joint_any": 8, "joint_all": 9, "joint_none": 10}


@pytest.fixture(scope="module")
def tk():
    lib = ctypes.CDLL(os.path.join(HERE, "lib", "libacpp_test_kernels.so"))
    vp = ctypes.c_void_p
    lib.tk_group_movement.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, vp, vp, vp,
                                      ctypes.c_uint64, ctypes.c_uint32, vp]
    return lib


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.float32, np.float64])
@pytest.mark.parametrize("local_size", [64, 96, 256, 1024])
@pytest.mark.parametrize("level", [0, 1])
def test_movement(gpu, tk, dtype, local_size, level):
    torch, _ = gpu
    rng = np.random.default_rng(local_size + level)
    n = local_size * 4
    x = rng.integers(-1000, 1000, n).astype(dtype)
    src = rng.integers(0, local_size + 70, n).astype(np.uint32)   # some out of range
    xd, sd = dev(torch, x), dev(torch, src)
    od = torch.empty_like(xd)
    q = torch.cuda.current_stream().cuda_stream
    params = {"broadcast": [0, 5, 31], "shift_left": [1, 3, 40], "shift_right": [1, 7, 33],
              "permute_xor": [1, 6, 32, 100], "select": [0], "any": [990, -2000], "all": [-990, -2000],
              "none": [990, 2000], "joint_any": [990], "joint_all": [-2000], "joint_none": [2000]}
    for name, which in WHICH.items():
        for param in params[name]:
            if name == "broadcast" and param >= min(local_size, 64 if level == 0 else local_size):
                continue
            if name == "broadcast" and level == 0 and local_size % 64 and param >= local_size % 64:
                continue  # the partial wave has fewer lanes than `param`
            pv = param if param >= 0 else param  # thresholds may be negative
            rc = tk.tk_group_movement(TAGS[dtype], level, which, ctypes.c_uint32(pv & 0xffffffff),
                                      xd.data_ptr(), sd.data_ptr(), od.data_ptr(), n, local_size, q)
            assert rc == 0
            torch.cuda.synchronize()
            thr = np.array(pv & 0xffffffff, dtype=np.uint32).astype(np.int64)
            thr = thr if thr < 2**31 else thr - 2**32
            got = host(od)
            exp = O.group_move(level, which, int(thr if which >= 5 else pv), x, src, local_size)
            assert np.array_equal(got, exp), (name, param, local_size, level)
