"""The exclusive int32 scan that allocates nothing (acpp_exclusive_scan_i32_noalloc).

Called through the C-ABI with ctypes and compared bit for bit with the
entry point it varies, acpp_exclusive_scan_i32 (scratch from an allocation
group): two calls in a row, the second with a smaller input, so the second
runs on the static scratch the first left behind; then the size limit.
"""
import ctypes

import numpy as np
import pytest

import opensycl_amd as acpp
from opensycl_amd import _abi
from opensycl_amd._abi import LaunchConfig

pytestmark = pytest.mark.gpu


def _noalloc(lib, torch, x, y, init, config=None):
    return lib.acpp_exclusive_scan_i32_noalloc(
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), ctypes.c_void_p(x.data_ptr()),
        ctypes.c_uint64(x.numel()), ctypes.c_void_p(y.data_ptr()), ctypes.c_int32(init),
        ctypes.c_int32(acpp.PLUS), ctypes.byref(config) if config is not None else None)


def test_noalloc_scan_matches_the_group_scan(gpu, q, scratch):
    torch, _ = gpu
    lib = _abi.load()
    g = torch.Generator(device="cuda")
    g.manual_seed(21)
    for n, init in (((1 << 26) + 12345, 7), (300001, -3)):
        print(f"noalloc scan n={n}", flush=True)
        x = torch.randint(-(1 << 15), (1 << 15) + 1, (n,), dtype=torch.int32, device="cuda", generator=g)
        want = torch.empty_like(x)
        acpp.exclusive_scan(q, scratch, x, want, init)
        got = torch.full_like(x, 99)
        _abi.check(_noalloc(lib, torch, x, got, init))
        q.wait()
        assert torch.equal(got, want), n


def test_noalloc_scan_size_limit(gpu, q, scratch):
    # 64 work-items x 2 rows x 4 int32 = 512-element tiles: the static
    # scratch holds 16384 tiles, so 2^23 elements fit and 2^23 + 1 do not
    torch, _ = gpu
    lib = _abi.load()
    cfg = LaunchConfig(64, 2, 0, 0)
    n = 1 << 23
    x = torch.ones(n + 1, dtype=torch.int32, device="cuda")
    got = torch.empty_like(x)
    _abi.check(_noalloc(lib, torch, x[:n], got[:n], 0, cfg))
    want = torch.empty(n, dtype=torch.int32, device="cuda")
    acpp.exclusive_scan(q, scratch, x[:n], want, 0, config=cfg)
    q.wait()
    assert torch.equal(got[:n], want)
    st = _noalloc(lib, torch, x, got, 0, cfg)
    assert st == _abi.ERROR_INVALID_ARGUMENT
    msg = lib.acpp_last_error_message().decode()
    assert "16384 tiles" in msg, msg


def test_noalloc_scan_under_graph_capture(gpu, q, scratch):
    # The callers this entry point is for: a scan captured into a hipGraph
    # (through torch's graph capture) and replayed.  Under capture the
    # library clears its static scratch with a kernel node; every replay,
    # with new input written between replays and eager work in between, must
    # match the allocation-group scan of the same input.
    torch, _ = gpu
    lib = _abi.load()
    n = (1 << 22) + 333
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    x = torch.randint(-(1 << 15), (1 << 15) + 1, (n,), dtype=torch.int32, device="cuda", generator=g)
    y = torch.empty_like(x)
    _abi.check(_noalloc(lib, torch, x, y, 1))  # eager first call: per-device set-up
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            st = _noalloc(lib, torch, x, y, 1)
    assert st == 0, lib.acpp_last_error_message().decode()
    torch.cuda.current_stream().wait_stream(side)
    for rep in range(4):
        x.copy_(torch.randint(-(1 << 15), (1 << 15) + 1, (n,), dtype=torch.int32, device="cuda", generator=g))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        want = torch.empty_like(x)
        acpp.exclusive_scan(q, scratch, x, want, 1)
        q.wait()
        assert torch.equal(y, want), rep
        # eager use of the same static scratch between replays
        z = torch.empty_like(x)
        _abi.check(_noalloc(lib, torch, x, z, 1))
        torch.cuda.synchronize()
        assert torch.equal(z, want), rep
