"""Allocation-group entry points captured into a hipGraph and replayed.

A scan, a reduction and a copy_if on one allocation group are captured with
torch's graph capture, after one eager call of each has put blocks of the
right sizes into the cache (capture must not allocate).  Every replay, with
new input written between replays and eager calls on the same group in
between, must give what the numpy restatement gives for the same input, bit
for bit.  Under capture the library clears each sync region with a kernel
node, so every replay starts from a clean region whatever ran before it.
"""
import numpy as np
import pytest

import opensycl_amd as acpp

pytestmark = pytest.mark.gpu


def _expected(xn):
    scan = np.cumsum(xn.astype(np.int64)).astype(np.int32)  # int32 wrap-around
    total = np.int64(xn.astype(np.int64).sum()).astype(np.int32)
    kept = xn[(xn & 1) == 0]
    return scan, total, kept


def test_group_calls_replay_under_graph_capture(gpu):
    torch, cache = gpu
    n = (1 << 22) + 333
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    x = torch.randint(-(1 << 20), (1 << 20) + 1, (n,), dtype=torch.int32, device="cuda", generator=gen)
    ys = torch.empty_like(x)
    yr = torch.zeros(1, dtype=torch.int32, device="cuda")
    yc = torch.empty_like(x)
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")

    side = torch.cuda.Stream()
    qs = acpp.Queue(side)
    g = acpp.AllocationGroup(cache, 0)

    def calls(d_scan, d_red, d_cif, d_cnt):
        acpp.inclusive_scan(qs, g, x, d_scan)
        acpp.reduce(qs, g, x, d_red)
        acpp.copy_if(qs, g, x, d_cif, acpp.PRED_MASK_EQ, 0, 1, d_cnt.view(torch.uint64))

    # Eager warm-up on the capture stream: the cache then holds a block of
    # every size the captured calls need.
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        calls(ys, yr, yc, cnt)
    qs.wait()
    g.release()

    print("capturing scan + reduce + copy_if", flush=True)
    graph = torch.cuda.CUDAGraph()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            calls(ys, yr, yc, cnt)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()

    es = torch.empty_like(x)
    er = torch.zeros(1, dtype=torch.int32, device="cuda")
    ec = torch.empty_like(x)
    ecnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    for rep in range(4):
        print(f"replay {rep}", flush=True)
        x.copy_(torch.randint(-(1 << 20), (1 << 20) + 1, (n,), dtype=torch.int32, device="cuda", generator=gen))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        acpp.check_asynchronous_errors(0)
        scan, total, kept = _expected(x.cpu().numpy())
        assert np.array_equal(ys.cpu().numpy(), scan), rep
        assert int(yr.item()) == int(total), rep
        k = int(cnt.item())
        assert k == kept.size, rep
        assert np.array_equal(yc[:k].cpu().numpy(), kept), rep
        # eager calls on the same group between replays (new regions)
        with torch.cuda.stream(side):
            calls(es, er, ec, ecnt)
        qs.wait()
        assert torch.equal(es, ys) and torch.equal(er, yr) and int(ecnt.item()) == k, rep
        assert torch.equal(ec[:k], yc[:k]), rep
    assert g.watchdog_trips() == 0
    del graph
    torch.cuda.synchronize()
    g.close()
