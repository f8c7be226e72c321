#!/usr/bin/env python3
"""Benchmark of the device-wide primitives on one MI355X.

Workload (the bandwidth-scale configuration):
  exclusive_scan(plus, init 0) and reduce(plus, init 0) over n = 2^28 int32
  drawn uniform in [-2^15, 2^15], plus transform_reduce (dot product) of two
  2^28 fp32 vectors drawn N(0,1).  One step = one pass of the three calls.

value = algorithmic bytes of one step x steps / wall time of the timed steps
        (inputs resident in HBM), in GB/s.
Algorithmic bytes per call: exclusive_scan 8n, reduce 4n, dot 8n.

The roofline object describes the dominant kernel (the scan): achieved GB/s
from HIP events recorded around every scan launch inside the timed region.
The cpu_baseline object times the oracle's restatement of the reference's
OpenMP-backend algorithms on this host's cores for the same workload and
checks the GPU results against it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "scan/reduce HBM GB/s vs ~8 TB/s peak; segmented-batch speedup vs plain launches"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--skip-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--ops", default="scan,reduce,dot",
                    help="comma-separated subset of scan,reduce,dot (profiling runs)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import opensycl_amd as acpp

    torch.cuda.set_device(0)
    n = 1 << args.log2n
    ops = [o for o in args.ops.split(",") if o]
    log(f"n=2^{args.log2n}, ops={ops}, steps={args.steps}, warmup={args.warmup}")

    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    x = torch.randint(-(1 << 15), (1 << 15) + 1, (n,), dtype=torch.int32, device="cuda",
                      generator=gen)
    y = torch.empty_like(x)
    gen.manual_seed(1)
    a = torch.randn(n, dtype=torch.float32, device="cuda", generator=gen)
    gen.manual_seed(2)
    b = torch.randn(n, dtype=torch.float32, device="cuda", generator=gen)
    red_out = torch.zeros(1, dtype=torch.int32, device="cuda")
    dot_out = torch.zeros(1, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    cache = acpp.AllocationCache()
    scratch = acpp.AllocationGroup(cache, 0)
    q = acpp.Queue()
    bytes_per = {"scan": 8 * n, "reduce": 4 * n, "dot": 8 * n}

    def run_op(op):
        if op == "scan":
            acpp.exclusive_scan(q, scratch, x, y, np.int32(0), acpp.PLUS)
        elif op == "reduce":
            acpp.reduce(q, scratch, x, red_out, np.int32(0), acpp.PLUS)
        else:
            acpp.transform_reduce(q, scratch, a, b, dot_out, 0.0)

    for _ in range(args.warmup):
        for op in ops:
            run_op(op)
    torch.cuda.synchronize()
    log("warmup done")

    stream = torch.cuda.current_stream()
    ev = {op: [] for op in ops}
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record(stream)
    for _ in range(args.steps):
        for op in ops:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            run_op(op)
            e1.record(stream)
            ev[op].append((e0, e1))
    t_end.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_end)
    log(f"timed {args.steps} steps in {total_ms:.3f} ms")

    per_op = {}
    for op in ops:
        ms = [e0.elapsed_time(e1) for e0, e1 in ev[op]]
        avg = sum(ms) / len(ms)
        per_op[op] = {
            "avg_us": round(avg * 1e3, 2),
            "min_us": round(min(ms) * 1e3, 2),
            "algorithmic_bytes": bytes_per[op],
            "gbs": round(bytes_per[op] / (avg * 1e-3) / 1e9, 1),
            "frac_of_peak": round(bytes_per[op] / (avg * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
        }

    step_bytes = sum(bytes_per[op] for op in ops)
    ms_per_step = total_ms / args.steps
    value = step_bytes * args.steps / (total_ms * 1e-3) / 1e9

    dom = "scan" if "scan" in ops else ops[0]
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get(f"{dom}_2^{args.log2n}")
        except Exception:
            traffic = None
    roofline = {
        "kernel": {"scan": "scan_kernel<int32, plus, 8 rows>",
                   "reduce": "reduce_kernel<int32, plus>",
                   "dot": "reduce_kernel<float, dot>"}[dom],
        "bound": "hbm",
        "achieved": per_op[dom]["gbs"],
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(per_op[dom]["gbs"] / PEAK_HBM_GBS, 4),
        "traffic": traffic,
    }

    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+fp32",
        "data": "synthetic (torch device RNG: int32 U[-2^15,2^15] seed 0; fp32 N(0,1) seeds 1, 2)",
        "config": {
            "workload": "exclusive_scan + reduce over 2^%d int32, transform_reduce dot of two "
                        "2^%d fp32" % (args.log2n, args.log2n),
            "n": n,
            "ops": ops,
            "step_algorithmic_bytes": step_bytes,
        },
        "ops": per_op,
        "roofline": roofline,
        "segmented_speedup": None,
    }

    # GPU results of the last step, checked below against the CPU leg.
    torch.cuda.synchronize()
    if not args.skip_cpu:
        result["cpu_baseline"] = cpu_baseline(x, y, a, b, red_out, dot_out, ops, bytes_per)
    else:
        result["cpu_baseline"] = None
    scratch.close()
    cache.close()
    print(json.dumps(result), flush=True)


def cpu_baseline(x, y, a, b, red_out, dot_out, ops, bytes_per):
    """Time the oracle's OpenMP-order restatement of the reference's CPU path
    on the same data (the full 2^28 workload), and check the GPU results."""
    import numpy as np

    import oracle as O

    threads = O.max_threads()
    log(f"cpu baseline on {threads} threads")
    xh = x.cpu().numpy()
    ah = a.cpu().numpy()
    bh = b.cpu().numpy()
    times = {}
    checks = {}
    if "scan" in ops:
        t0 = time.perf_counter()
        yc = O.scan_omp_order(xh, True, np.int32(0), O.PLUS, threads)
        times["scan"] = time.perf_counter() - t0
        checks["scan_bit_exact"] = bool(np.array_equal(y.cpu().numpy(), yc))
        del yc
    if "reduce" in ops:
        t0 = time.perf_counter()
        rc = O.reduce_omp_order(xh, np.int32(0), O.PLUS, threads)
        times["reduce"] = time.perf_counter() - t0
        checks["reduce_bit_exact"] = bool(int(red_out.item()) == int(rc))
    if "dot" in ops:
        t0 = time.perf_counter()
        dc = O.dot_omp_order(ah, bh, np.float32(0), threads)
        times["dot"] = time.perf_counter() - t0
        exact = O.dot_fp64(ah, bh)
        checks["dot_rel_err_vs_fp64_gpu"] = abs(dot_out.item() - exact) / max(abs(exact), 1e-30)
        checks["dot_rel_err_vs_fp64_cpu"] = abs(float(dc) - exact) / max(abs(exact), 1e-30)
    total_bytes = sum(bytes_per[o] for o in times)
    total_t = sum(times.values())
    log(f"cpu baseline: {total_t:.2f} s, checks {checks}")
    return {
        "value": round(total_bytes / total_t / 1e9, 2),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": "one full step of the same workload (2^28 elements per call) on the GPU "
                  "results' own inputs; OpenMP restatement of the reference's OMP-backend "
                  "reduce (8 items/work-item, 128-wide trees) and decoupled look-back scan "
                  "(1024x16 tiles)",
        "seconds": {k: round(v, 3) for k, v in times.items()},
        "checks": checks,
    }


if __name__ == "__main__":
    main()
