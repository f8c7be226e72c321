"""Learned offload decision of the stdpar entry points.

The reference decides whether a std::*(par_unseq) call runs on the device by
interpolating sampled host and device run times linearly between the two
nearest sampled problem sizes (include/hipSYCL/std/stdpar/detail/
offload_heuristic_db.hpp:247-318, used by offload.hpp:384-594).  Here the
library's MLP (the acpp_mlp_* HIP kernels, the same network as the launch
cost model) predicts the log run time of a call on each side from the
operation, element type and size; the library evaluates it once per device
into a table of crossover sizes (opensycl_amd/csrc/stdpar_offload.cpp).

  collect()  -- measures host (sequential std::) and offloaded run times of
                every operation class and element type over a range of sizes,
                on pinned host memory, with the native sampler
                (opensycl_amd/bench/stdpar_sample.cpp);
  train()    -- fits the MLP on the GPU with the HIP training kernels;
  features() -- the 16 model inputs, the same as stdpar_features() in
                stdpar_offload.cpp.
The trained model is stored in opensycl_amd/data/offload_model.json, from
which the build compiles it into the library.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time

import numpy as np

from ._abi import MLP_INPUTS
from .build import LIB_DIR

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
MODEL_PATH = os.path.join(DATA, "offload_model.json")
SAMPLES_PATH = os.path.join(DATA, "offload_samples.json")

OPS = {"reduce": 0, "transform_reduce": 1, "binary_transform_reduce": 2, "inclusive_scan": 3,
       "exclusive_scan": 4}
DTYPES = {"int32": (0, 4, 0), "int64": (1, 8, 0), "float32": (2, 4, 1), "float64": (3, 8, 1)}


def features(op: int, elem_bytes: int, is_float: int, log2n: float, device_side: int) -> np.ndarray:
    """Raw (unstandardised) inputs of one row; stdpar_features() in
    stdpar_offload.cpp builds the same values."""
    lb = log2n + math.log2(elem_bytes)
    f = np.zeros(MLP_INPUTS)
    f[0] = log2n
    f[1] = log2n * log2n / 40.0
    f[2] = lb
    f[3] = elem_bytes
    f[4] = float(is_float)
    f[5 + op] = 1.0
    f[10] = float(device_side)
    f[11] = log2n if device_side else 0.0
    f[12] = lb if device_side else 0.0
    return f


def _sampler():
    lib = ctypes.CDLL(os.path.join(LIB_DIR, "libacpp_bench.so"))
    fn = lib.acpp_bench_stdpar_sample
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_int
    return fn


def sample(op: str, dtype: str, n: int, reps: int = 5) -> tuple[float, float]:
    """(host microseconds, device microseconds) of one call, medians."""
    fn = _sampler()
    h, d = ctypes.c_double(), ctypes.c_double()
    code, _, _ = DTYPES[dtype]
    rc = fn(OPS[op], code, n, reps, ctypes.byref(h), ctypes.byref(d))
    if rc != 0:
        raise RuntimeError(f"stdpar sampler failed ({rc}) for {op} {dtype} n={n}")
    return h.value, d.value


def collect(log2_sizes=None, reps: int = 5, log=print) -> list[dict]:
    """Samples of every operation and element type; half-integer log2 sizes
    are included so that the model sees sizes between powers of two."""
    if log2_sizes is None:
        log2_sizes = [x / 2 for x in range(8, 53)]  # 2^4 .. 2^26
    # torch's HIP runtime must be initialised before the sampler's: the other
    # order leaves torch without a device in this process.
    import torch
    torch.cuda.init()
    rows = []
    t0 = time.time()
    for op in OPS:
        for dtype in DTYPES:
            for l2 in log2_sizes:
                n = int(round(2 ** l2))
                h, d = sample(op, dtype, n, reps)
                rows.append({"op": op, "dtype": dtype, "n": n, "host_us": h, "device_us": d})
            log(f"[offload] {op} {dtype}: done ({time.time() - t0:.0f} s)")
    return rows


def design(rows: list[dict]) -> tuple[np.ndarray, np.ndarray]:
    """(features, time in us) for both sides of every sample."""
    X, y = [], []
    for r in rows:
        _, b, fl = DTYPES[r["dtype"]]
        l2 = math.log2(r["n"])
        for side, t in ((0, r["host_us"]), (1, r["device_us"])):
            X.append(features(OPS[r["op"]], b, fl, l2, side))
            y.append(t)
    return np.array(X), np.array(y)


def held_out(rows: list[dict]) -> tuple[list[dict], list[dict]]:
    """Training rows: integer log2 sizes; evaluation rows: the half-integer
    sizes between them, which the model never sees."""
    train, test = [], []
    for r in rows:
        (train if abs(math.log2(r["n"]) - round(math.log2(r["n"]))) < 1e-9 else test).append(r)
    return train, test


def train(rows: list[dict], steps: int = 4000, seed: int = 0):
    from .cost_model import CostModel
    X, y = design(rows)
    m = CostModel(seed=seed)
    m.fit(X, y, steps=steps, lr=3e-3, weight_decay=1e-5)
    return m


def decisions(model, rows: list[dict]) -> np.ndarray:
    """1 where the model offloads the sampled call."""
    X, _ = design(rows)
    pred = model.predict_log(X).reshape(-1, 2)  # (host, device) per row
    return (pred[:, 1] < pred[:, 0]).astype(int)


def regret(rows: list[dict], offload: np.ndarray) -> np.ndarray:
    """t(chosen) / t(best) - 1 per sample."""
    h = np.array([r["host_us"] for r in rows])
    d = np.array([r["device_us"] for r in rows])
    chosen = np.where(offload == 1, d, h)
    return chosen / np.minimum(h, d) - 1.0


def interpolation_decisions(train_rows: list[dict], rows: list[dict]) -> np.ndarray:
    """The reference's rule on the same samples: linear interpolation of the
    two sampled sizes nearest to n, per operation and element type, for each
    side (offload_heuristic_db.hpp:247-318)."""
    out = []
    for r in rows:
        pts = sorted((t["n"], t["host_us"], t["device_us"]) for t in train_rows
                     if t["op"] == r["op"] and t["dtype"] == r["dtype"])
        near = sorted(pts, key=lambda p: abs(p[0] - r["n"]))[:2]
        (n0, h0, d0), (n1, h1, d1) = sorted(near)
        w = (r["n"] - n0) / (n1 - n0)
        h = h0 + w * (h1 - h0)
        d = d0 + w * (d1 - d0)
        out.append(1 if d < h else 0)
    return np.array(out)


def save(model, rows: list[dict], extra: dict | None = None) -> None:
    d = model.to_dict()
    d["kind"] = "stdpar_offload"
    if extra:
        d.update(extra)
    with open(MODEL_PATH + ".tmp", "w") as f:
        json.dump(d, f)
    os.replace(MODEL_PATH + ".tmp", MODEL_PATH)
    with open(SAMPLES_PATH + ".tmp", "w") as f:
        json.dump(rows, f)
    os.replace(SAMPLES_PATH + ".tmp", SAMPLES_PATH)


def main() -> None:
    """Collect samples, train, report held-out regret, save (GPU box)."""
    rows = collect()
    with open(SAMPLES_PATH + ".tmp", "w") as f:
        json.dump(rows, f)
    os.replace(SAMPLES_PATH + ".tmp", SAMPLES_PATH)
    tr, te = held_out(rows)
    model = train(tr)
    dec = decisions(model, te)
    reg = regret(te, dec)
    base = regret(te, np.ones(len(te), dtype=int))
    interp = regret(te, interpolation_decisions(tr, te))
    report = {"held_out": len(te), "mean_regret": float(reg.mean()), "max_regret": float(reg.max()),
              "always_offload_mean_regret": float(base.mean()),
              "interpolation_mean_regret": float(interp.mean())}
    print(json.dumps(report))
    save(model, rows, {"report": report})


if __name__ == "__main__":
    main()
