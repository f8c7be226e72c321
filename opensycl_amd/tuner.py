"""Autotuner for the segmented scans and reductions, and the learned
configuration choice built on it.

* ``CONFIGS`` -- the search space: work-group size x rows (16-byte vectors
  per work-item) of the persistent segmented kernel.
* ``autotune`` -- exhaustive search: times every configuration on one
  problem shape on the GPU (median of several launches, HIP events on the
  queue's stream) and returns the timings.
* ``features`` -- the cost model's 16 inputs for (shape, configuration):
  element count, segment count, moments of the log segment lengths, element
  type, operation, and the configuration's tile geometry.
* ``collect`` -- timing database over random shapes (the training and
  held-out data of the cost model).
* ``LearnedTuner`` -- picks the configuration with the lowest time predicted
  by the cost model (opensycl_amd.cost_model), evaluated by the HIP forward
  kernel; ``regret`` compares its picks with the exhaustive optimum.

No counterpart in the reference: its launch geometry is fixed
(numeric.hpp:144-156; decoupled_lookback_scan.hpp:481-488).

Command line (on a GPU):
  python -m opensycl_amd.tuner collect --shapes 500 --out timings.json
  python -m opensycl_amd.tuner train --db timings.json --out cost_model.json
  python -m opensycl_amd.tuner evaluate --db timings.json --model cost_model.json
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from dataclasses import asdict, dataclass
from typing import Optional

import numpy as np

from ._abi import MLP_INPUTS
from .cost_model import CostModel

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_MODEL = os.path.join(DATA_DIR, "cost_model.json")

WORK_GROUP_SIZES = (64, 128, 256, 512, 1024)
ROWS = (2, 4, 8)
CONFIGS = [(wg, r) for wg in WORK_GROUP_SIZES for r in ROWS]
# The single fixed tile the library used before its size rule, kept as a
# comparison point.
DEFAULT_CONFIG = (512, 4)


def default_config(kind: str, n: int, dtype: str = "i32") -> tuple:
    """The library's own choice when no configuration is given: the size
    and element-size rule of segmented.hip (default_work_group /
    default_rows)."""
    wide = DTYPES.get(dtype, 4) == 8
    if n < (1 << 16):
        return (128, 2) if kind == "reduce" else (256, 2)
    if n < (1 << 20):
        return (256, 2) if kind == "reduce" else ((1024, 2) if wide else (512, 2))
    if kind == "reduce":
        return (1024, 2) if n < (1 << 24) else ((512, 4) if wide else (256, 4))
    return (512, 8) if n < (1 << 26) else (1024, 8)


KINDS = ("scan", "reduce")
DTYPES = {"i32": 4, "i64": 8, "f32": 4, "f64": 8}


@dataclass
class Shape:
    kind: str            # "scan" (segmented inclusive scan) or "reduce"
    dtype: str           # i32 / i64 / f32 / f64
    op: str              # "plus" or "max"
    lengths: list        # segment lengths

    @property
    def n(self) -> int:
        return int(np.sum(np.asarray(self.lengths, dtype=np.int64)))

    @property
    def num_segments(self) -> int:
        return len(self.lengths)

    def offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(np.asarray(self.lengths, dtype=np.uint64))]).astype(np.uint64)


@dataclass
class ShapeSummary:
    """What the timing database keeps of a shape: its type, operation, sizes
    and the 11 shape features (the segment lengths themselves are not
    needed by the cost model)."""
    kind: str
    dtype: str
    op: str
    n: int
    num_segments: int
    stats: list


def summarize(shape: Shape) -> ShapeSummary:
    return ShapeSummary(shape.kind, shape.dtype, shape.op, shape.n, shape.num_segments,
                        [float(v) for v in shape_features(shape)])


def shape_features(shape) -> np.ndarray:
    """The 11 shape features (element count, segment statistics, type, op)."""
    if isinstance(shape, ShapeSummary):
        return np.asarray(shape.stats, dtype=np.float64)
    lens = np.asarray(shape.lengths, dtype=np.float64)
    ll = np.log2(lens + 1.0)
    mu = float(ll.mean()) if len(ll) else 0.0
    sd = float(ll.std()) if len(ll) else 0.0
    skew = float(((ll - mu) ** 3).mean() / sd ** 3) if sd > 1e-9 else 0.0
    return np.array([
        math.log2(shape.n + 1.0),
        math.log2(shape.num_segments + 1.0),
        mu,
        sd,
        float(ll.max()) if len(ll) else 0.0,
        float(np.clip(skew, -5, 5)),
        float(np.mean(lens == 0)) if len(lens) else 0.0,
        math.log2(DTYPES[shape.dtype]),
        1.0 if shape.dtype[0] == "f" else 0.0,
        1.0 if shape.kind == "scan" else 0.0,
        1.0 if shape.op == "plus" else 0.0,
    ])


def config_features(shape, cfg) -> np.ndarray:
    """The 5 configuration features (geometry relative to the shape)."""
    wg, rows = cfg
    tile = wg * rows * (16 // DTYPES[shape.dtype])
    tiles = shape.n / tile
    return np.array([
        math.log2(wg),
        math.log2(rows),
        math.log2(tile),
        math.log2(tiles + 1.0),
        math.log2(shape.num_segments / (tiles + 1.0) + 1.0),
    ])


def features(shape, cfg) -> np.ndarray:
    f = np.concatenate([shape_features(shape), config_features(shape, cfg)])
    assert f.shape == (MLP_INPUTS,)
    return f


def sample_shape(rng: np.random.Generator, max_elems: int = 1 << 28) -> Shape:
    """A random segmented workload: log-uniform segment count and median
    length, log-normal spread, clipped to [0 or 1, 2^20] per segment."""
    kind = KINDS[rng.integers(0, 2)]
    dtype = list(DTYPES)[rng.integers(0, 4)]
    op = "plus" if rng.random() < 0.75 else "max"
    S = int(2 ** rng.uniform(0, 16))
    median = 2 ** rng.uniform(0, 14)
    sigma = rng.uniform(0, 2.5)
    lo = 0 if rng.random() < 0.2 else 1
    lens = np.clip(np.round(np.exp(rng.normal(math.log(median), sigma, S))), lo, 1 << 20).astype(np.int64)
    total = np.cumsum(lens)
    if total[-1] > max_elems:
        S = max(1, int(np.searchsorted(total, max_elems)))
        lens = lens[:S]
    if lens.sum() == 0:
        lens[0] = 1
    return Shape(kind, dtype, op, lens.tolist())


def lognormal_workload(kind: str, dtype: str, S: int = 65536, seed: int = 0) -> Shape:
    """65,536 segments, lengths log-normal (median 512, sigma 2) clipped to
    [1, 2^20]."""
    rng = np.random.default_rng(seed)
    lens = np.clip(np.round(np.exp(rng.normal(math.log(512), 2.0, S))), 1, 1 << 20).astype(np.int64)
    return Shape(kind, dtype, "plus", lens.tolist())


# ---------------------------------------------------------------------------
# Timing on the GPU
# ---------------------------------------------------------------------------

def _torch_dtype(tag):
    import torch
    return {"i32": torch.int32, "i64": torch.int64, "f32": torch.float32, "f64": torch.float64}[tag]


class _Problem:
    """Device buffers for one shape."""

    def __init__(self, shape: Shape, seed: int = 0, device: int = 0):
        import torch
        dev = f"cuda:{device}"
        n = shape.n
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        dt = _torch_dtype(shape.dtype)
        if shape.dtype[0] == "f":
            self.x = torch.randn(max(n, 1), dtype=dt, device=dev, generator=g)[:n]
        else:
            self.x = torch.randint(-1000, 1001, (max(n, 1),), dtype=dt, device=dev, generator=g)[:n]
        self.offsets = torch.from_numpy(shape.offsets()).to(dev)
        self.out = torch.empty(n if shape.kind == "scan" else shape.num_segments, dtype=dt, device=dev)
        self.shape = shape

    def run(self, q, scratch, cfg) -> None:
        from . import algorithms as A
        op = A.PLUS if self.shape.op == "plus" else A.MAX
        config = (cfg[0], cfg[1], 0, 0)
        if self.shape.kind == "scan":
            A.segmented_inclusive_scan(q, scratch, self.x, self.out, self.offsets, op, config=config)
        else:
            A.segmented_reduce(q, scratch, self.x, self.offsets, self.out, 0, op, config=config)


def time_config(problem: _Problem, q, scratch, cfg, reps: int = 5) -> float:
    """Median time of `reps` launches in microseconds (HIP events on the
    queue's stream), after one untimed launch."""
    import torch
    problem.run(q, scratch, cfg)
    scratch.release()
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(q.stream)
        problem.run(q, scratch, cfg)
        e1.record(q.stream)
        e1.synchronize()
        scratch.release()
        times.append(e0.elapsed_time(e1) * 1e3)
    return float(np.median(times))


def autotune(shape: Shape, configs=CONFIGS, reps: int = 5, device: int = 0, seed: int = 0) -> dict:
    """Exhaustive search: {config: median time (us)} for every config."""
    from .algorithms import AllocationCache, AllocationGroup, Queue
    q = Queue(device=device)
    cache = AllocationCache()
    prob = _Problem(shape, seed, device)
    out = {}
    with AllocationGroup(cache, device) as scratch:
        for cfg in configs:
            out[tuple(cfg)] = time_config(prob, q, scratch, cfg, reps)
    del prob
    cache.close()
    return out


def _key(cfg) -> str:
    return f"{cfg[0]}x{cfg[1]}"


def _unkey(k: str):
    a, b = k.split("x")
    return (int(a), int(b))


def collect(num_shapes: int, seed: int = 0, reps: int = 5, path: Optional[str] = None,
            device: int = 0, progress: bool = True) -> list:
    """Timing database: the 65,536-segment log-normal workload for every
    kind and element type, then random shapes up to 2^28 elements, one
    record per shape with every config's time.  Written to `path` (JSON) as
    it grows."""
    rng = np.random.default_rng(seed)
    fixed = [lognormal_workload(k, d) for k in KINDS for d in DTYPES]
    db = []
    t0 = time.time()
    for i in range(num_shapes):
        shape = fixed[i] if i < len(fixed) else sample_shape(rng)
        times = autotune(shape, reps=reps, device=device, seed=seed * 100003 + i)
        db.append({"shape": asdict(summarize(shape)), "times_us": {_key(c): t for c, t in times.items()}})
        if progress and (i + 1) % 10 == 0:
            print(f"[tuner] {i + 1}/{num_shapes} shapes timed ({time.time() - t0:.0f} s)", flush=True)
        if path and ((i + 1) % 50 == 0 or i + 1 == num_shapes):
            save_db(db, path)
    return db


def save_db(db: list, path: str) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(db, f)
    os.replace(tmp, path)


def load_db(path: str) -> list:
    with open(path) as f:
        return json.load(f)


def record_shape(rec: dict) -> ShapeSummary:
    return ShapeSummary(**rec["shape"])


def training_set(db: list):
    """(features, time) pairs of every (shape, config) in the database."""
    X, y = [], []
    for rec in db:
        shape = record_shape(rec)
        for k, t in rec["times_us"].items():
            X.append(features(shape, _unkey(k)))
            y.append(t)
    return np.array(X), np.array(y)


# ---------------------------------------------------------------------------
# Learned choice
# ---------------------------------------------------------------------------
class LearnedTuner:
    """Configuration choice by the cost model: the config with the lowest
    predicted time.  The exhaustive autotuner is the fallback (``exhaustive``)
    and the judge of its regret."""

    def __init__(self, model: Optional[CostModel] = None, path: str = DEFAULT_MODEL):
        if model is None:
            if not os.path.exists(path):
                raise FileNotFoundError(f"no trained cost model at {path}")
            model = CostModel.load(path)
        self.model = model

    def choose_many(self, shapes: list, device: int = 0) -> list:
        X = np.array([features(s, c) for s in shapes for c in CONFIGS])
        pred = self.model.predict_log(X, device).reshape(len(shapes), len(CONFIGS))
        return [CONFIGS[int(np.argmin(p))] for p in pred]

    def choose(self, shape: Shape, device: int = 0):
        return self.choose_many([shape], device)[0]

    @staticmethod
    def exhaustive(shape: Shape, device: int = 0):
        t = autotune(shape, device=device)
        return min(t, key=t.get)


class NativeTuner:
    """The same choice through the C-ABI (acpp_cost_model_choose_segmented):
    features computed by the library from host offsets, network evaluated
    by its forward kernel.  What a C or C++ caller uses."""

    _ELEM = {"i32": 0, "i64": 1, "f32": 2, "f64": 3}

    def __init__(self, model: Optional[CostModel] = None, path: str = DEFAULT_MODEL, device: int = 0):
        import ctypes
        from . import _abi
        if model is None:
            model = CostModel.load(path)
        self._lib = _abi.load()
        p = np.ascontiguousarray(model.params, dtype=np.float32)
        xm = np.ascontiguousarray(model.x_mean, dtype=np.float64)
        xs = np.ascontiguousarray(model.x_std, dtype=np.float64)
        h = ctypes.c_void_p()
        _abi.check(self._lib.acpp_cost_model_create(device, p.ctypes.data, xm.ctypes.data, xs.ctypes.data,
                                                    model.y_mean, model.y_std, ctypes.byref(h)))
        self._h = h

    def choose(self, shape: Shape, stream: int = 0):
        import ctypes
        from . import _abi
        off = np.ascontiguousarray(shape.offsets(), dtype=np.uint64)
        cfg = _abi.LaunchConfig()
        pred = ctypes.c_float()
        _abi.check(self._lib.acpp_cost_model_choose_segmented(
            self._h, stream, off.ctypes.data, shape.num_segments, self._ELEM[shape.dtype],
            0 if shape.kind == "scan" else 1, 0 if shape.op == "plus" else 3, ctypes.byref(cfg),
            ctypes.byref(pred)))
        return (cfg.work_group_size, cfg.items_per_thread), float(pred.value)

    def close(self):
        if self._h:
            self._lib.acpp_cost_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def regret(db: list, picks: list) -> dict:
    """Regret t(pick)/t(best) - 1 over the records, with the same statistic
    for the library's default (its size rule, default_config: the method
    without the cost model) and for the former single fixed tile
    (DEFAULT_CONFIG), and the mean speed-up of the picks over the default."""
    r, rd, rf, gain = [], [], [], []
    for rec, pick in zip(db, picks):
        t = rec["times_us"]
        best = min(t.values())
        tp = t[_key(pick)]
        sh = rec.get("shape", {})
        dflt = (default_config(sh["kind"], sh["n"], sh.get("dtype", "i32"))
                if "kind" in sh and "n" in sh else DEFAULT_CONFIG)
        td = t[_key(dflt)]
        tf = t[_key(DEFAULT_CONFIG)]
        r.append(tp / best - 1.0)
        rd.append(td / best - 1.0)
        rf.append(tf / best - 1.0)
        gain.append(td / tp)
    r, rd, rf, gain = np.array(r), np.array(rd), np.array(rf), np.array(gain)
    return {
        "shapes": len(r),
        "mean_regret": float(r.mean()), "median_regret": float(np.median(r)),
        "p90_regret": float(np.quantile(r, 0.9)), "max_regret": float(r.max()),
        "default_mean_regret": float(rd.mean()), "default_median_regret": float(np.median(rd)),
        "default_p90_regret": float(np.quantile(rd, 0.9)),
        "fixed_512x4_mean_regret": float(rf.mean()),
        "geomean_speedup_vs_default": float(np.exp(np.log(gain).mean())),
        "exact_best_fraction": float(np.mean(r < 1e-12)),
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m opensycl_amd.tuner")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--shapes", type=int, default=500)
    c.add_argument("--seed", type=int, default=0)
    c.add_argument("--reps", type=int, default=5)
    c.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--db", required=True)
    t.add_argument("--train", type=int, default=400, help="first N records train, the rest are held out")
    t.add_argument("--steps", type=int, default=4000)
    t.add_argument("--lr", type=float, default=3e-3)
    t.add_argument("--out", required=True)
    e = sub.add_parser("evaluate")
    e.add_argument("--db", required=True)
    e.add_argument("--train", type=int, default=400)
    e.add_argument("--model", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "collect":
        collect(a.shapes, a.seed, a.reps, a.out)
        return 0
    db = load_db(a.db)
    train_db, held = db[:a.train], db[a.train:]
    if a.cmd == "train":
        X, y = training_set(train_db)
        m = CostModel(seed=0)
        hist = m.fit(X, y, steps=a.steps, lr=a.lr, log_every=max(1, a.steps // 20))
        print("[tuner] training loss:", " ".join(f"{v:.4f}" for v in hist), flush=True)
        m.save(a.out, {"train_records": len(train_db), "steps": a.steps, "lr": a.lr,
                       "configs": [list(c) for c in CONFIGS]})
        return 0
    m = CostModel.load(a.model)
    tuner = LearnedTuner(m)
    for name, part in (("train", train_db), ("held-out", held)):
        if not part:
            continue
        picks = tuner.choose_many([record_shape(r) for r in part])
        print(f"[tuner] {name}: " + json.dumps(regret(part, picks)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
