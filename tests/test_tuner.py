"""Host logic of the autotuner and learned tuner (no GPU): features, shape
sampling, regret bookkeeping, model persistence; and the committed
MI355X timing database + trained model, evaluated with the FP64 oracle's
forward pass."""
import json
import math
import os

import numpy as np
import pytest

from opensycl_amd import tuner as T
from opensycl_amd.cost_model import CostModel
import oracle.mlp as M


def test_config_space_and_default():
    assert len(T.CONFIGS) == 15
    assert T.DEFAULT_CONFIG in T.CONFIGS
    assert all(wg % 64 == 0 and 64 <= wg <= 1024 for wg, _ in T.CONFIGS)


def test_features_shape_and_values():
    s = T.Shape("scan", "i64", "plus", [0, 3, 1000, 1])
    f = T.features(s, (256, 4))
    assert f.shape == (16,)
    assert f[0] == pytest.approx(math.log2(1005))
    assert f[1] == pytest.approx(math.log2(5))
    assert f[6] == pytest.approx(0.25)        # empty fraction
    assert f[7] == 3 and f[8] == 0 and f[9] == 1 and f[10] == 1
    assert f[11] == 8 and f[12] == 2
    assert f[13] == pytest.approx(math.log2(256 * 4 * 2))


def test_sample_shapes_bounded():
    rng = np.random.default_rng(0)
    for _ in range(200):
        s = T.sample_shape(rng)
        assert 1 <= s.n <= 1 << 28
        assert s.num_segments >= 1
        assert max(s.lengths) <= 1 << 20
        off = s.offsets()
        assert off[0] == 0 and off[-1] == s.n and np.all(np.diff(off.astype(np.int64)) >= 0)


def test_lognormal_workload():
    s = T.lognormal_workload("scan", "i64")
    assert s.num_segments == 65536
    assert min(s.lengths) >= 1 and max(s.lengths) <= 1 << 20
    assert 1 << 26 < s.n < 1 << 29


def test_regret_bookkeeping():
    d = "%dx%d" % T.DEFAULT_CONFIG
    db = [{"shape": {}, "times_us": {d: 10.0, "1024x8": 5.0, "64x2": 20.0}},
          {"shape": {}, "times_us": {d: 8.0, "1024x8": 8.0, "64x2": 4.0}}]
    r = T.regret(db, [(1024, 8), T.DEFAULT_CONFIG])
    assert r["mean_regret"] == pytest.approx((0.0 + 1.0) / 2)
    assert r["default_mean_regret"] == pytest.approx((1.0 + 1.0) / 2)
    assert r["geomean_speedup_vs_default"] == pytest.approx(math.sqrt(2.0 * 1.0))


def test_model_roundtrip(tmp_path):
    m = CostModel(seed=3)
    m.x_mean = np.arange(16.0)
    path = str(tmp_path / "m.json")
    m.save(path, {"note": 1})
    m2 = CostModel.load(path)
    assert np.array_equal(m.params, m2.params) and np.array_equal(m.x_mean, m2.x_mean)


DB = os.path.join(T.DATA_DIR, "segmented_timings.json")


@pytest.mark.skipif(not (os.path.exists(DB) and os.path.exists(T.DEFAULT_MODEL)),
                    reason="timing database / trained model not committed")
def test_committed_model_regret_on_held_out_shapes():
    # The committed model was trained on the first 400 records of the MI355X
    # timing database; the last 100 are held out.  Predictions here use the
    # FP64 oracle forward pass with the committed parameters.
    db = T.load_db(DB)
    with open(T.DEFAULT_MODEL) as f:
        meta = json.load(f)
    held = db[meta["train_records"]:]
    assert len(held) >= 100
    m = CostModel.load(T.DEFAULT_MODEL)
    picks = []
    for rec in held:
        s = T.record_shape(rec)
        X = np.array([T.features(s, c) for c in T.CONFIGS])
        y, _, _ = M.forward(m.params, (X - m.x_mean) / m.x_std)
        picks.append(T.CONFIGS[int(np.argmin(y))])
    r = T.regret(held, picks)
    assert r["mean_regret"] <= 0.10, r
    # the learned choice beats the library's fixed default on unseen shapes
    assert r["mean_regret"] < r["default_mean_regret"], r
    assert r["geomean_speedup_vs_default"] > 1.0, r


def test_library_default_rule_on_committed_timings():
    # The size rule the library applies without a configuration
    # (segmented.hip default_work_group / default_rows) against the best of
    # the 15 configurations on the committed MI355X timings: held-out mean
    # regret below 10 %, and better than the former single fixed tile.
    db = T.load_db(os.path.join(T.DATA_DIR, "segmented_timings.json"))
    held = db[400:]
    picks = [T.default_config(r["shape"]["kind"], r["shape"]["n"], r["shape"]["dtype"]) for r in held]
    r = T.regret(held, picks)
    assert r["mean_regret"] < 0.10, r
    assert r["mean_regret"] < r["fixed_512x4_mean_regret"], r
    assert T.default_config("scan", 1 << 28, "i64") == (1024, 8)
    assert T.default_config("scan", 1 << 24, "i32") == (512, 8)
    assert T.default_config("scan", 1 << 18, "i64") == (1024, 2)
    assert T.default_config("reduce", 1 << 28, "f32") == (256, 4)
    assert T.default_config("reduce", 1 << 28, "f64") == (512, 4)
    assert T.default_config("reduce", 1000) == (128, 2)
