"""Learned cost model: a small MLP that predicts the run time of a launch
configuration on a problem shape, trained and evaluated on the GPU by the
library's own HIP kernels (acpp_mlp_adam_step / acpp_mlp_forward,
opensycl_amd/csrc/mlp.hip).

Inputs are 16 features (shape features followed by configuration
features, see ``opensycl_amd.tuner.features``), standardised with the
training set's mean and deviation; the target is the standardised natural
log of the measured time in microseconds.  Nothing here has a counterpart in
the reference, whose launch geometry is fixed (numeric.hpp:144-156,
decoupled_lookback_scan.hpp:481-488).
"""
from __future__ import annotations

import ctypes
import json
import math
import os
from typing import Optional

import numpy as np

from . import _abi
from ._abi import MLP_HIDDEN, MLP_INPUTS, MLP_PARAMS, check

_OFF_W1 = 0
_OFF_B1 = MLP_HIDDEN * MLP_INPUTS
_OFF_W2 = _OFF_B1 + MLP_HIDDEN
_OFF_B2 = _OFF_W2 + MLP_HIDDEN * MLP_HIDDEN
_OFF_W3 = _OFF_B2 + MLP_HIDDEN
_OFF_B3 = _OFF_W3 + MLP_HIDDEN


def glorot_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights and zero biases, float32."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_PARAMS, dtype=np.float32)
    h, d = MLP_HIDDEN, MLP_INPUTS
    p[_OFF_W1:_OFF_B1] = rng.uniform(-1, 1, h * d) * math.sqrt(6.0 / (d + h))
    p[_OFF_W2:_OFF_B2] = rng.uniform(-1, 1, h * h) * math.sqrt(6.0 / (h + h))
    p[_OFF_W3:_OFF_B3] = rng.uniform(-1, 1, h) * math.sqrt(6.0 / (h + 1))
    return p


def _torch():
    import torch
    return torch


class CostModel:
    """MLP cost model with its input/target standardisation."""

    def __init__(self, params: Optional[np.ndarray] = None, x_mean=None, x_std=None,
                 y_mean: float = 0.0, y_std: float = 1.0, seed: int = 0):
        self.params = glorot_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        if self.params.shape != (MLP_PARAMS,):
            raise ValueError(f"expected {MLP_PARAMS} parameters")
        self.x_mean = np.zeros(MLP_INPUTS) if x_mean is None else np.asarray(x_mean, dtype=np.float64)
        self.x_std = np.ones(MLP_INPUTS) if x_std is None else np.asarray(x_std, dtype=np.float64)
        self.y_mean = float(y_mean)
        self.y_std = float(y_std)
        self.history: list[float] = []

    # -- standardisation -------------------------------------------------
    def _x(self, X: np.ndarray) -> np.ndarray:
        X = np.asarray(X, dtype=np.float64).reshape(-1, MLP_INPUTS)
        return ((X - self.x_mean) / self.x_std).astype(np.float32)

    # -- training on the GPU ---------------------------------------------
    def fit(self, X: np.ndarray, time_us: np.ndarray, steps: int = 3000, lr: float = 3e-3,
            weight_decay: float = 1e-4, device: int = 0, log_every: int = 0) -> list[float]:
        """Full-batch Adam on (features, measured time) pairs with the HIP
        training kernels.  Returns the loss every ``log_every`` steps."""
        from .algorithms import AllocationCache, AllocationGroup, Queue
        torch = _torch()
        X = np.asarray(X, dtype=np.float64)
        y = np.log(np.asarray(time_us, dtype=np.float64))
        self.x_mean = X.mean(0)
        self.x_std = np.where(X.std(0) > 1e-12, X.std(0), 1.0)
        self.y_mean = float(y.mean())
        self.y_std = float(y.std()) if y.std() > 1e-12 else 1.0
        xd = torch.from_numpy(self._x(X)).to(f"cuda:{device}")
        td = torch.from_numpy(((y - self.y_mean) / self.y_std).astype(np.float32)).to(f"cuda:{device}")
        pd = torch.from_numpy(self.params.copy()).to(f"cuda:{device}")
        md = torch.zeros_like(pd)
        vd = torch.zeros_like(pd)
        loss = torch.zeros(1, dtype=torch.float32, device=f"cuda:{device}")
        lib = _abi.load()
        q = Queue(device=device)
        cache = AllocationCache()
        hist = []
        with AllocationGroup(cache, device) as g:
            for step in range(1, steps + 1):
                check(lib.acpp_mlp_adam_step(g.handle, q.handle, pd.data_ptr(), md.data_ptr(), vd.data_ptr(),
                                             step, xd.data_ptr(), td.data_ptr(), xd.shape[0], ctypes.c_float(lr),
                                             ctypes.c_float(weight_decay),
                                             loss.data_ptr() if log_every and step % log_every == 0 else None))
                if log_every and step % log_every == 0:
                    q.wait()
                    hist.append(float(loss.item()))
                g.release()
            q.wait()
        cache.close()
        self.params = pd.cpu().numpy()
        self.history = hist
        return hist

    # -- inference on the GPU --------------------------------------------
    def predict_log(self, X: np.ndarray, device: int = 0) -> np.ndarray:
        """Predicted natural log of time in microseconds, via the HIP
        forward kernel."""
        from .algorithms import Queue
        torch = _torch()
        xs = self._x(X)
        xd = torch.from_numpy(xs).to(f"cuda:{device}")
        pd = torch.from_numpy(self.params).to(f"cuda:{device}")
        yd = torch.empty(xs.shape[0], dtype=torch.float32, device=f"cuda:{device}")
        q = Queue(device=device)
        check(_abi.load().acpp_mlp_forward(q.handle, pd.data_ptr(), xd.data_ptr(), xs.shape[0], yd.data_ptr()))
        q.wait()
        return yd.cpu().numpy().astype(np.float64) * self.y_std + self.y_mean

    def predict_us(self, X: np.ndarray, device: int = 0) -> np.ndarray:
        return np.exp(self.predict_log(X, device))

    # -- persistence ------------------------------------------------------
    def to_dict(self) -> dict:
        return {"params": self.params.tolist(), "x_mean": self.x_mean.tolist(), "x_std": self.x_std.tolist(),
                "y_mean": self.y_mean, "y_std": self.y_std}

    @staticmethod
    def from_dict(d: dict) -> "CostModel":
        return CostModel(np.array(d["params"], dtype=np.float32), d["x_mean"], d["x_std"], d["y_mean"],
                         d["y_std"])

    def save(self, path: str, extra: Optional[dict] = None) -> None:
        d = self.to_dict()
        if extra:
            d.update(extra)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f)
        os.replace(tmp, path)

    @staticmethod
    def load(path: str) -> "CostModel":
        with open(path) as f:
            return CostModel.from_dict(json.load(f))
