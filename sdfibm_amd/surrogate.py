"""Learned volume-fraction surrogate: training on the GPU, evaluation, parameter files.

The network (11 -> 32 -> 32 -> 1, ReLU; BF16 weights and activations, FP32 accumulation, MFMA
inference) maps a cut cell's 8 corner signed distances / h and the unit SDF gradient at its
centre to the solid volume fraction of the reference's apex method (src/geometrictools.cpp:47-72).
Training data are generated on the device from random spheres and ellipsoids (2-12 cells);
the forward/backward passes and Adam are this library's kernels.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib

N_PARAMS = 1473
N_IN, N_HID = 11, 32
PARAMS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "surrogate_params.npy")

__all__ = ["N_PARAMS", "init_params", "samples", "evaluate", "grad", "train", "load_params", "PARAMS_PATH"]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases (b1 slightly positive to keep the ReLUs alive)."""
    rng = np.random.default_rng(seed)
    w1 = rng.uniform(-1, 1, (N_HID, N_IN)) * np.sqrt(6.0 / N_IN)
    w2 = rng.uniform(-1, 1, (N_HID, N_HID)) * np.sqrt(6.0 / N_HID)
    w3 = rng.uniform(-1, 1, N_HID) * np.sqrt(3.0 / N_HID)
    return np.concatenate([w1.ravel(), np.full(N_HID, 0.01), w2.ravel(), np.full(N_HID, 0.01), w3, [0.5]]
                          ).astype(np.float32)


def samples(seed: int, n: int):
    """(features [n, 16], labels [n]) of the training distribution, generated on the GPU."""
    X = np.zeros((n, 16), dtype=np.float32)
    y = np.zeros(n, dtype=np.float32)
    _lib.check(_lib.load().sdfibm_surrogate_samples(int(seed), int(n), _p(X), _p(y)))
    return X, y


def evaluate(params, X) -> np.ndarray:
    """The deployed MFMA inference on feature rows [n, 16]."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 16)
    out = np.zeros(X.shape[0], dtype=np.float32)
    _lib.check(_lib.load().sdfibm_surrogate_eval(_p(p), _p(X), X.shape[0], _p(out)))
    return out


def grad(params, X, y):
    """(gradient of 0.5 sum (net - y)^2, sum (net - y)^2) from the training kernels."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 16)
    y = np.ascontiguousarray(y, dtype=np.float32)
    g = np.zeros(N_PARAMS, dtype=np.float32)
    ss = np.zeros(1)
    _lib.check(_lib.load().sdfibm_surrogate_grad(_p(p), _p(X), _p(y), X.shape[0], _p(g), _p(ss)))
    return g, float(ss[0])


def train(params=None, *, seed: int = 1, n_train: int = 10_000_000, n_val: int = 200_000, batch: int = 65536,
          epochs: int = 8, lr: float = 3e-3):
    """Train on the GPU; returns (params, report dict)."""
    p = np.ascontiguousarray(init_params(seed) if params is None else params, dtype=np.float32).copy()
    cfg = _lib.SurrogateTrainConfig(seed=seed, n_train=n_train, n_val=n_val, batch=batch, epochs=epochs, lr=lr)
    rep = np.zeros(4)
    _lib.check(_lib.load().sdfibm_surrogate_train(C.byref(cfg), _p(p), _p(rep)))
    return p, dict(train_mse=rep[0], val_mse=rep[1], val_max_err=rep[2], seconds=rep[3])


def load_params(path: str = PARAMS_PATH) -> np.ndarray:
    p = np.load(path, allow_pickle=False).astype(np.float32)
    if p.size != N_PARAMS:
        raise ValueError(f"{path}: expected {N_PARAMS} parameters")
    return p
