"""The learned warm start: device evaluation, gradients and Adam steps of the network, its file
format, and the training recipe of the dam-break workloads.

The network (16 features -> 64 -> 64 -> 4, ReLU) runs in BF16 on the matrix cores with FP32
accumulation; its parameters are one float32 blob of ``MLP_PARAMS`` values (layout in
include/akua.h).  Every function here calls the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Sequence, Tuple

import numpy as np

from . import _lib

MLP_PARAMS = _lib.MLP_PARAMS
MLP_TRAINABLE = _lib.MLP_TRAINABLE
HERE = os.path.dirname(os.path.abspath(__file__))
NETWORK_DIR = os.path.join(HERE, "networks")


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _blob(params) -> np.ndarray:
    return np.ascontiguousarray(params, np.float32).reshape(MLP_PARAMS)


def mlp_forward(params, features) -> np.ndarray:
    L = _lib.load()
    F = np.ascontiguousarray(features, np.float32).reshape(-1, 16)
    y = np.zeros((len(F), 4), np.float32)
    _lib.check(L.akua_mlp_forward(_p(_blob(params)), _p(F), len(F), _p(y)))
    return y


def mlp_gradient(params, features, targets, weights=(1.0, 1.0, 1.0, 1.0)) -> Tuple[float, np.ndarray]:
    L = _lib.load()
    F = np.ascontiguousarray(features, np.float32).reshape(-1, 16)
    T = np.ascontiguousarray(targets, np.float32).reshape(-1, 4)
    w = np.ascontiguousarray(weights, np.float32).reshape(4)
    g = np.zeros(MLP_TRAINABLE, np.float32)
    loss = np.zeros(1, np.float32)
    _lib.check(L.akua_mlp_gradient(_p(_blob(params)), _p(F), _p(T), len(F), _p(w), _p(g), _p(loss)))
    return float(loss[0]), g


def mlp_adam_step(params, m, v, t: int, lr: float, features, targets, weights=(1.0, 1.0, 1.0, 1.0)):
    """One Adam step on the device; returns (params, m, v, loss) as new arrays."""
    L = _lib.load()
    p = _blob(params).copy()
    m = np.ascontiguousarray(m, np.float32).reshape(MLP_TRAINABLE).copy()
    v = np.ascontiguousarray(v, np.float32).reshape(MLP_TRAINABLE).copy()
    F = np.ascontiguousarray(features, np.float32).reshape(-1, 16)
    T = np.ascontiguousarray(targets, np.float32).reshape(-1, 4)
    w = np.ascontiguousarray(weights, np.float32).reshape(4)
    loss = np.zeros(1, np.float32)
    _lib.check(L.akua_mlp_adam_step(_p(p), _p(m), _p(v), int(t), float(lr), _p(F), _p(T), len(F),
                                    _p(w), _p(loss)))
    return p, m, v, float(loss[0])


def save_network(path: str, params, meta: dict) -> None:
    """Parameters as float32 .npy plus a small JSON file of how they were made."""
    import json
    np.save(path, _blob(params))
    with open(os.path.splitext(path)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


def load_network(path: str) -> np.ndarray:
    p = np.load(path, allow_pickle=False)
    if p.shape != (MLP_PARAMS,) or p.dtype != np.float32:
        raise ValueError(f"{path}: not a warm-start network blob")
    return p
