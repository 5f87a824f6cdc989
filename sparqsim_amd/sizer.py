"""The learned output sizer: Python side of the engine's MLP (sparq_mlp_*,
sparq_sizer_* in include/sparq_hip.h).

The network (16 features -> 64 -> 64 -> 64 -> 1, ReLU, FP32) predicts log2(G)/32
for the group count G of Hadamard_Int, Hadamard_Int_Full and dense group
unitaries (Rot_GeneralUnitary / Rot_GeneralStatePrep).  With the sizer enabled
those operators allocate for 2^(32*y + margin) groups and do not wait for the
exact count; the exact count, computed on the device anyway, is read back with
the final size, and an under-prediction reruns the exact path, so states are
identical with and without it.

Workflow: `collect()` runs workloads with logging on and returns the
(features, targets) the engine recorded; `train()` fits the network on the GPU
with the engine's own forward/backward/Adam kernels; `fit_margin()` picks the
safety margin from the residuals; `save()` / `load_file()` / `activate()`.
"""
from __future__ import annotations

import ctypes
import os
from contextlib import contextmanager

import numpy as np

FEATURES = 16
HIDDEN = 64
PARAMS = HIDDEN * FEATURES + HIDDEN + 2 * (HIDDEN * HIDDEN + HIDDEN) + HIDDEN + 1
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sizer_weights.npz")

_lib = None


def _engine():
    global _lib
    if _lib is None:
        from . import ENGINE_LIBRARY

        lib = ctypes.CDLL(ENGINE_LIBRARY)
        f = ctypes.POINTER(ctypes.c_float)
        lib.sparq_mlp_forward.argtypes = [f, f, ctypes.c_uint32, f]
        lib.sparq_mlp_grad.argtypes = [f, f, f, ctypes.c_uint32, ctypes.c_float, f, f]
        lib.sparq_mlp_train.argtypes = [f, f, f, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_uint64, f]
        lib.sparq_sizer_load.argtypes = [f, ctypes.c_uint32, ctypes.c_float]
        lib.sparq_sizer_enable.argtypes = [ctypes.c_int]
        lib.sparq_sizer_log.argtypes = [ctypes.c_int]
        lib.sparq_sizer_log_fetch.argtypes = [f, f, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int]
        lib.sparq_sizer_stats.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        lib.sparq_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        msg = _engine().sparq_last_error().decode()
        raise (ValueError if rc == 3 else RuntimeError)(f"sizer call failed ({rc}): {msg}")


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised parameters (zero biases), packed in the engine's layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((FEATURES, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, 1)):
        parts.append(rng.standard_normal(fan_in * fan_out) * np.sqrt(2.0 / fan_in))
        parts.append(np.zeros(fan_out))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == PARAMS
    return p


def forward(params: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Network output on the GPU for x[n, 16]."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FEATURES)
    y = np.zeros(len(x), dtype=np.float32)
    _check(_engine().sparq_mlp_forward(_fp(p), _fp(x), len(x), _fp(y)))
    return y


def grad(params: np.ndarray, x: np.ndarray, y: np.ndarray, alpha: float = 4.0):
    """(loss, gradient) of the asymmetric loss over all samples, on the GPU."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FEATURES)
    y = np.ascontiguousarray(y, dtype=np.float32)
    g = np.zeros(PARAMS, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    _check(_engine().sparq_mlp_grad(_fp(p), _fp(x), _fp(y), len(y), alpha, _fp(g), _fp(loss)))
    return float(loss[0]), g


def train(params: np.ndarray, x: np.ndarray, y: np.ndarray, steps: int = 2000, batch: int = 256, lr: float = 1e-3,
          alpha: float = 4.0, seed: int = 0):
    """Adam on the GPU; returns (new parameters, per-step minibatch loss)."""
    p = np.array(params, dtype=np.float32, copy=True)
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FEATURES)
    y = np.ascontiguousarray(y, dtype=np.float32)
    hist = np.zeros(steps, dtype=np.float32)
    _check(_engine().sparq_mlp_train(_fp(p), _fp(x), _fp(y), len(y), steps, batch, lr, alpha, seed, _fp(hist)))
    return p, hist


def fit_margin(params: np.ndarray, x: np.ndarray, y: np.ndarray, quantile: float = 0.995, slack: float = 0.05) -> float:
    """Margin (log2 units) that covers the given quantile of under-predictions."""
    under = (np.asarray(y, dtype=np.float64) - forward(params, x)) * 32.0
    return float(max(0.0, np.quantile(under, quantile)) + slack)


def activate(params: np.ndarray, margin: float, on: bool = True) -> None:
    p = np.ascontiguousarray(params, dtype=np.float32)
    _check(_engine().sparq_sizer_load(_fp(p), p.size, float(margin)))
    _check(_engine().sparq_sizer_enable(1 if on else 0))


def enable(on: bool) -> None:
    _check(_engine().sparq_sizer_enable(1 if on else 0))


def stats() -> dict:
    """Operators sized by the exact count, by the network, and network misses (reruns)."""
    out = (ctypes.c_uint64 * 3)()
    _check(_engine().sparq_sizer_stats(out))
    return {"exact": int(out[0]), "predicted": int(out[1]), "reruns": int(out[2])}


@contextmanager
def logging():
    """Record (features, log2(G)/32) of every sized operator run inside the block."""
    _check(_engine().sparq_sizer_log(1))
    try:
        yield
    finally:
        _check(_engine().sparq_sizer_log(0))


def fetch_logs(clear: bool = True):
    n = ctypes.c_uint32(0)
    f = ctypes.POINTER(ctypes.c_float)
    _check(_engine().sparq_sizer_log_fetch(ctypes.cast(None, f), ctypes.cast(None, f), 0, ctypes.byref(n), 0))
    x = np.zeros((n.value, FEATURES), dtype=np.float32)
    y = np.zeros(n.value, dtype=np.float32)
    _check(_engine().sparq_sizer_log_fetch(_fp(x), _fp(y), n.value, ctypes.byref(n), 1 if clear else 0))
    return x, y


def save(path: str, params: np.ndarray, margin: float, **meta) -> None:
    np.savez(path, params=np.asarray(params, dtype=np.float32), margin=np.float64(margin),
             **{k: np.asarray(v) for k, v in meta.items()})


def load_file(path: str = DEFAULT_WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), float(z["margin"])


def activate_default() -> bool:
    """Load and enable the committed network if present; returns whether it was."""
    if not os.path.exists(DEFAULT_WEIGHTS):
        return False
    p, m = load_file(DEFAULT_WEIGHTS)
    activate(p, m)
    return True
