"""Learned grow proposal: host mirror of the snc_hip_mlp_* C ABI.

The MLP (16 -> 64 -> 64 -> 1, BF16 weights, FP32 accumulation) scores
candidate synapses for the grow step of structural co-training; it is
trained on the GPU on logged grow outcomes (did the grown synapse survive
the next prune). The reference's grow heuristic remains the fallback and
the baseline it is compared with (src/structure/connectome.cpp:164-211).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

N_FEATURES = 16
N_PARAMS = 5313
NOALLOC_MAX_EXAMPLES = 1024  # SNC_MLP_NOALLOC_MAX_EXAMPLES

_SIGS = {
    "snc_hip_mlp_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "snc_hip_mlp_destroy": (None, [C.c_void_p]),
    "snc_hip_mlp_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "snc_hip_mlp_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "snc_hip_mlp_score": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64,
                                    C.POINTER(C.c_float), C.POINTER(C.c_double)]),
    "snc_hip_mlp_grad": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                   C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_double)]),
    "snc_hip_mlp_grad_noalloc": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                           C.c_int64, C.POINTER(C.c_float),
                                           C.POINTER(C.c_double)]),
    "snc_hip_mlp_train": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                    C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                                    C.POINTER(C.c_double)]),
}
L.register_signatures(_SIGS)


class GrowMLP:
    def __init__(self, seed: int = 1):
        self._h = C.c_void_p()
        L.check(L.lib().snc_hip_mlp_create(seed, C.byref(self._h)))
        L.track(self)
        self.last_kernel_ms = 0.0

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(N_PARAMS, np.float32)
        L.check(L.lib().snc_hip_mlp_get_params(self._h, L.ptr(p, C.c_float)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == N_PARAMS
        L.check(L.lib().snc_hip_mlp_set_params(self._h, L.ptr(p, C.c_float)))

    def score(self, features) -> np.ndarray:
        X = np.ascontiguousarray(features, np.float32).reshape(-1, N_FEATURES)
        out = np.zeros(X.shape[0], np.float32)
        ms = C.c_double()
        L.check(L.lib().snc_hip_mlp_score(self._h, L.ptr(X, C.c_float), X.shape[0],
                                          L.ptr(out, C.c_float), C.byref(ms)))
        self.last_kernel_ms = ms.value
        return out

    def grad(self, features, labels):
        X = np.ascontiguousarray(features, np.float32).reshape(-1, N_FEATURES)
        y = np.ascontiguousarray(labels, np.float32)
        g = np.zeros(N_PARAMS, np.float32)
        loss = C.c_double()
        L.check(L.lib().snc_hip_mlp_grad(self._h, L.ptr(X, C.c_float), L.ptr(y, C.c_float),
                                         X.shape[0], L.ptr(g, C.c_float), C.byref(loss)))
        return loss.value, g

    def grad_noalloc(self, features, labels):
        """grad() through snc_hip_mlp_grad_noalloc: no device allocation, at
        most NOALLOC_MAX_EXAMPLES examples, the GPU's one fixed scratch array."""
        X = np.ascontiguousarray(features, np.float32).reshape(-1, N_FEATURES)
        y = np.ascontiguousarray(labels, np.float32)
        g = np.zeros(N_PARAMS, np.float32)
        loss = C.c_double()
        L.check(L.lib().snc_hip_mlp_grad_noalloc(self._h, L.ptr(X, C.c_float),
                                                 L.ptr(y, C.c_float), X.shape[0],
                                                 L.ptr(g, C.c_float), C.byref(loss)))
        return loss.value, g

    def train(self, features, labels, epochs=20, batch=256, lr=1e-3, seed=1) -> float:
        X = np.ascontiguousarray(features, np.float32).reshape(-1, N_FEATURES)
        y = np.ascontiguousarray(labels, np.float32)
        loss = C.c_double()
        L.check(L.lib().snc_hip_mlp_train(self._h, L.ptr(X, C.c_float), L.ptr(y, C.c_float),
                                          X.shape[0], epochs, batch, lr, seed, C.byref(loss)))
        return loss.value

    def close(self):
        if self._h:
            L.lib().snc_hip_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
