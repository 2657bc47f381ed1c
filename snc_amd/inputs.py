"""Input side of the spiking path: InputEvent streams and their generators.

EVENT_DTYPE has the memory layout of InputEvent (include/snc/encoders.hpp:21-26)
so event arrays pass to the C ABI without copying.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

EVENT_DTYPE = np.dtype([("sample_id", "<i4"), ("channel_id", "<i4"), ("time", "<i4"),
                        ("amplitude", "<f4")])

ENCODERS = {"direct": 0, "poisson": 1, "latency": 2}


def _events_from_buf(buf) -> np.ndarray:
    lib = L.lib()
    p = C.POINTER(L.Event)()
    n = C.c_int64()
    L.check(lib.snc_event_buf_data(buf, C.byref(p), C.byref(n)))
    if n.value:
        raw = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n.value * 16,))
        ev = raw.view(EVENT_DTYPE).copy()
    else:
        ev = np.zeros(0, dtype=EVENT_DTYPE)
    lib.snc_event_buf_free(buf)
    return ev


def event_ptr(ev: np.ndarray):
    ev = np.ascontiguousarray(ev, dtype=EVENT_DTYPE)
    return ev, (ev.ctypes.data_as(C.POINTER(L.Event)) if ev.size else None)


def encode(encoder, features, num_steps, *, gain=1.0, max_rate=0.5, threshold=0.0, seed=0,
           sample_id=0) -> np.ndarray:
    """encode() of the reference (encoders.cpp:25-71), same mt19937_64 draws."""
    enc = ENCODERS[encoder] if isinstance(encoder, str) else int(encoder)
    x = np.ascontiguousarray(features, dtype=np.float32)
    out = C.c_void_p()
    L.check(L.lib().snc_encode(enc, L.ptr(x, C.c_float), x.size, num_steps, gain, max_rate,
                               threshold, seed, sample_id, C.byref(out)))
    return _events_from_buf(out)


def background_drive(n, steps, p, amplitude=1.0, seed=1) -> np.ndarray:
    """Philox Bernoulli drive on every neuron (channel = neuron id)."""
    out = C.c_void_p()
    L.check(L.lib().snc_gen_background(n, steps, p, amplitude, seed, C.byref(out)))
    return _events_from_buf(out)


def images(n, pixels=784, zero_frac=0.7, seed=1) -> np.ndarray:
    out = np.zeros((n, pixels), dtype=np.uint8)
    L.lib().snc_gen_images(n, pixels, zero_frac, seed, L.ptr(out, C.c_uint8))
    return out


def prototype_images(n, pixels=784, classes=10, zero_frac=0.7, noise=0.2, seed=1):
    X = np.zeros((n, pixels), dtype=np.uint8)
    y = np.zeros(n, dtype=np.int32)
    L.lib().snc_gen_prototype_images(n, pixels, classes, zero_frac, noise, seed,
                                     L.ptr(X, C.c_uint8), L.ptr(y, C.c_int32))
    return X, y


def event_streams(n, channels=700, classes=20, n_centres=20, sd=15.0, events_per_sample=5000,
                  t_min=80, t_max=120, seed=1):
    lengths = np.zeros(n, dtype=np.int32)
    labels = np.zeros(n, dtype=np.int32)
    out = C.c_void_p()
    L.check(L.lib().snc_gen_event_streams(n, channels, classes, n_centres, sd, events_per_sample,
                                          t_min, t_max, seed, L.ptr(lengths, C.c_int32),
                                          L.ptr(labels, C.c_int32), C.byref(out)))
    return _events_from_buf(out), lengths, labels


def make_synthetic(n, dim, classes, noise, seed):
    """make_synthetic (datasets.cpp:11-30)."""
    X = np.zeros((n, dim), dtype=np.float32)
    y = np.zeros(n, dtype=np.int32)
    L.lib().snc_make_synthetic(n, dim, classes, noise, seed, L.ptr(X, C.c_float),
                               L.ptr(y, C.c_int32))
    return X, y
