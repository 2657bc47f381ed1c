"""ctypes binding of libsnc_hip.so (the C ABI in include/snc_hip.h, snc_gen.h).

The library is loaded from this package directory only. There is no
fallback: if the shared object is missing, importing a function that needs it
raises ``HipBackendMissing``.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsnc_hip.so")

SNC_OK = 0
SNC_E_INVALID = -1
SNC_E_NODEVICE = -2
SNC_E_HIP = -3
SNC_E_NOMEM = -4
SNC_E_UNSUPPORTED = -5

VARIANT_ATOMIC = 5
VARIANT_BUCKET = 6
VARIANT_SORT = 7
VARIANT_AUTO = 8
VARIANT_AUTO_EXACT = 9
VARIANT_PULL = 10

ROLE_INTERNAL, ROLE_INPUT, ROLE_OUTPUT = 0, 1, 2
KERNEL_NAMES = ["k_hub_step", "k_step", "sort_step", "k_persist"]
RESET_TO_VALUE, RESET_SUBTRACT = 0, 1


class HipBackendMissing(RuntimeError):
    """libsnc_hip.so is not built (run __graft_entry__.build())."""


class HipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"snc_hip error {code}: {msg}")
        self.code = code


class GraphView(C.Structure):
    _fields_ = [
        ("num_neurons", C.c_int32),
        ("num_synapses", C.c_int32),
        ("row_ptr", C.POINTER(C.c_int32)),
        ("post_ids", C.POINTER(C.c_int32)),
        ("weights", C.POINTER(C.c_float)),
        ("delays", C.POINTER(C.c_int32)),
        ("sign", C.POINTER(C.c_int8)),
        ("role", C.POINTER(C.c_uint8)),
        ("channel", C.POINTER(C.c_int32)),
    ]


class Lif(C.Structure):
    _fields_ = [
        ("threshold", C.c_float),
        ("decay", C.c_float),
        ("reset", C.c_float),
        ("refractory", C.c_int32),
        ("reset_mode", C.c_int32),
    ]


class Event(C.Structure):
    _fields_ = [
        ("sample_id", C.c_int32),
        ("channel_id", C.c_int32),
        ("time", C.c_int32),
        ("amplitude", C.c_float),
    ]


class ForwardStats(C.Structure):
    _fields_ = [
        ("spikes", C.c_int64),
        ("synaptic_events", C.c_int64),
        ("overflow", C.c_int32),
        ("used_variant", C.c_int32),
        ("ms", C.c_double),
    ]


class ForwardCapture(C.Structure):
    _fields_ = [
        ("logits", C.POINTER(C.c_float)),
        ("neuron_spikes", C.POINTER(C.c_int32)),
        ("raster", C.POINTER(C.c_uint32)),
        ("final_v", C.POINTER(C.c_float)),
        ("step_spikes", C.POINTER(C.c_int64)),
    ]


class GraphProfile(C.Structure):
    _fields_ = [
        ("num_neurons", C.c_int32),
        ("num_synapses", C.c_int64),
        ("max_delay", C.c_int32),
        ("ring_len", C.c_int32),
        ("max_in_degree", C.c_int32),
        ("max_out_degree", C.c_int32),
        ("mean_in_degree", C.c_double),
        ("heavy_fraction", C.c_double),
        ("heavy_threshold", C.c_int32),
        ("num_heavy", C.c_int32),
    ]


class SheetParams(C.Structure):
    _fields_ = [
        ("n_layers", C.c_int32),
        ("layers", C.POINTER(C.c_int32)),
        ("fan_in", C.c_int32),
        ("recurrent_fan_in", C.c_int32),
        ("dense_readout", C.c_int32),
        ("radius", C.c_float),
        ("sheet", C.c_float),
        ("w_std", C.c_float),
        ("delay", C.c_int32),
        ("max_delay", C.c_int32),
        ("seed", C.c_uint64),
    ]


class HubParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("out_degree", C.c_int32),
        ("radius", C.c_float),
        ("hub_fraction", C.c_float),
        ("hub_alpha", C.c_float),
        ("hub_min", C.c_int32),
        ("hub_max", C.c_int32),
        ("w_std", C.c_float),
        ("max_delay", C.c_int32),
        ("all_input", C.c_int32),
        ("seed", C.c_uint64),
    ]


class PlanConfig(C.Structure):
    _fields_ = [("variant", C.c_int32), ("block_neurons", C.c_int32),
                ("heavy_threshold", C.c_int32), ("persistent", C.c_int32)]


TUNE_MAX_CANDIDATES = 24
TUNE_ERROR_LEN = 120


class TuneReport(C.Structure):
    _fields_ = [("n_candidates", C.c_int32), ("chosen", C.c_int32), ("trial_steps", C.c_int32),
                ("search_ms", C.c_double), ("cand", PlanConfig * TUNE_MAX_CANDIDATES),
                ("ms_per_step", C.c_double * TUNE_MAX_CANDIDATES),
                ("status", C.c_int32 * TUNE_MAX_CANDIDATES),
                ("error", (C.c_char * TUNE_ERROR_LEN) * TUNE_MAX_CANDIDATES)]


_P = C.c_void_p
_SIGS = {
    "snc_hip_last_error": (C.c_char_p, []),
    "snc_hip_available": (C.c_int, []),
    "snc_hip_graph_upload": (C.c_int, [C.POINTER(GraphView), C.POINTER(_P)]),
    "snc_hip_graph_free": (None, [_P]),
    "snc_hip_graph_profile_get": (C.c_int, [_P, C.POINTER(GraphProfile)]),
    "snc_hip_plan_create": (C.c_int, [_P, C.POINTER(Event), C.c_int64, C.c_int32, C.POINTER(Lif),
                                      C.c_int32, C.POINTER(_P)]),
    "snc_hip_plan_create_config": (C.c_int, [_P, C.POINTER(Event), C.c_int64, C.c_int32,
                                             C.POINTER(Lif), C.POINTER(PlanConfig),
                                             C.POINTER(_P)]),
    "snc_hip_plan_get_config": (C.c_int, [_P, C.POINTER(PlanConfig)]),
    "snc_hip_graph_tune_report": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(TuneReport)]),
    "snc_hip_plan_create_batch": (C.c_int, [_P, C.POINTER(Event), C.c_int64, C.c_int32, C.c_int32,
                                            C.POINTER(Lif), C.c_int32, C.POINTER(PlanConfig),
                                            C.POINTER(_P)]),
    "snc_hip_plan_num_samples": (C.c_int32, [_P]),
    "snc_hip_plan_sample_stats": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "snc_hip_plan_run": (C.c_int, [_P, C.POINTER(ForwardCapture), C.POINTER(ForwardStats)]),
    "snc_hip_plan_launch": (C.c_int, [_P]),
    "snc_hip_plan_variant": (C.c_int32, [_P]),
    "snc_hip_plan_kernel_time": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "snc_hip_plan_free": (None, [_P]),
    "snc_hip_plan_profile": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "snc_hip_synchronize": (C.c_int, []),
    "snc_hip_forward": (C.c_int, [C.POINTER(GraphView), C.POINTER(Event), C.c_int64, C.c_int32,
                                  C.POINTER(Lif), C.c_int32, C.POINTER(ForwardCapture),
                                  C.POINTER(ForwardStats)]),
    # generators (snc_gen.h)
    "snc_graph_buf_view": (C.c_int, [_P, C.POINTER(GraphView)]),
    "snc_graph_buf_free": (None, [_P]),
    "snc_graph_buf_from_view": (C.c_int, [C.POINTER(GraphView), C.POINTER(_P)]),
    "snc_event_buf_data": (C.c_int, [_P, C.POINTER(C.POINTER(Event)), C.POINTER(C.c_int64)]),
    "snc_event_buf_free": (None, [_P]),
    "snc_graph_validate": (C.c_int, [C.POINTER(GraphView)]),
    "snc_gen_sheet": (C.c_int, [C.POINTER(SheetParams), C.POINTER(_P)]),
    "snc_gen_hub_sheet": (C.c_int, [C.POINTER(HubParams), C.POINTER(_P)]),
    "snc_gen_random_recurrent": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                           C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                                           C.POINTER(_P)]),
    "snc_gen_static_snc": (C.c_int, [C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_float,
                                     C.c_int32, C.c_float, C.c_uint64, C.POINTER(_P)]),
    "snc_gen_random_sparse": (C.c_int, [C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_float,
                                        C.c_int32, C.c_float, C.c_uint64, C.POINTER(_P)]),
    "snc_gen_dense": (C.c_int, [C.POINTER(C.c_int32), C.c_int32, C.c_float, C.c_int32,
                                C.POINTER(_P)]),
    "snc_gen_background": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_float, C.c_uint64,
                                     C.POINTER(_P)]),
    "snc_gen_images": (None, [C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                              C.POINTER(C.c_uint8)]),
    "snc_gen_prototype_images": (None, [C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                        C.c_uint64, C.POINTER(C.c_uint8), C.POINTER(C.c_int32)]),
    "snc_gen_event_streams": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                        C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                        C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                        C.POINTER(_P)]),
    "snc_encode": (C.c_int, [C.c_int32, C.POINTER(C.c_float), C.c_int32, C.c_int32, C.c_float,
                             C.c_float, C.c_float, C.c_uint64, C.c_int32, C.POINTER(_P)]),
    "snc_make_synthetic": (None, [C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                                  C.POINTER(C.c_float), C.POINTER(C.c_int32)]),
    "snc_hip_graph_memory": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "snc_hip_plan_memory": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "snc_hip_measure_expansion": (C.c_int, [_P, C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                            C.POINTER(C.c_double)]),
    "snc_hip_measure_scatter": (C.c_int, [C.c_int32, C.c_int64, C.c_int64, C.c_int32, C.c_int32,
                                          C.POINTER(C.c_double)]),
    # SNCG files (snc_gen.h)
    "snc_graph_io_channels": (C.c_int, [C.POINTER(GraphView), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32)]),
    "snc_graph_write_sncg": (C.c_int, [C.POINTER(GraphView), C.c_char_p]),
    "snc_graph_read_sncg": (C.c_int, [C.c_char_p, C.c_float, C.POINTER(_P),
                                      C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "snc_graph_spread_delays": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                          C.c_int32, C.c_uint64]),
}

_lib = None

# Objects holding library handles (graphs, plans, sessions, models). They are
# closed at interpreter exit, before the HIP runtime's own teardown, instead
# of by __del__ during module teardown, when the runtime may already be gone.
_LIVE = weakref.WeakSet()


def track(obj):
    """Register an object with a close() method for closing at exit."""
    _LIVE.add(obj)
    return obj


@atexit.register
def _close_live():
    # plans, sessions and models first, graphs (which plans use) last
    for obj in sorted(list(_LIVE), key=lambda o: bool(getattr(o, "_close_last", False))):
        try:
            obj.close()
        except Exception:
            pass


def lib():
    """The loaded library; raises HipBackendMissing if it was never built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HipBackendMissing(f"{LIB_PATH} not found: build the HIP backend first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def register_signatures(sigs: dict):
    """Add entry points bound by another module (binds now if already loaded)."""
    _SIGS.update(sigs)
    if _lib is not None:
        for name, (res, args) in sigs.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args


def exported_symbols():
    return list(_SIGS)


def check(rc: int):
    if rc != SNC_OK:
        msg = lib().snc_hip_last_error().decode(errors="replace")
        raise HipError(rc, msg)


def ptr(arr, ctype):
    """ctypes pointer to a contiguous numpy array (or None)."""
    if arr is None:
        return None
    return arr.ctypes.data_as(C.POINTER(ctype))
