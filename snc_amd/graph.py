"""SNNGraph (include/snc/snn_graph.hpp:43-77) as numpy arrays, plus generators.

The CSR is over OUTGOING synapses with rows sorted by pre id; per-neuron
role / sign / channel are parallel arrays. Generators are the C++ ones in
libsnc_hip.so (include/snc_gen.h), so every host regenerates identical arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


@dataclass
class SNNGraph:
    row_ptr: np.ndarray   # int32 [N+1]
    post_ids: np.ndarray  # int32 [S]
    weights: np.ndarray   # float32 [S]
    delays: np.ndarray    # int32 [S]
    sign: np.ndarray      # int8 [N]
    role: np.ndarray      # uint8 [N] (0 internal, 1 input, 2 output)
    channel: np.ndarray   # int32 [N]

    def __post_init__(self):
        self.row_ptr = np.ascontiguousarray(self.row_ptr, dtype=np.int32)
        self.post_ids = np.ascontiguousarray(self.post_ids, dtype=np.int32)
        self.weights = np.ascontiguousarray(self.weights, dtype=np.float32)
        self.delays = np.ascontiguousarray(self.delays, dtype=np.int32)
        self.sign = np.ascontiguousarray(self.sign, dtype=np.int8)
        self.role = np.ascontiguousarray(self.role, dtype=np.uint8)
        self.channel = np.ascontiguousarray(self.channel, dtype=np.int32)

    @property
    def num_neurons(self) -> int:
        return int(self.row_ptr.shape[0] - 1)

    @property
    def num_synapses(self) -> int:
        return int(self.post_ids.shape[0])

    @property
    def num_input_channels(self) -> int:
        m = self.channel[self.role == L.ROLE_INPUT]
        return int(m.max() + 1) if m.size else 0

    @property
    def num_output_channels(self) -> int:
        m = self.channel[self.role == L.ROLE_OUTPUT]
        return int(m.max() + 1) if m.size else 0

    @property
    def max_delay(self) -> int:
        return int(self.delays.max()) if self.delays.size else 1

    def view(self) -> L.GraphView:
        v = L.GraphView()
        v.num_neurons = self.num_neurons
        v.num_synapses = self.num_synapses
        v.row_ptr = L.ptr(self.row_ptr, C.c_int32)
        v.post_ids = L.ptr(self.post_ids, C.c_int32)
        v.weights = L.ptr(self.weights, C.c_float)
        v.delays = L.ptr(self.delays, C.c_int32)
        v.sign = L.ptr(self.sign, C.c_int8)
        v.role = L.ptr(self.role, C.c_uint8)
        v.channel = L.ptr(self.channel, C.c_int32)
        return v

    def validate(self) -> tuple[bool, str]:
        """SNNGraph::validate (snn_graph.cpp:30-56)."""
        lib = L.lib()
        ok = lib.snc_graph_validate(C.byref(self.view()))
        return bool(ok), ("" if ok else lib.snc_hip_last_error().decode())

    def in_degree(self) -> np.ndarray:
        return np.bincount(self.post_ids, minlength=self.num_neurons).astype(np.int64)

    def pre_ids(self) -> np.ndarray:
        return np.repeat(np.arange(self.num_neurons, dtype=np.int32), np.diff(self.row_ptr))


DELAYS_RANDOM, DELAYS_DISTANCE = 0, 1
SNCG_MAGIC = 0x534E4347


def write_sncg(g: SNNGraph, path) -> None:
    """Write g as an SNCG file, the format of the reference's snc_export
    (src/bench/snc_export.cpp:107-123): topology, delays, roles and channels;
    no weights."""
    L.check(L.lib().snc_graph_write_sncg(C.byref(g.view()), str(path).encode()))


def read_sncg(path, weight: float = 0.5) -> tuple[SNNGraph, int, int]:
    """load_graph (python/snc/graph.py:11-24) into an SNNGraph: every weight
    is `weight` and every sign +1, since the file holds neither. Returns
    (graph, num_input_channels, num_output_channels) from the header."""
    out, n_in, n_out = C.c_void_p(), C.c_int32(), C.c_int32()
    L.check(L.lib().snc_graph_read_sncg(str(path).encode(), weight, C.byref(out),
                                        C.byref(n_in), C.byref(n_out)))
    return _from_buf(out), int(n_in.value), int(n_out.value)


def spread_delays(g: SNNGraph, mode: int, delay_max: int, layers=(), seed: int = 1) -> SNNGraph:
    """snc_export's heterogeneous delays (snc_export.cpp:81-101) on a copy of g."""
    lib = L.lib()
    buf = C.c_void_p()
    L.check(lib.snc_graph_buf_from_view(C.byref(g.view()), C.byref(buf)))
    lay = np.ascontiguousarray(layers, dtype=np.int32)
    rc = lib.snc_graph_spread_delays(buf, mode, delay_max, L.ptr(lay, C.c_int32), lay.size, seed)
    if rc:
        lib.snc_graph_buf_free(buf)
        L.check(rc)
    return _from_buf(buf)


def _from_buf(buf) -> SNNGraph:
    lib = L.lib()
    v = L.GraphView()
    L.check(lib.snc_graph_buf_view(buf, C.byref(v)))
    n, m = v.num_neurons, v.num_synapses

    def arr(p, count, dt):
        if count == 0:
            return np.zeros(0, dtype=dt)
        return np.ctypeslib.as_array(p, shape=(count,)).astype(dt, copy=True)

    g = SNNGraph(arr(v.row_ptr, n + 1, np.int32), arr(v.post_ids, m, np.int32),
                 arr(v.weights, m, np.float32), arr(v.delays, m, np.int32),
                 arr(v.sign, n, np.int8), arr(v.role, n, np.uint8), arr(v.channel, n, np.int32))
    lib.snc_graph_buf_free(buf)
    return g


def _call_gen(fn, *args) -> SNNGraph:
    out = C.c_void_p()
    L.check(fn(*args, C.byref(out)))
    return _from_buf(out)


def sheet_graph(layers, fan_in, *, recurrent_fan_in=0, dense_readout=False, radius=3.0,
                sheet=28.0, w_std=1.0, delay=1, max_delay=0, seed=1) -> SNNGraph:
    """Layered 2-D-sheet graph with exp(-d/r) partner draws (snc_gen_sheet)."""
    lay = np.ascontiguousarray(layers, dtype=np.int32)
    p = L.SheetParams(len(lay), L.ptr(lay, C.c_int32), fan_in, recurrent_fan_in,
                      1 if dense_readout else 0, radius, sheet, w_std, delay, max_delay, seed)
    return _call_gen(L.lib().snc_gen_sheet, C.byref(p))


def hub_sheet_graph(width, height, out_degree=100, *, radius=3.0, hub_fraction=0.01,
                    hub_alpha=2.1, hub_min=100, hub_max=50000, w_std=0.05, max_delay=32,
                    all_input=True, seed=1) -> SNNGraph:
    """Delivery stress graph with power-law hubs (snc_gen_hub_sheet)."""
    p = L.HubParams(width, height, out_degree, radius, hub_fraction, hub_alpha, hub_min, hub_max,
                    w_std, max_delay, 1 if all_input else 0, seed)
    return _call_gen(L.lib().snc_gen_hub_sheet, C.byref(p))


def random_recurrent_graph(n_in, n_hidden, n_out, density_in, density_rec, delay_min, delay_max,
                           w_std, seed=1) -> SNNGraph:
    return _call_gen(L.lib().snc_gen_random_recurrent, n_in, n_hidden, n_out, density_in,
                     density_rec, delay_min, delay_max, w_std, seed)


def static_snc(layers, budget, weight=0.5, delay=1, inhib=0.2, seed=1) -> SNNGraph:
    """make_static_snc (snn_graph.cpp:185-236)."""
    lay = np.ascontiguousarray(layers, dtype=np.int32)
    return _call_gen(L.lib().snc_gen_static_snc, L.ptr(lay, C.c_int32), len(lay), budget, weight,
                     delay, inhib, seed)


def random_sparse(layers, budget, weight=0.5, delay=1, inhib=0.2, seed=1) -> SNNGraph:
    """make_random_sparse (snn_graph.cpp:288-326)."""
    lay = np.ascontiguousarray(layers, dtype=np.int32)
    return _call_gen(L.lib().snc_gen_random_sparse, L.ptr(lay, C.c_int32), len(lay), budget,
                     weight, delay, inhib, seed)


def dense(layers, weight=0.5, delay=1) -> SNNGraph:
    """make_dense (snn_graph.cpp:267-286)."""
    lay = np.ascontiguousarray(layers, dtype=np.int32)
    return _call_gen(L.lib().snc_gen_dense, L.ptr(lay, C.c_int32), len(lay), weight, delay)
