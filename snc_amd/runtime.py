"""Runtime layer mirror (include/snc/runtime.hpp) for the hip-* backends.

``forward(g, events, num_steps, lif, backend)`` keeps the reference signature
and result shape (ForwardResult: logits + RunStats + used_backend). Only the
hip-* backends live here; the reference's own ``forward`` keeps cpu/openmp and
its "unavailable -> cpu" fallback (cpu_backend.cpp:195-207). This module never
falls back: a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from enum import IntEnum

import numpy as np

from . import _lib as L
from .graph import SNNGraph
from .inputs import event_ptr


class Backend(IntEnum):
    """snc::Backend (runtime.hpp:21-27) extended with the hip-* strategies."""
    Cpu = 0
    OpenMP = 1
    CudaAtomic = 2
    CudaBucket = 3
    CudaSort = 4
    HipAtomic = L.VARIANT_ATOMIC
    HipBucket = L.VARIANT_BUCKET
    HipSort = L.VARIANT_SORT
    HipAuto = L.VARIANT_AUTO


_NAMES = {
    Backend.Cpu: "cpu", Backend.OpenMP: "openmp", Backend.CudaAtomic: "cuda-atomic",
    Backend.CudaBucket: "cuda-bucket", Backend.CudaSort: "cuda-sort",
    Backend.HipAtomic: "hip-atomic", Backend.HipBucket: "hip-bucket",
    Backend.HipSort: "hip-sort", Backend.HipAuto: "hip-auto",
}


def parse_backend(name: str) -> Backend:
    for b, s in _NAMES.items():
        if s == name:
            return b
    raise ValueError(f"unknown backend: {name}")


def backend_name(b: Backend) -> str:
    return _NAMES.get(Backend(b), "?")


def backend_is_hip(b: Backend) -> bool:
    return Backend(b) in (Backend.HipAtomic, Backend.HipBucket, Backend.HipSort, Backend.HipAuto)


@dataclass
class LIFParams:
    """LIFParams (runtime.hpp:34-39) plus the reset mode."""
    threshold: float = 1.0
    decay: float = 0.9
    reset: float = 0.0
    refractory: int = 1
    reset_mode: int = L.RESET_TO_VALUE

    def c(self) -> L.Lif:
        return L.Lif(self.threshold, self.decay, self.reset, self.refractory, self.reset_mode)


@dataclass
class RunStats:
    spikes: int = 0
    synaptic_events: int = 0
    overflow: int = 0
    ms: float = 0.0


@dataclass
class ForwardResult:
    logits: np.ndarray
    stats: RunStats
    used_backend: Backend
    neuron_spikes: np.ndarray | None = None
    raster: np.ndarray | None = None       # uint32 [T][ceil(N/32)]
    final_v: np.ndarray | None = None
    step_spikes: np.ndarray | None = None
    extra: dict = field(default_factory=dict)


def synchronize():
    L.check(L.lib().snc_hip_synchronize())


def available() -> bool:
    """hip::available(): library present and a gfx950 device usable."""
    try:
        return bool(L.lib().snc_hip_available())
    except L.HipBackendMissing:
        return False


class HipGraph:
    """A graph uploaded once to HBM (snc_hip_graph_upload)."""

    def __init__(self, g: SNNGraph):
        self.graph = g
        self._h = C.c_void_p()
        L.check(L.lib().snc_hip_graph_upload(C.byref(g.view()), C.byref(self._h)))

    def profile(self) -> dict:
        p = L.GraphProfile()
        L.check(L.lib().snc_hip_graph_profile_get(self._h, C.byref(p)))
        return {k: getattr(p, k) for k, _ in p._fields_}

    def close(self):
        if self._h:
            L.lib().snc_hip_graph_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ForwardPlan:
    """Graph + one sample's events + device state, reusable across runs."""

    def __init__(self, hg: HipGraph, events, num_steps: int, lif: LIFParams, backend: Backend):
        if not backend_is_hip(backend):
            raise ValueError(f"{backend_name(backend)} is not a hip backend")
        self.hg = hg
        self.num_steps = int(num_steps)
        ev, p = event_ptr(events)
        self._events = ev
        self._h = C.c_void_p()
        lc = lif.c()
        L.check(L.lib().snc_hip_plan_create(hg._h, p, ev.size, self.num_steps, C.byref(lc),
                                            int(backend), C.byref(self._h)))

    @property
    def variant(self) -> Backend:
        return Backend(L.lib().snc_hip_plan_variant(self._h))

    def run(self, capture_raster=False, capture_v=False) -> ForwardResult:
        g = self.hg.graph
        n = g.num_neurons
        logits = np.zeros(max(g.num_output_channels, 1), dtype=np.float32)
        nspk = np.zeros(n, dtype=np.int32)
        words = (n + 31) // 32
        raster = np.zeros((self.num_steps, words), dtype=np.uint32) if capture_raster else None
        steps = np.zeros(self.num_steps, dtype=np.int64) if capture_raster else None
        fv = np.zeros(n, dtype=np.float32) if capture_v else None
        cap = L.ForwardCapture(L.ptr(logits, C.c_float), L.ptr(nspk, C.c_int32),
                               L.ptr(raster, C.c_uint32), L.ptr(fv, C.c_float),
                               L.ptr(steps, C.c_int64))
        st = L.ForwardStats()
        L.check(L.lib().snc_hip_plan_run(self._h, C.byref(cap), C.byref(st)))
        return ForwardResult(logits[: g.num_output_channels], RunStats(st.spikes, st.synaptic_events,
                                                                       st.overflow, st.ms),
                             Backend(st.used_variant), nspk, raster, fv, steps)

    def launch(self):
        """Enqueue one run without synchronising (throughput timing)."""
        L.check(L.lib().snc_hip_plan_launch(self._h))

    def kernel_time(self):
        avg = C.c_double()
        k = C.c_int64()
        L.check(L.lib().snc_hip_plan_kernel_time(self._h, C.byref(avg), C.byref(k)))
        return avg.value, k.value

    def profile(self) -> dict:
        """Per-kernel device time of one instrumented run: {name: (total_ms, launches)}."""
        nk = len(L.KERNEL_NAMES)
        ms = (C.c_double * nk)()
        k = (C.c_int64 * nk)()
        L.check(L.lib().snc_hip_plan_profile(self._h, ms, k))
        return {L.KERNEL_NAMES[i]: (ms[i], k[i]) for i in range(nk) if k[i]}

    def close(self):
        if self._h:
            L.lib().snc_hip_plan_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def forward(g: SNNGraph, events, num_steps: int, lif: LIFParams, backend: Backend,
            capture_raster=False, capture_v=False) -> ForwardResult:
    """hip::forward: upload, run one sample, read back. Raises when unavailable."""
    if not backend_is_hip(backend):
        raise ValueError(f"{backend_name(backend)}: the hip runtime serves hip-* backends only")
    hg = HipGraph(g)
    try:
        plan = ForwardPlan(hg, events, num_steps, lif, backend)
        try:
            return plan.run(capture_raster=capture_raster, capture_v=capture_v)
        finally:
            plan.close()
    finally:
        hg.close()
