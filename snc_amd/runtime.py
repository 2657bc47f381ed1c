"""Runtime layer mirror (include/snc/runtime.hpp) for the hip-* backends.

``forward(g, events, num_steps, lif, backend)`` keeps the reference signature
and result shape (ForwardResult: logits + RunStats + used_backend). Only the
hip-* backends live here; the reference's own ``forward`` keeps cpu/openmp and
its "unavailable -> cpu" fallback (cpu_backend.cpp:195-207). This module never
falls back: a missing library or device raises.
"""
from __future__ import annotations

import hashlib
from collections import OrderedDict

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
    HipAutoExact = L.VARIANT_AUTO_EXACT
    HipPull = L.VARIANT_PULL


_NAMES = {
    Backend.Cpu: "cpu", Backend.OpenMP: "openmp", Backend.CudaAtomic: "cuda-atomic",
    Backend.CudaBucket: "cuda-bucket", Backend.CudaSort: "cuda-sort",
    Backend.HipAtomic: "hip-atomic", Backend.HipBucket: "hip-bucket",
    Backend.HipSort: "hip-sort", Backend.HipAuto: "hip-auto",
    Backend.HipAutoExact: "hip-auto-exact", Backend.HipPull: "hip-pull",
}


def parse_backend(name: str) -> Backend:
    for b, s in _NAMES.items():
        if s == name:
            return b
    raise ValueError(f"unknown backend: {name}")


def backend_name(b: Backend) -> str:
    return _NAMES.get(Backend(b), "?")


def backend_is_hip(b: Backend) -> bool:
    return Backend(b) in (Backend.HipAtomic, Backend.HipBucket, Backend.HipSort, Backend.HipAuto,
                          Backend.HipAutoExact, Backend.HipPull)


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


SCATTER_ATOMIC_ADD, SCATTER_ATOMIC_OR, SCATTER_STORE, SCATTER_LOAD = 0, 1, 2, 3


def measure_scatter(kind: int, target_bytes: int, ops: int, blocks: int, reps: int = 200) -> float:
    """Best average microseconds per launch of `ops` scattered accesses of one
    kind (SCATTER_*) over target_bytes, from `blocks` blocks of 512 threads,
    graph-launched `reps` times: the device's floor for a step kernel that
    delivers `ops` events (snc_hip_measure_scatter)."""
    us = C.c_double()
    L.check(L.lib().snc_hip_measure_scatter(kind, int(target_bytes), int(ops), int(blocks),
                                            int(reps), C.byref(us)))
    return us.value


def available() -> bool:
    """hip::available(): library present and a gfx950 device usable."""
    try:
        return bool(L.lib().snc_hip_available())
    except L.HipBackendMissing:
        return False


@dataclass
class PlanConfig:
    """snc_plan_config: strategy and launch shape of a forward plan."""
    variant: Backend = Backend.HipBucket
    block_neurons: int = 256
    heavy_threshold: int = 512
    persistent: bool = False

    def c(self):
        return L.PlanConfig(int(self.variant), self.block_neurons, self.heavy_threshold,
                            1 if self.persistent else 0)

    @staticmethod
    def from_c(c) -> "PlanConfig":
        return PlanConfig(Backend(c.variant), c.block_neurons, c.heavy_threshold,
                          bool(c.persistent))


class HipGraph:
    """A graph uploaded once to HBM (snc_hip_graph_upload)."""

    _close_last = True  # plans built on it close before it at exit

    def __init__(self, g: SNNGraph):
        self.graph = g
        self._h = C.c_void_p()
        L.check(L.lib().snc_hip_graph_upload(C.byref(g.view()), C.byref(self._h)))
        L.track(self)

    def profile(self) -> dict:
        p = L.GraphProfile()
        L.check(L.lib().snc_hip_graph_profile_get(self._h, C.byref(p)))
        return {k: getattr(p, k) for k, _ in p._fields_}

    def memory(self) -> int:
        """Device bytes this graph holds (snc_hip_graph_memory)."""
        b = C.c_int64()
        L.check(L.lib().snc_hip_graph_memory(self._h, C.byref(b)))
        return int(b.value)

    def measure_expansion(self, fired, reps: int = 200) -> float:
        """Best average microseconds per launch of the hip-atomic step
        kernel's delivery alone for these fired rows
        (snc_hip_measure_expansion)."""
        f = np.ascontiguousarray(fired, np.int32)
        us = C.c_double()
        L.check(L.lib().snc_hip_measure_expansion(self._h, L.ptr(f, C.c_int32), f.size, int(reps),
                                                  C.byref(us)))
        return us.value

    def tune_report(self, exact_only=False, batched=False) -> dict:
        """The autotuner's search for this graph (after a hip-auto plan):
        candidates with their measured ms per step, and the chosen one."""
        r = L.TuneReport()
        L.check(L.lib().snc_hip_graph_tune_report(self._h, 1 if exact_only else 0,
                                                  1 if batched else 0, C.byref(r)))
        cands = [(PlanConfig.from_c(r.cand[i]), r.ms_per_step[i]) for i in range(r.n_candidates)]
        failed = [(PlanConfig.from_c(r.cand[i]), int(r.status[i]),
                   r.error[i].value.decode(errors="replace"))
                  for i in range(r.n_candidates) if r.status[i] != L.SNC_OK]
        return {"candidates": cands, "chosen": r.chosen, "trial_steps": r.trial_steps,
                "search_ms": r.search_ms, "failed": failed}

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
    """Graph + inputs + device state, reusable across runs.

    num_samples=None: one sample (every event belongs to it). num_samples=B:
    a batched plan of B independent samples stepped together; an event
    belongs to sample ``sample_id``, and results carry a leading sample axis.
    """

    def __init__(self, hg: HipGraph, events, num_steps: int, lif: LIFParams,
                 backend: Backend = Backend.HipAuto, config: PlanConfig | None = None,
                 num_samples: int | None = None):
        """backend: a hip-* strategy with its default launch shape, or hip-auto /
        hip-auto-exact (autotuned); config: an explicit PlanConfig instead."""
        if config is None and not backend_is_hip(backend):
            raise ValueError(f"{backend_name(backend)} is not a hip backend")
        self.hg = hg
        self.num_steps = int(num_steps)
        self.batched = num_samples is not None
        self.num_samples = int(num_samples) if self.batched else 1
        ev, p = event_ptr(events)
        self._events = ev
        self._h = C.c_void_p()
        lc = lif.c()
        if self.batched:
            cc = config.c() if config is not None else None
            L.check(L.lib().snc_hip_plan_create_batch(
                hg._h, p, ev.size, self.num_samples, self.num_steps, C.byref(lc), int(backend),
                C.byref(cc) if cc is not None else None, C.byref(self._h)))
        elif config is not None:
            cc = config.c()
            L.check(L.lib().snc_hip_plan_create_config(hg._h, p, ev.size, self.num_steps,
                                                       C.byref(lc), C.byref(cc),
                                                       C.byref(self._h)))
        else:
            L.check(L.lib().snc_hip_plan_create(hg._h, p, ev.size, self.num_steps, C.byref(lc),
                                                int(backend), C.byref(self._h)))
        L.track(self)

    @property
    def variant(self) -> Backend:
        return Backend(L.lib().snc_hip_plan_variant(self._h))

    @property
    def config(self) -> PlanConfig:
        c = L.PlanConfig()
        L.check(L.lib().snc_hip_plan_get_config(self._h, C.byref(c)))
        return PlanConfig.from_c(c)

    def run(self, capture_raster=False, capture_v=False) -> ForwardResult:
        g = self.hg.graph
        n, B, T = g.num_neurons, self.num_samples, self.num_steps
        C_out = max(g.num_output_channels, 1)
        words = (n + 31) // 32
        logits = np.zeros((B, C_out), dtype=np.float32)
        nspk = np.zeros((B, n), dtype=np.int32)
        raster = np.zeros((B, T, words), dtype=np.uint32) if capture_raster else None
        steps = np.zeros((B, T), dtype=np.int64) if capture_raster else None
        fv = np.zeros((B, n), dtype=np.float32) if capture_v else None
        cap = L.ForwardCapture(L.ptr(logits, C.c_float), L.ptr(nspk, C.c_int32),
                               L.ptr(raster, C.c_uint32), L.ptr(fv, C.c_float),
                               L.ptr(steps, C.c_int64))
        st = L.ForwardStats()
        L.check(L.lib().snc_hip_plan_run(self._h, C.byref(cap), C.byref(st)))
        logits = logits[:, : g.num_output_channels]
        if not self.batched:
            logits, nspk = logits[0], nspk[0]
            raster = raster[0] if raster is not None else None
            steps = steps[0] if steps is not None else None
            fv = fv[0] if fv is not None else None
        return ForwardResult(logits, RunStats(st.spikes, st.synaptic_events, st.overflow, st.ms),
                             Backend(st.used_variant), nspk, raster, fv, steps)

    def sample_stats(self):
        """(spikes, synaptic_events) per sample of the last run."""
        sp = np.zeros(self.num_samples, np.int64)
        ev = np.zeros(self.num_samples, np.int64)
        L.check(L.lib().snc_hip_plan_sample_stats(self._h, L.ptr(sp, C.c_int64),
                                                  L.ptr(ev, C.c_int64)))
        return sp, ev

    def memory(self) -> int:
        """Device bytes this plan holds (snc_hip_plan_memory)."""
        b = C.c_int64()
        L.check(L.lib().snc_hip_plan_memory(self._h, C.byref(b)))
        return int(b.value)

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


_GRAPH_CACHE: "OrderedDict[bytes, HipGraph]" = OrderedDict()
_GRAPH_CACHE_SIZE = 4


def _graph_key(g: SNNGraph) -> bytes:
    h = hashlib.blake2b(digest_size=20)
    for a in (g.row_ptr, g.post_ids, g.weights, g.delays, g.sign, g.role, g.channel):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.digest()


def resident_graph(g: SNNGraph) -> HipGraph:
    """The device copy of g, uploaded on first use and kept (with its
    autotuner results) for the next calls on a graph with the same contents."""
    key = _graph_key(g)
    hg = _GRAPH_CACHE.get(key)
    if hg is not None:
        _GRAPH_CACHE.move_to_end(key)
        return hg
    hg = HipGraph(g)
    _GRAPH_CACHE[key] = hg
    while len(_GRAPH_CACHE) > _GRAPH_CACHE_SIZE:
        _GRAPH_CACHE.popitem(last=False)[1].close()
    return hg


def forward_oneshot(g: SNNGraph, events, num_steps: int, lif: LIFParams, backend: Backend,
                    capture_raster=False, capture_v=False) -> ForwardResult:
    """cuda::forward's shape (cuda_backend.cu:186-361) through snc_hip_forward:
    upload the graph, build a plan, run, read back and free everything in one
    call. Nothing stays resident, so hip-auto searches on every call."""
    if not backend_is_hip(backend):
        raise ValueError(f"{backend_name(backend)}: the hip runtime serves hip-* backends only")
    n, T = g.num_neurons, int(num_steps)
    c_out = max(g.num_output_channels, 1)
    words = (n + 31) // 32
    logits = np.zeros(c_out, dtype=np.float32)
    nspk = np.zeros(n, dtype=np.int32)
    raster = np.zeros((T, words), dtype=np.uint32) if capture_raster else None
    steps = np.zeros(T, dtype=np.int64) if capture_raster else None
    fv = np.zeros(n, dtype=np.float32) if capture_v else None
    cap = L.ForwardCapture(L.ptr(logits, C.c_float), L.ptr(nspk, C.c_int32),
                           L.ptr(raster, C.c_uint32), L.ptr(fv, C.c_float),
                           L.ptr(steps, C.c_int64))
    st = L.ForwardStats()
    ev, p = event_ptr(events)
    lc = lif.c()
    L.check(L.lib().snc_hip_forward(C.byref(g.view()), p, ev.size, T, C.byref(lc), int(backend),
                                    C.byref(cap), C.byref(st)))
    return ForwardResult(logits[: g.num_output_channels],
                         RunStats(st.spikes, st.synaptic_events, st.overflow, st.ms),
                         Backend(st.used_variant), nspk, raster, fv, steps)


def forward(g: SNNGraph, events, num_steps: int, lif: LIFParams, backend: Backend,
            capture_raster=False, capture_v=False) -> ForwardResult:
    """hip::forward: run one sample, read back. The graph stays resident
    between calls on the same contents (content hash), so the autotuner's
    search for hip-auto runs once per graph. Raises when unavailable."""
    if not backend_is_hip(backend):
        raise ValueError(f"{backend_name(backend)}: the hip runtime serves hip-* backends only")
    hg = resident_graph(g)
    plan = ForwardPlan(hg, events, num_steps, lif, backend)
    try:
        return plan.run(capture_raster=capture_raster, capture_v=capture_v)
    finally:
        plan.close()
