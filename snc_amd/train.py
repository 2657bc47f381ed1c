"""e-prop trainer session mirror of cudatrain:: (include/snc/cuda_trainer.hpp:24-57).

``HipTrainSession`` keeps the reference lifecycle -- create, set/get weights,
train_epoch, evaluate, reset/collect stats, destroy -- on top of the C ABI
(snc_hip_train_*). Three execution modes share one definition of the math:
PLAIN runs one sample at a time with per-step kernel launches; BATCHED runs a
whole minibatch of uneven-length samples in one persistent launch; SERIAL runs
one sample at a time, each as one launch of that persistent kernel (the
per-sample baseline of BATCHED, with bit-identical weights).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .graph import SNNGraph
from .inputs import EVENT_DTYPE
from .runtime import LIFParams

PLAIN, BATCHED, SERIAL = 0, 1, 2
ELIG_F32, ELIG_F64 = 0, 1
READOUT_SPIKING, READOUT_LEAKY = 0, 1


class TrainConfigC(C.Structure):
    _fields_ = [("num_steps", C.c_int32), ("lif", L.Lif), ("lr", C.c_float),
                ("w_init", C.c_float), ("w_max", C.c_float), ("surrogate_scale", C.c_float),
                ("feedback_scale", C.c_float), ("train_hidden", C.c_int32),
                ("readout_mode", C.c_int32), ("readout_psi", C.c_float), ("encoder", C.c_int32),
                ("enc_gain", C.c_float), ("enc_max_rate", C.c_float),
                ("enc_threshold", C.c_float), ("seed", C.c_uint64),
                ("reward_mode", C.c_int32), ("reward_baseline_decay", C.c_float)]


class TrainStatsC(C.Structure):
    _fields_ = [("spikes", C.c_int64), ("synaptic_events", C.c_int64), ("loss_sum", C.c_double),
                ("correct", C.c_int32), ("samples", C.c_int32), ("ms", C.c_double)]


@dataclass
class TrainConfig:
    """TrainConfig (include/snc/trainer.hpp:30-48) plus the readout mode."""
    num_steps: int = 30
    lif: LIFParams = None
    lr: float = 0.02
    w_init: float = 0.3
    w_max: float = 4.0
    surrogate_scale: float = 0.3
    feedback_scale: float = 1.0
    train_hidden: bool = True
    readout_mode: int = READOUT_SPIKING
    readout_psi: float = 1.0
    encoder: int = 1           # 0 direct, 1 poisson (Philox on device), 2 latency
    enc_gain: float = 1.0
    enc_max_rate: float = 0.5
    enc_threshold: float = 0.0
    seed: int = 1
    reward_mode: bool = False           # three-factor reward rule (trainer.cpp:165-179)
    reward_baseline_decay: float = 0.99

    def c(self) -> TrainConfigC:
        lif = self.lif or LIFParams()
        return TrainConfigC(self.num_steps, lif.c(), self.lr, self.w_init, self.w_max,
                            self.surrogate_scale, self.feedback_scale,
                            1 if self.train_hidden else 0, self.readout_mode, self.readout_psi,
                            self.encoder, self.enc_gain, self.enc_max_rate, self.enc_threshold,
                            self.seed, 1 if self.reward_mode else 0, self.reward_baseline_decay)


@dataclass
class EpochStats:
    """EpochStats (trainer.hpp:50-57); test_acc filled by the caller."""
    loss: float = 0.0
    train_acc: float = 0.0
    test_acc: float = 0.0
    spikes: int = 0
    synaptic_events: int = 0
    energy: float = 0.0
    ms: float = 0.0
    samples: int = 0


_SIGS = {
    "snc_hip_train_create": (C.c_int, [C.POINTER(L.GraphView), C.POINTER(TrainConfigC), C.c_int32,
                                       C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "snc_hip_train_destroy": (None, [C.c_void_p]),
    "snc_hip_train_set_weights": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64]),
    "snc_hip_train_get_weights": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64]),
    "snc_hip_train_upload_features": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int32,
                                                C.c_int32, C.POINTER(C.c_int32),
                                                C.POINTER(C.c_int32)]),
    "snc_hip_train_upload_events": (C.c_int, [C.c_void_p, C.POINTER(L.Event), C.c_int64,
                                              C.c_int32, C.POINTER(C.c_int32),
                                              C.POINTER(C.c_int32)]),
    "snc_hip_train_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                     C.c_int32, C.c_int32, C.POINTER(C.c_float),
                                     C.POINTER(C.c_uint32), C.POINTER(TrainStatsC)]),
    "snc_hip_train_epoch": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                      C.c_int32, C.POINTER(TrainStatsC)]),
    "snc_hip_train_evaluate": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                         C.POINTER(TrainStatsC)]),
    "snc_hip_train_profile": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                        C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]),
    "snc_hip_train_reset_stats": (C.c_int, [C.c_void_p]),
    "snc_hip_train_collect_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64),
                                              C.POINTER(C.c_int64)]),
}


def _bind():
    lib = L.lib()
    if not getattr(lib, "_train_bound", False):
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib._train_bound = True
    return lib


L.register_signatures(_SIGS)


def _stats(st: TrainStatsC) -> EpochStats:
    n = max(1, st.samples)
    return EpochStats(loss=st.loss_sum / n, train_acc=st.correct / n, spikes=st.spikes,
                      synaptic_events=st.synaptic_events,
                      energy=1.0 * st.spikes + 0.2 * st.synaptic_events, ms=st.ms,
                      samples=st.samples)


# Phases of the batched kernel (SNC_TRAIN_PHASES, snc_hip_train_profile).
PHASE_NAMES = ("setup", "lif_trace_history", "delivery_bits", "prefix_scan", "arrival_staging",
               "fold_inject", "epilogue", "eligibility", "history_stream", "contrib_store",
               "queue_wait")
PHASES = len(PHASE_NAMES)


class HipTrainSession:
    """cudatrain:: session on gfx950 (create / set_weights / get_weights /
    train_epoch / evaluate / reset_stats / collect_stats / destroy)."""

    def __init__(self, g: SNNGraph, cfg: TrainConfig, batch: int, max_steps: int | None = None,
                 elig: int = ELIG_F64):
        self.graph = g
        self.cfg = cfg
        self.batch = int(batch)
        self.max_steps = int(max_steps or cfg.num_steps)
        self.classes = max(1, g.num_output_channels)
        self._h = C.c_void_p()
        lib = _bind()
        cc = cfg.c()
        L.check(lib.snc_hip_train_create(C.byref(g.view()), C.byref(cc), self.batch,
                                         self.max_steps, elig, C.byref(self._h)))
        L.track(self)
        self.n_data = 0

    # weights ---------------------------------------------------------------
    def get_weights(self) -> np.ndarray:
        w = np.zeros(self.graph.num_synapses, np.float32)
        L.check(_bind().snc_hip_train_get_weights(self._h, L.ptr(w, C.c_float), w.size))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        L.check(_bind().snc_hip_train_set_weights(self._h, L.ptr(w, C.c_float), w.size))

    # datasets ----------------------------------------------------------------
    def upload_features(self, X, labels, lengths=None):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(labels, np.int32)
        ln = None if lengths is None else np.ascontiguousarray(lengths, np.int32)
        L.check(_bind().snc_hip_train_upload_features(self._h, L.ptr(X, C.c_float), X.shape[0],
                                                      X.shape[1], L.ptr(y, C.c_int32),
                                                      L.ptr(ln, C.c_int32)))
        self.n_data = X.shape[0]

    def upload_events(self, events, n, labels, lengths=None):
        ev = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
        y = np.ascontiguousarray(labels, np.int32)
        ln = None if lengths is None else np.ascontiguousarray(lengths, np.int32)
        evp = ev.ctypes.data_as(C.POINTER(L.Event)) if ev.size else None
        L.check(_bind().snc_hip_train_upload_events(self._h, evp, ev.size, n, L.ptr(y, C.c_int32),
                                                    L.ptr(ln, C.c_int32)))
        self.n_data = n

    # training ----------------------------------------------------------------
    def step(self, idx, epoch=0, mode=BATCHED, apply=True, want_logits=False,
             want_raster=False):
        idx = np.ascontiguousarray(idx, np.int32)
        B = idx.size
        logits = np.zeros((B, self.classes), np.float32) if want_logits else None
        words = (self.graph.num_neurons + 31) // 32
        raster = np.zeros((B, self.max_steps, words), np.uint32) if want_raster else None
        st = TrainStatsC()
        L.check(_bind().snc_hip_train_step(self._h, L.ptr(idx, C.c_int32), B, epoch, mode,
                                           1 if apply else 0, L.ptr(logits, C.c_float),
                                           L.ptr(raster, C.c_uint32), C.byref(st)))
        return _stats(st), logits, raster

    def train_epoch(self, epoch, mode=BATCHED, idx=None) -> EpochStats:
        """One shuffled pass over idx (default: the whole uploaded pool)."""
        st = TrainStatsC()
        if idx is None:
            L.check(_bind().snc_hip_train_epoch(self._h, None, 0, epoch, mode, C.byref(st)))
        else:
            idx = np.ascontiguousarray(idx, np.int32)
            L.check(_bind().snc_hip_train_epoch(self._h, L.ptr(idx, C.c_int32), idx.size, epoch,
                                                mode, C.byref(st)))
        return _stats(st)

    def evaluate(self, idx=None, mode=BATCHED) -> float:
        if idx is None:
            idx = np.arange(self.n_data, dtype=np.int32)
        idx = np.ascontiguousarray(idx, np.int32)
        st = TrainStatsC()
        L.check(_bind().snc_hip_train_evaluate(self._h, L.ptr(idx, C.c_int32), idx.size, mode,
                                               C.byref(st)))
        return st.correct / max(1, st.samples)

    # structural statistics -----------------------------------------------------
    def reset_stats(self):
        L.check(_bind().snc_hip_train_reset_stats(self._h))

    def profile(self, idx, epoch=0) -> dict:
        """Phase split of one batched minibatch (snc_hip_train_profile): per
        phase, the mean and largest time (us) over the minibatch's workgroups,
        and the launch's device time; no update is applied."""
        i = np.ascontiguousarray(idx, np.int32)
        mean = np.zeros(PHASES, np.float64)
        mx = np.zeros(PHASES, np.float64)
        ms = C.c_double()
        L.check(_bind().snc_hip_train_profile(self._h, L.ptr(i, C.c_int32), i.size, int(epoch),
                                              L.ptr(mean, C.c_double), L.ptr(mx, C.c_double),
                                              C.byref(ms)))
        return {"kernel_ms": ms.value,
                "phases": {name: {"mean_us": float(mean[k]), "max_us": float(mx[k])}
                           for k, name in enumerate(PHASE_NAMES)}}

    def collect_stats(self):
        syn = np.zeros(self.graph.num_synapses, np.int64)
        neur = np.zeros(self.graph.num_neurons, np.int64)
        L.check(_bind().snc_hip_train_collect_stats(self._h, L.ptr(syn, C.c_int64),
                                                    L.ptr(neur, C.c_int64)))
        return syn, neur

    def close(self):
        if self._h:
            _bind().snc_hip_train_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
