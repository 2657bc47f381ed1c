"""Structural slow path: Connectome mirror (include/snc/connectome.hpp:56-97)
and the two-timescale co-training loop (src/training/cotrain_main.cpp:247-285)
with the learned grow proposal.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .graph import _from_buf

GROW_HEURISTIC, GROW_LEARNED, GROW_LEARNED_POOL = 0, 1, 2


class StructConfigC(C.Structure):
    _fields_ = [("grow_per_epoch", C.c_int32), ("delay", C.c_int32), ("w_grow_init", C.c_float),
                ("locality_window", C.c_float), ("protect_rounds", C.c_int32),
                ("max_attempts_mult", C.c_int32), ("reward_modulated", C.c_int32),
                ("reward_chance", C.c_float), ("reward_floor", C.c_float), ("seed", C.c_uint64),
                ("explore", C.c_float), ("on_device", C.c_int32)]


class StructReportC(C.Structure):
    _fields_ = [("pruned", C.c_int32), ("grown", C.c_int32), ("synapses", C.c_int32),
                ("candidates_scored", C.c_int64), ("score_ms", C.c_double)]


@dataclass
class StructConfig:
    """StructConfig (connectome.hpp:33-48), same defaults."""
    grow_per_epoch: int = 200
    delay: int = 1
    w_grow_init: float = 0.02  # negative: the surviving synapses' median |w| at each step
    locality_window: float = 0.2
    protect_rounds: int = 2
    max_attempts_mult: int = 8
    reward_modulated: bool = False
    reward_chance: float = 0.1
    reward_floor: float = 0.1
    seed: int = 1
    explore: float = 0.0   # logging runs: share of non-greedy (uniform) grow choices
    on_device: bool = False  # sheet connectomes: prune selection and grow candidates on the GPU

    def c(self):
        return StructConfigC(self.grow_per_epoch, self.delay, self.w_grow_init,
                             self.locality_window, self.protect_rounds, self.max_attempts_mult,
                             1 if self.reward_modulated else 0, self.reward_chance,
                             self.reward_floor, self.seed, self.explore,
                             1 if self.on_device else 0)


_SIGS = {
    "snc_conn_layered_local": (C.c_int, [C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int32,
                                         C.c_float, C.c_uint64, C.POINTER(C.c_void_p)]),
    "snc_conn_sheet_recurrent": (C.c_int, [C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int32,
                                           C.c_float, C.c_int32, C.c_float, C.c_uint64,
                                           C.POINTER(C.c_void_p)]),
    "snc_conn_prune_to": (C.c_int, [C.c_void_p, C.c_int32]),
    "snc_conn_free": (None, [C.c_void_p]),
    "snc_conn_num_synapses": (C.c_int32, [C.c_void_p]),
    "snc_conn_to_graph": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "snc_conn_edge_weights": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64]),
    "snc_conn_set_edge_weights": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64]),
    "snc_conn_structural_update": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_int64,
                                             C.POINTER(C.c_int64), C.c_int64,
                                             C.POINTER(StructConfigC), C.c_float, C.c_int32,
                                             C.c_void_p, C.c_double, C.c_int32,
                                             C.POINTER(StructReportC)]),
    "snc_conn_grow_log_size": (C.c_int64, [C.c_void_p]),
    "snc_conn_grow_log": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                    C.c_int64]),
    "snc_conn_grow_log_rank": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64]),
    "snc_conn_candidate_features": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32),
                                              C.POINTER(C.c_int32), C.c_int64,
                                              C.POINTER(C.c_int64), C.c_int64, C.c_double,
                                              C.c_int32, C.POINTER(C.c_float)]),
}
L.register_signatures(_SIGS)


class Connectome:
    def __init__(self, layers, synapse_budget, delay=1, w_init=0.02, seed=1):
        lay = np.ascontiguousarray(layers, np.int32)
        self._h = C.c_void_p()
        L.check(L.lib().snc_conn_layered_local(L.ptr(lay, C.c_int32), lay.size, synapse_budget,
                                               delay, w_init, seed, C.byref(self._h)))
        L.track(self)
        self.layers = list(layers)

    @classmethod
    def sheet_recurrent(cls, layers, fan_in, rec_fan_in=0, radius=0.1, delay=1, w_init=0.02,
                        seed=1) -> "Connectome":
        """Layers on 2-D grids with recurrent hidden layers
        (snc_conn_sheet_recurrent): each neuron of layer l >= 1 draws fan_in
        partners from layer l-1 and, if hidden, rec_fan_in from its own layer,
        p ~ exp(-d / radius) in unit-square coordinates."""
        lay = np.ascontiguousarray(layers, np.int32)
        obj = cls.__new__(cls)
        obj._h = C.c_void_p()
        L.check(L.lib().snc_conn_sheet_recurrent(L.ptr(lay, C.c_int32), lay.size, fan_in,
                                                 rec_fan_in, radius, delay, w_init, seed,
                                                 C.byref(obj._h)))
        L.track(obj)
        obj.layers = list(layers)
        return obj

    def prune_to(self, budget):
        """Remove the smallest-|w| synapses until at most `budget` remain."""
        L.check(L.lib().snc_conn_prune_to(self._h, int(budget)))

    @property
    def num_synapses(self) -> int:
        return int(L.lib().snc_conn_num_synapses(self._h))

    def to_graph(self):
        out = C.c_void_p()
        L.check(L.lib().snc_conn_to_graph(self._h, C.byref(out)))
        return _from_buf(out)

    def edge_weights(self) -> np.ndarray:
        w = np.zeros(self.num_synapses, np.float32)
        L.check(L.lib().snc_conn_edge_weights(self._h, L.ptr(w, C.c_float), w.size))
        return w

    def set_edge_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        L.check(L.lib().snc_conn_set_edge_weights(self._h, L.ptr(w, C.c_float), w.size))

    def structural_update(self, syn_deliveries, neuron_fires, cfg: StructConfig, reward=1.0,
                          mode=GROW_HEURISTIC, model=None, activity_norm=1.0, log=False):
        syn = np.ascontiguousarray(syn_deliveries, np.int64)
        neur = np.ascontiguousarray(neuron_fires, np.int64)
        rep = StructReportC()
        cc = cfg.c()
        L.check(L.lib().snc_conn_structural_update(
            self._h, L.ptr(syn, C.c_int64), syn.size, L.ptr(neur, C.c_int64), neur.size,
            C.byref(cc), reward, mode, model._h if model is not None else None, activity_norm,
            1 if log else 0, C.byref(rep)))
        return {k: getattr(rep, k) for k, _ in rep._fields_}

    def grow_log(self, rank=False):
        """(features, labels): labels 1/0 survived/pruned at the first eligible
        prune (-1 undecided); with rank=True the |w| percentile at that prune
        instead (-1 undecided)."""
        n = int(L.lib().snc_conn_grow_log_size(self._h))
        X = np.zeros((n, 16), np.float32)
        y = np.zeros(n, np.int32)
        L.check(L.lib().snc_conn_grow_log(self._h, L.ptr(X, C.c_float), L.ptr(y, C.c_int32), n))
        if rank:
            r = np.zeros(n, np.float32)
            L.check(L.lib().snc_conn_grow_log_rank(self._h, L.ptr(r, C.c_float), n))
            return X, r
        return X, y

    def candidate_features(self, pre, post, neuron_fires, activity_norm=1.0, delay=1):
        pre = np.ascontiguousarray(pre, np.int32)
        post = np.ascontiguousarray(post, np.int32)
        neur = np.ascontiguousarray(neuron_fires, np.int64)
        F = np.zeros((pre.size, 16), np.float32)
        L.check(L.lib().snc_conn_candidate_features(self._h, L.ptr(pre, C.c_int32),
                                                    L.ptr(post, C.c_int32), pre.size,
                                                    L.ptr(neur, C.c_int64), neur.size,
                                                    activity_norm, delay, L.ptr(F, C.c_float)))
        return F

    def close(self):
        if self._h:
            L.lib().snc_conn_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cotrain(con: Connectome, make_session, outer=12, inner=2, scfg: StructConfig | None = None,
            mode=GROW_HEURISTIC, model=None, log=False, train_idx=None, test_idx=None,
            activity_norm=1.0, on_round=None):
    """Two-timescale loop of cotrain_main.cpp:247-285: per outer round, compile
    the connectome to a frozen graph, train weights for `inner` epochs on the
    GPU (make_session(graph) -> HipTrainSession with the dataset uploaded),
    carry weights back, and run one structural update from the round's
    activity (not after the last round). Epochs cover train_idx (default: the
    session's whole pool). Returns per-round records."""
    scfg = scfg or StructConfig()
    rounds = []
    epoch = 0
    for r in range(outer):
        g = con.to_graph()
        sess = make_session(g)
        if r > 0:
            sess.set_weights(con.edge_weights())
        sess.reset_stats()
        st = None
        for _ in range(inner):
            epoch += 1
            st = sess.train_epoch(epoch, idx=train_idx)
        con.set_edge_weights(sess.get_weights())
        syn, neur = sess.collect_stats()
        test_acc = sess.evaluate(test_idx) if test_idx is not None else float("nan")
        rec = {"round": r, "synapses": con.num_synapses, "train_acc": st.train_acc,
               "test_acc": test_acc, "spikes": st.spikes}
        if r < outer - 1:
            rep = con.structural_update(syn, neur, scfg, reward=float(st.train_acc), mode=mode,
                                        model=model, activity_norm=activity_norm, log=log)
            rec.update(rep)
        rounds.append(rec)
        if on_round:
            on_round(rec)
        sess.close()
    return rounds
