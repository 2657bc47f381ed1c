"""Learned grow proposal, end to end: log grow outcomes from the substrate's
own heuristic co-training runs, train the MLP on the GPU with its own
forward/backward kernels, and compare learned against heuristic growth at an
equal synapse budget over several seeds.

Workload (structural co-training with a learned grow proposal): the batched
e-prop network's data and training setup (784 inputs, 512 hidden LIF, 10
leaky readout; prototype images with 20 % pixel noise, Poisson p = pixel/255
* 0.2, T_b ~ U[50, 200], lr 1.0, minibatch 256) as a layered local connectome
(Connectome::layered_local semantics) under a 3000-synapse budget; grow =
prune = 200 per structural epoch, 12 structural epochs of one training epoch
each. Logging runs explore: a share of freed posts regrows from a uniformly
drawn window candidate, so the log holds outcomes of non-greedy choices too.

kind="sheet" runs the same loop on the 2-D recurrent sheet connectome
(SHEET below) with the structural step on the GPU. kind="c4" is the
structural co-training workload as stated (C4 below): the batched e-prop
network itself, thinned to a 3000-synapse budget, with the learned rule
scoring the whole candidate pool every structural step.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import configs as K
from .grow import GrowMLP
from .structure import (GROW_HEURISTIC, GROW_LEARNED, GROW_LEARNED_POOL, Connectome,
                        StructConfig, cotrain)
from .train import HipTrainSession

LAYERS = [784, 512, 10]


def workload(n_train=2048, n_test=512, seed=1):
    """Data and trainer configuration of the co-training runs."""
    w = K.c2(n_train=n_train, n_test=n_test, seed=seed)
    cfg = dataclasses.replace(w.cfg, w_init=0.3)
    return w, cfg


# The sheet variant: 784 -> 512 (recurrent) -> 10 on 2-D grids, 32 partners
# from the previous layer and 8 recurrent per hidden neuron (p ~ exp(-d/0.2)),
# signed N(0, 0.2) start weights, grow = prune = 300 per round within discs of
# radius 0.1, structural step on the GPU.
SHEET = dict(fan_in=32, rec_fan_in=8, radius=0.2, w_std=0.2, grow=300, window=0.1)

# The c4 base: the batched e-prop network's sheets (784 -> 512 recurrent ->
# 10; 32 partners from the previous layer and 56 recurrent per hidden
# neuron, p ~ exp(-d / r), r = 3 spacings of the 28-wide input grid; the
# readout draws 32 hidden partners) with signed N(0, 0.2) weights, thinned
# by |w| to a 3000-synapse budget; grow = prune = 200 per structural epoch.
# Grow windows of radius 1.5 cover the whole previous (and own) layer, so
# the candidate pool is every pair the architecture allows that is not yet
# connected: about 668k pairs, all scored by the learned rule each step.
# "local_window" is the heuristic's own locality (a disc of radius 0.1),
# run as a second baseline.
C4 = dict(fan_in=32, rec_fan_in=56, radius=3.0 / 27.0, w_std=0.2, budget=3000, grow=200,
          window=1.5, local_window=0.1)


def run_cotrain(seed, mode, w, cfg, model=None, log=False, explore=0.0, outer=12, budget=3000,
                grow=200, on_round=None, kind="layered", window=None, w_grow=0.02):
    """One co-training run: returns (connectome, per-round records). kind
    "layered" is the reference's Connectome; "sheet" the 2-D recurrent one;
    "c4" the batched e-prop network thinned to C4's budget. `window`
    overrides the kind's grow window."""
    if kind == "c4":
        con = Connectome.sheet_recurrent(LAYERS, fan_in=C4["fan_in"], rec_fan_in=C4["rec_fan_in"],
                                         radius=C4["radius"], seed=seed)
        con.set_edge_weights((np.random.default_rng(seed + 1000).standard_normal(
            con.num_synapses) * C4["w_std"]).astype(np.float32))
        con.prune_to(C4["budget"])
        grow, window, on_device = C4["grow"], window or C4["window"], True
    elif kind == "sheet":
        con = Connectome.sheet_recurrent(LAYERS, fan_in=SHEET["fan_in"],
                                         rec_fan_in=SHEET["rec_fan_in"], radius=SHEET["radius"],
                                         seed=seed)
        con.set_edge_weights((np.random.default_rng(seed + 1000).standard_normal(
            con.num_synapses) * SHEET["w_std"]).astype(np.float32))
        grow, window, on_device = SHEET["grow"], SHEET["window"], True
    else:
        con = Connectome(LAYERS, budget, 1, 0.02, seed)
        window, on_device = 0.2, False

    def make(g):
        s = HipTrainSession(g, dataclasses.replace(cfg, seed=seed), batch=w.batch,
                            max_steps=w.max_steps)
        s.upload_features(w.X, w.labels, w.lengths)
        if kind in ("sheet", "c4"):
            s.set_weights(con.edge_weights())  # signed start weights from structural epoch 0
        return s

    scfg = StructConfig(grow_per_epoch=grow, delay=1, w_grow_init=w_grow, locality_window=window,
                        protect_rounds=2, seed=seed, explore=explore, on_device=on_device)
    n = w.labels.size
    train_idx = np.arange(w.n_train, dtype=np.int32)
    test_idx = np.arange(w.n_train, n, dtype=np.int32)
    steps = float(w.lengths[: w.n_train].sum())
    rounds = cotrain(con, make, outer=outer, inner=1, scfg=scfg, mode=mode, model=model, log=log,
                     train_idx=train_idx, test_idx=test_idx, activity_norm=steps,
                     on_round=on_round)
    return con, rounds


def collect_log(seeds, w, cfg, explore=0.5, outer=12, kind="layered", label="survival",
                model=None, w_grow=0.02):
    """Labelled grow events from heuristic runs (or, with a model, from
    learned-mode runs that use it): (features, survived) or, with
    label="rank", (features, |w| percentile at the edge's first eligible
    prune) -- a graded target instead of a mostly-1 binary one."""
    Xs, ys = [], []
    learned_mode = GROW_LEARNED_POOL if kind == "c4" else GROW_LEARNED
    for s in seeds:
        mode = GROW_HEURISTIC if model is None else learned_mode
        con, _ = run_cotrain(s, mode, w, cfg, model=model, log=True, explore=explore,
                             outer=outer, kind=kind, w_grow=w_grow)
        X, y = con.grow_log(rank=label == "rank")
        keep = y >= 0
        Xs.append(X[keep])
        ys.append(y[keep])
    return np.concatenate(Xs), np.concatenate(ys).astype(np.float32)


def ab_compare(seeds=(1, 2, 3, 4, 5), train_seeds=(101, 102, 103), outer=12, mlp_epochs=40,
               explore=0.5, log=print, kind="layered", label="survival", onpolicy=0,
               w_grow=0.02):
    """Learned vs heuristic grow at equal budget. Returns a dict with the
    per-seed final test accuracies and synapse counts, the MLP's loss before
    and after training, and its scoring rate on one million candidates."""
    t0 = time.time()
    w, cfg = workload()
    Xl, yl = collect_log(train_seeds, w, cfg, explore=explore, outer=outer, kind=kind,
                         label=label, w_grow=w_grow)
    log(f"grow log: {Xl.shape[0]} labelled events, mean {label} label {yl.mean():.3f} "
        f"({time.time() - t0:.1f}s)")
    mlp = GrowMLP(seed=11)
    loss0, _ = mlp.grad(Xl, yl)
    loss = mlp.train(Xl, yl, epochs=mlp_epochs, batch=256, lr=2e-3, seed=5)
    log(f"MLP BCE {loss0:.4f} -> {loss:.4f}")
    # On-policy rounds: log the outcomes of the model's own choices (learned
    # mode, same exploration) and retrain on everything logged so far.
    for k in range(onpolicy):
        Xo, yo = collect_log(train_seeds, w, cfg, explore=explore, outer=outer, kind=kind,
                             label=label, model=mlp, w_grow=w_grow)
        Xl, yl = np.concatenate([Xl, Xo]), np.concatenate([yl, yo])
        loss = mlp.train(Xl, yl, epochs=mlp_epochs, batch=256, lr=2e-3, seed=6 + k)
        log(f"on-policy round {k + 1}: +{Xo.shape[0]} events (mean label {yo.mean():.3f}), "
            f"MLP BCE {loss:.4f}")
    res = {"label": label, "onpolicy_rounds": onpolicy, "w_grow_init": w_grow, "log_examples": int(Xl.shape[0]),
           "log_label_mean": float(yl.mean()),
           "mlp_loss_before": float(loss0), "mlp_loss_after": float(loss),
           "heuristic": [], "learned": [], "synapses": [], "candidates_scored_per_step": []}
    learned_mode = GROW_LEARNED_POOL if kind == "c4" else GROW_LEARNED
    if kind == "c4":
        res["heuristic_local"] = []
    for s in seeds:
        _, rh = run_cotrain(s, GROW_HEURISTIC, w, cfg, outer=outer, kind=kind, w_grow=w_grow)
        _, rl = run_cotrain(s, learned_mode, w, cfg, model=mlp, outer=outer, kind=kind,
                            w_grow=w_grow)
        res["heuristic"].append(rh[-1]["test_acc"])
        res["learned"].append(rl[-1]["test_acc"])
        syn = [rh[-1]["synapses"], rl[-1]["synapses"]]
        msg = ""
        if kind == "c4":
            # the heuristic at its own locality as a second baseline
            _, rk = run_cotrain(s, GROW_HEURISTIC, w, cfg, outer=outer, kind=kind,
                                window=C4["local_window"], w_grow=w_grow)
            res["heuristic_local"].append(rk[-1]["test_acc"])
            syn.append(rk[-1]["synapses"])
            msg = f" heuristic-local {rk[-1]['test_acc']:.4f}"
        res["synapses"].append(syn)
        cs = [r["candidates_scored"] for r in rl if "candidates_scored" in r]
        res["candidates_scored_per_step"].append(float(np.mean(cs)) if cs else 0.0)
        log(f"seed {s}: heuristic {rh[-1]['test_acc']:.4f} learned {rl[-1]['test_acc']:.4f}{msg} "
            f"synapses {syn} candidates/step {res['candidates_scored_per_step'][-1]:.0f}")
    F = np.random.default_rng(3).standard_normal((1 << 20, 16)).astype(np.float32)
    mlp.score(F)
    mlp.score(F)
    res["score_1M_kernel_ms"] = mlp.last_kernel_ms
    res["score_candidates_per_s"] = (1 << 20) / (mlp.last_kernel_ms / 1e3)
    h, l = np.array(res["heuristic"]), np.array(res["learned"])
    res["heuristic_mean"], res["learned_mean"] = float(h.mean()), float(l.mean())
    res["gain_mean_pt"] = float(100.0 * (l - h).mean())
    res["wins"] = int((l > h).sum())
    if kind == "c4":
        k = np.array(res["heuristic_local"])
        res["heuristic_local_mean"] = float(k.mean())
        res["gain_vs_local_mean_pt"] = float(100.0 * (l - k).mean())
        res["wins_vs_local"] = int((l > k).sum())
    res["wall_s"] = time.time() - t0
    return res


if __name__ == "__main__":
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(description="learned vs heuristic grow at equal budget")
    ap.add_argument("--seeds", default="1,2,3,4,5")
    ap.add_argument("--train-seeds", default="101,102,103")
    ap.add_argument("--outer", type=int, default=12)
    ap.add_argument("--mlp-epochs", type=int, default=40)
    ap.add_argument("--explore", type=float, default=0.5)
    ap.add_argument("--kind", choices=["layered", "sheet", "c4"], default="layered")
    ap.add_argument("--label", choices=["survival", "rank"], default="survival")
    ap.add_argument("--onpolicy", type=int, default=0,
                    help="rounds of logging the model's own choices and retraining")
    ap.add_argument("--w-grow", type=float, default=0.02,
                    help="starting weight of a grown synapse (StructConfig.w_grow_init; "
                         "negative: the surviving synapses' median |w|)")
    a = ap.parse_args()
    out = ab_compare(tuple(int(x) for x in a.seeds.split(",")),
                     tuple(int(x) for x in a.train_seeds.split(",")), a.outer, a.mlp_epochs,
                     a.explore, log=lambda m: print(m, file=sys.stderr, flush=True), kind=a.kind,
                     label=a.label, onpolicy=a.onpolicy, w_grow=a.w_grow)
    print(json.dumps(out))
