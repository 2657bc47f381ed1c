"""Learned grow proposal, end to end: log grow outcomes from heuristic
co-training runs, train the MLP on the GPU, and compare learned against
heuristic growth at an equal synapse budget over several seeds.

Workload: the batched e-prop network 784 -> 512 -> 10 (leaky readout,
prototype images, Poisson input) as a layered local connectome under a
3000-synapse budget (Connectome::layered_local semantics, 1-D positions),
grow = prune = 200 per structural epoch, 12 structural epochs.
"""
from __future__ import annotations

import time

import numpy as np

from . import configs as K
from . import inputs as I
from .grow import GrowMLP
from .structure import GROW_HEURISTIC, GROW_LEARNED, Connectome, StructConfig, cotrain
from .train import HipTrainSession, READOUT_LEAKY, TrainConfig


def dataset(n_train=1024, n_test=512, steps=60, seed=7):
    Xu, y = I.prototype_images(n_train + n_test, 784, 10, 0.7, 0.2, seed=seed)
    X = Xu.astype(np.float32) / np.float32(255.0)
    return X, y, np.full(n_train + n_test, steps, np.int32), n_train


def make_session_factory(X, y, lengths, batch=256, steps=60, seed=1, lr=1.0):
    cfg = TrainConfig(num_steps=steps, lif=K.lif_tau20(), lr=lr, w_init=0.3, w_max=4.0,
                      surrogate_scale=0.3, feedback_scale=1.0, readout_mode=READOUT_LEAKY,
                      readout_psi=1.0, encoder=1, enc_gain=1.0, enc_max_rate=0.2, seed=seed)
    def make(g):
        s = HipTrainSession(g, cfg, batch=batch, max_steps=steps)
        s.upload_features(X, y, lengths)
        s.n_data = X.shape[0]
        return s

    return make, cfg


def run_cotrain(seed, mode, model=None, log=False, data=None, outer=12, inner=1, budget=3000,
                grow=200, steps=60, batch=256):
    X, y, lengths, n_train = data or dataset(steps=steps)
    make, _ = make_session_factory(X, y, lengths, batch=batch, steps=steps, seed=seed)
    con = Connectome([784, 512, 10], budget, 1, 0.3, seed)
    scfg = StructConfig(grow_per_epoch=grow, delay=1, w_grow_init=0.02, locality_window=0.2,
                        protect_rounds=2, seed=seed)
    train_idx = np.arange(n_train, dtype=np.int32)
    test_idx = np.arange(n_train, X.shape[0], dtype=np.int32)
    rounds = cotrain(con, make, outer=outer, inner=inner, scfg=scfg, mode=mode, model=model,
                     log=log, train_idx=train_idx, test_idx=test_idx,
                     activity_norm=float(n_train * steps))
    return con, rounds


def collect_log(seeds=(101, 102, 103), data=None, **kw):
    Xs, ys = [], []
    for s in seeds:
        con, _ = run_cotrain(s, GROW_HEURISTIC, log=True, data=data, **kw)
        X, y = con.grow_log()
        keep = y >= 0
        Xs.append(X[keep])
        ys.append(y[keep])
    return np.concatenate(Xs), np.concatenate(ys).astype(np.float32)


def ab_compare(seeds=(1, 2, 3, 4, 5), train_seeds=(101, 102, 103), outer=12, steps=60,
               mlp_epochs=60):
    """Returns a dict with per-seed final test accuracies for both growth
    rules at the same budget, the MLP's training loss and scoring rate."""
    t0 = time.time()
    data = dataset(steps=steps)
    Xl, yl = collect_log(train_seeds, data=data, outer=outer, steps=steps)
    mlp = GrowMLP(seed=11)
    loss0, _ = mlp.grad(Xl, yl)
    loss = mlp.train(Xl, yl, epochs=mlp_epochs, batch=256, lr=2e-3, seed=5)
    res = {"log_examples": int(Xl.shape[0]), "log_positive_rate": float(yl.mean()),
           "mlp_loss_before": loss0, "mlp_loss_after": loss, "heuristic": [], "learned": [],
           "synapses": []}
    for s in seeds:
        _, rh = run_cotrain(s, GROW_HEURISTIC, data=data, outer=outer, steps=steps)
        _, rl = run_cotrain(s, GROW_LEARNED, model=mlp, data=data, outer=outer, steps=steps)
        res["heuristic"].append(rh[-1]["test_acc"])
        res["learned"].append(rl[-1]["test_acc"])
        res["synapses"].append((rh[-1]["synapses"], rl[-1]["synapses"]))
    # scoring throughput on one million synthetic candidates
    F = np.random.default_rng(3).standard_normal((1 << 20, 16)).astype(np.float32)
    mlp.score(F)
    mlp.score(F)
    res["score_1M_kernel_ms"] = mlp.last_kernel_ms
    res["score_candidates_per_s"] = (1 << 20) / (mlp.last_kernel_ms / 1e3)
    res["wall_s"] = time.time() - t0
    h, l = np.array(res["heuristic"]), np.array(res["learned"])
    res["heuristic_mean"], res["learned_mean"] = float(h.mean()), float(l.mean())
    res["gain_mean"] = float((l - h).mean())
    return res
