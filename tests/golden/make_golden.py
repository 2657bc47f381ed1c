"""Writes the golden fixtures under tests/golden/ from the reference's own CPU
code, compiled unchanged by oracle/Makefile into oracle/_ref/libsnc_ref.so.

Run in a container that has the reference tree (python tests/golden/make_golden.py).
Every fixture holds inputs and the reference's outputs only:
  fwd_*.npz      graph arrays, events, LIF, and the reference's logits, spike
                 and event totals, per-step spike totals, and per-step rasters
                 of all non-input neurons (public-API prefix runs)
  encode.npz     encode() outputs for direct / poisson / latency
  synthetic.npz  make_synthetic() output
  train_*.npz    one-sample e-prop steps of snc::Trainer: weights before and
                 after, neuron fires and synapse deliveries (train_reward.npz:
                 the three-factor reward rule, reward_mode = 1)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from snc_amd import graph as G  # noqa: E402
from snc_amd import inputs as I  # noqa: E402


class _Lif:
    def __init__(self, threshold, decay, reset, refractory):
        self.threshold, self.decay, self.reset, self.refractory = threshold, decay, reset, refractory
        self.reset_mode = 0


def graph_arrays(g):
    return dict(row_ptr=g.row_ptr, post_ids=g.post_ids, weights=g.weights, delays=g.delays,
                sign=g.sign, role=g.role, channel=g.channel)


def write_forward(name, g, events_list, T, lif):
    out = graph_arrays(g)
    out.update(T=np.int32(T), threshold=np.float32(lif.threshold), decay=np.float32(lif.decay),
               reset=np.float32(lif.reset), refractory=np.int32(lif.refractory))
    for s, ev in enumerate(events_list):
        r = O.ref_forward(g, ev, T, lif)
        ids, ras, sp = O.ref_raster(g, ev, T, lif)
        out[f"events_{s}"] = ev
        out[f"logits_{s}"] = r["logits"]
        out[f"spikes_{s}"] = np.int64(r["spikes"])
        out[f"synaptic_events_{s}"] = np.int64(r["events"])
        out[f"step_spikes_{s}"] = sp
        out[f"raster_ids_{s}"] = ids.astype(np.int32)
        out[f"raster_{s}"] = np.packbits(ras, axis=1, bitorder="little")
    out["n_samples"] = np.int32(len(events_list))
    np.savez_compressed(os.path.join(HERE, name), **out)
    print("wrote", name)


def main():
    if not O.ref_available():
        raise SystemExit("compiled reference missing: run make -C oracle with the reference tree")
    decay = float(np.float32(np.exp(-1.0 / 20.0)))
    lif = _Lif(1.0, decay, 0.0, 1)

    # c0-shaped sheet graph, signed N(0, 1/sqrt(32)) weights, delay 1.
    g0 = G.sheet_graph([784, 256, 10], 32, radius=3.0, sheet=28.0, w_std=1 / np.sqrt(32.0),
                       delay=1, seed=1)
    imgs = I.images(3, 784, 0.7, seed=1).astype(np.float32) / np.float32(255.0)
    evs = [O.ref_encode(1, imgs[i], 50, max_rate=0.2, seed=1 + i) for i in range(3)]
    write_forward("fwd_c0.npz", g0, evs, 50, lif)

    # Distance delays 1..8, inhibitory pres (sign -1), refractory 2.
    g1 = G.sheet_graph([196, 120, 10], 24, recurrent_fan_in=12, radius=2.5, sheet=14.0,
                       w_std=0.35, max_delay=8, seed=3)
    rng = np.random.default_rng(5)
    g1.sign[196:] = np.where(rng.random(g1.num_neurons - 196) < 0.2, -1, 1).astype(np.int8)
    imgs = I.images(2, 196, 0.5, seed=2).astype(np.float32) / np.float32(255.0)
    evs = [O.ref_encode(1, imgs[i], 60, max_rate=0.4, seed=10 + i) for i in range(2)]
    write_forward("fwd_delay_inhib.npz", g1, evs, 60, _Lif(1.0, 0.9, 0.0, 2))

    # Reference generator + direct encoding (the snc_bench defaults, weight .5).
    g2 = G.static_snc([64, 48, 8], 1500, 0.5, 3, 0.2, 4)
    X, _ = O.ref_make_synthetic(2, 64, 8, 0.15, 4)
    evs = [O.ref_encode(0, X[i], 30, seed=4 + i) for i in range(2)]
    write_forward("fwd_static_direct.npz", g2, evs, 30, _Lif(1.0, 0.9, 0.0, 1))

    x = np.random.default_rng(1).random(100).astype(np.float32)
    x[::7] = 0.0
    enc = {f"enc_{k}": O.ref_encode(k, x, 40, gain=1.5, max_rate=0.3, threshold=0.05, seed=77)
           for k in range(3)}
    np.savez_compressed(os.path.join(HERE, "encode.npz"), x=x, **enc)
    X, y = O.ref_make_synthetic(12, 20, 3, 0.25, 42)
    np.savez_compressed(os.path.join(HERE, "synthetic.npz"), X=X, y=y)
    print("wrote encode.npz synthetic.npz")

    write_train()


TRAIN_CASES = [("train_spiking.npz", 0.3, 0.08, 1, 0),
               ("train_direct.npz", 0.25, 0.05, 0, 0),
               ("train_reward.npz", 0.3, 0.5, 1, 1)]


def write_train(names=None):
    """snc::Trainer epochs on a small static-snc graph: weights and stats."""
    for name, w_init, lr, enc, reward in TRAIN_CASES:
        if names and name not in names:
            continue
        g = G.static_snc([64, 48, 8], 1500, 0.5, 1, 0.2, 4)
        X, y = O.ref_make_synthetic(12, 64, 8, 0.35, 3)
        cfg = O.train_config(num_steps=20, lr=lr, w_init=w_init, w_max=4.0, gamma=0.3,
                             feedback=1.0, encoder=enc, max_rate=0.5, seed=5,
                             reward_mode=bool(reward), reward_baseline_decay=0.9)
        r = O.RefTrainer(g, cfg)
        out = graph_arrays(g)
        out.update(X=X, y=y, num_steps=np.int32(20), lr=np.float32(lr), w_init=np.float32(w_init),
                   encoder=np.int32(enc), seed=np.uint64(5), w0=r.weights)
        if reward:
            out.update(reward_mode=np.int32(1), reward_baseline_decay=np.float32(0.9))
        for ep in (1, 2):
            st = r.train_epoch(X, y, ep, 8)
            syn, neur = r.stats()
            out[f"w_{ep}"] = r.weights
            out[f"syn_deliv_{ep}"] = syn
            out[f"neuron_fires_{ep}"] = neur
            out[f"epoch_{ep}"] = np.array([st["train_acc"], st["loss"], st["spikes"],
                                           st["events"]], dtype=np.float64)
        np.savez_compressed(os.path.join(HERE, name), **out)
        print("wrote", name)


if __name__ == "__main__":
    if len(sys.argv) > 1:  # e.g. make_golden.py train_reward.npz: only those train fixtures
        write_train(sys.argv[1:])
    else:
        main()
