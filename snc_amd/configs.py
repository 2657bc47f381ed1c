"""Benchmark workloads of the spiking path, regenerated from explicit seeds.

  c0  plain forward: 784 -> 256 -> 10 LIF on a 2-D sheet (32 partners per
      neuron, p ~ exp(-d/3)), 64 Poisson-encoded 28x28 u8 images, 50 steps.
  c1  delivery stress: 200k LIF on a 500 x 400 sheet, out-degree 100, 1 % hubs
      with power-law in-degree (alpha 2.1, up to 50k), delays 1..32 by
      distance, Bernoulli background drive p = 0.01, 1000 steps, batch 1.
  c2  batched e-prop: 784 -> 512 recurrent LIF -> 10 leaky readout.
  c3  delay-rich recurrent e-prop on 700-channel event streams.
LIF everywhere: tau_mem = 20 steps (decay = exp(-1/20) rounded once to
float32), threshold 1, refractory 1; subtract reset for c0/c2/c3 as their
configurations state, the reference's reset-to-0 for c1 (which states no
reset rule) so the reference's own CPU code can be timed on it unchanged.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import graph as G
from . import inputs as I
from .runtime import LIFParams

DECAY_TAU20 = float(np.float32(math.exp(-1.0 / 20.0)))


def lif_tau20(reset_mode=L.RESET_SUBTRACT) -> LIFParams:
    return LIFParams(threshold=1.0, decay=DECAY_TAU20, reset=0.0, refractory=1,
                     reset_mode=reset_mode)


@dataclass
class ForwardWorkload:
    name: str
    graph: G.SNNGraph
    samples: list          # list of event arrays (one per sample)
    num_steps: int
    lif: LIFParams


def c0_graph(seed=1) -> G.SNNGraph:
    return G.sheet_graph([784, 256, 10], 32, radius=3.0, sheet=28.0,
                         w_std=1.0 / math.sqrt(32.0), delay=1, seed=seed)


def c0_events(n_samples=64, num_steps=50, seed=1):
    imgs = I.images(n_samples, 784, 0.7, seed)
    feats = imgs.astype(np.float32) / np.float32(255.0)
    # Poisson, p = pixel/255 * 0.2, per-sample seed seed + i (snc_bench.cpp:197-198).
    return [I.encode("poisson", feats[i], num_steps, gain=1.0, max_rate=0.2, seed=seed + i)
            for i in range(n_samples)]


def c0(n_samples=64, reset_mode=L.RESET_SUBTRACT, seed=1) -> ForwardWorkload:
    return ForwardWorkload("c0_plain_forward_784-256-10", c0_graph(seed),
                           c0_events(n_samples, 50, seed), 50, lif_tau20(reset_mode))


C1_SHAPE = dict(width=500, height=400, out_degree=100)


def c1_graph(width=500, height=400, seed=1) -> G.SNNGraph:
    return G.hub_sheet_graph(width, height, 100, radius=3.0, hub_fraction=0.01, hub_alpha=2.1,
                             hub_min=100, hub_max=50000, w_std=0.05, max_delay=32,
                             all_input=True, seed=seed)


@dataclass
class TrainWorkload:
    name: str
    graph: G.SNNGraph
    cfg: object            # snc_amd.train.TrainConfig
    batch: int
    max_steps: int
    kind: str              # "features" or "events"
    X: np.ndarray | None
    events: np.ndarray | None
    labels: np.ndarray
    lengths: np.ndarray
    n_train: int
    init_weights: np.ndarray | None = None  # signed start weights (set_weights)


def _lengths(n, lo, hi, seed):
    from .rng_py import splitmix_uniform_ints
    return splitmix_uniform_ints(n, lo, hi, seed)


def c2(n_train=2048, n_test=512, batch=256, seed=1, w_std=0.2) -> TrainWorkload:
    """784 -> 512 recurrent LIF (32 input + 56 recurrent partners per hidden
    neuron, 2-D locality) -> 10 leaky readout, ~50k synapses; prototype images
    with 20 % pixel noise, Poisson p = pixel/255 * 0.2, T_b ~ U[50, 200].
    Training starts from the graph's signed N(0, 0.2) weights: the reference's
    all-positive uniform [0, w_init] start is bistable on this recurrent graph
    (silent at w_init <= 0.04, above 40 % firing at >= 0.07)."""
    from .train import READOUT_LEAKY, TrainConfig
    g = G.sheet_graph([784, 512, 10], 32, recurrent_fan_in=56, dense_readout=True, radius=3.0,
                      sheet=28.0, w_std=w_std, delay=1, seed=seed)
    n = n_train + n_test
    Xu, y = I.prototype_images(n, 784, 10, 0.7, 0.2, seed=seed + 10)
    X = Xu.astype(np.float32) / np.float32(255.0)
    lengths = _lengths(n, 50, 200, seed + 20)
    cfg = TrainConfig(num_steps=200, lif=lif_tau20(), lr=1.0, w_init=0.0, w_max=4.0,
                      surrogate_scale=0.3, feedback_scale=1.0, readout_mode=READOUT_LEAKY,
                      readout_psi=1.0, encoder=1, enc_gain=1.0, enc_max_rate=0.2, seed=seed)
    return TrainWorkload("c2_batched_eprop_784-512r-10", g, cfg, batch, 200, "features", X,
                         None, y, lengths, n_train, g.weights.copy())


def c3(n_train=1024, n_test=256, batch=128, seed=1) -> TrainWorkload:
    """700-channel event streams (~5k events, 20 Gaussian centres, sd 15,
    T ~ U[80, 120]) -> 1024 recurrent LIF (10 % input and recurrent density,
    delays 0..30 mapped to 1..31) -> 20 leaky readout."""
    from .train import READOUT_LEAKY, TrainConfig
    g = G.random_recurrent_graph(700, 1024, 20, 0.1, 0.1, 1, 31, 1.0 / math.sqrt(102.0),
                                 seed=seed)
    n = n_train + n_test
    ev, lengths, y = I.event_streams(n, 700, 20, 20, 15.0, 5000, 80, 120, seed=seed + 30)
    # The configuration states no learning rate for this workload; 3.0 learns
    # the temporal-order task in a few epochs (held-out accuracy 0.08 -> 0.91
    # after 5 epochs, 0.996 after 8; lr 1.0 reaches 0.79 after 8).
    cfg = TrainConfig(num_steps=120, lif=lif_tau20(), lr=3.0, w_init=0.0, w_max=4.0,
                      surrogate_scale=0.3, feedback_scale=1.0, readout_mode=READOUT_LEAKY,
                      readout_psi=1.0, encoder=0, seed=seed)
    return TrainWorkload("c3_delay_rich_eprop_700-1024r-20", g, cfg, batch, 120, "events", None,
                         ev, y, lengths, n_train, g.weights.copy())


def c1(width=500, height=400, num_steps=1000, seed=1, reset_mode=L.RESET_TO_VALUE):
    g = c1_graph(width, height, seed)
    drive = I.background_drive(g.num_neurons, num_steps, 0.01, 1.0, seed + 1)
    return ForwardWorkload(f"c1_delivery_stress_{g.num_neurons}n_{g.num_synapses}syn", g,
                           [drive], num_steps, lif_tau20(reset_mode))
