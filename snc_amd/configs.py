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


def c1(width=500, height=400, num_steps=1000, seed=1, reset_mode=L.RESET_TO_VALUE):
    g = c1_graph(width, height, seed)
    drive = I.background_drive(g.num_neurons, num_steps, 0.01, 1.0, seed + 1)
    return ForwardWorkload(f"c1_delivery_stress_{g.num_neurons}n_{g.num_synapses}syn", g,
                           [drive], num_steps, lif_tau20(reset_mode))
