"""Small hand-built graphs and inputs shared by the CPU and GPU tests."""
from __future__ import annotations

import numpy as np

from snc_amd.graph import SNNGraph
from snc_amd.inputs import EVENT_DTYPE


def make_graph(n, edges, roles=None, channels=None, signs=None) -> SNNGraph:
    """edges: iterable of (pre, post, weight, delay); CSR rows keep the given
    order within a pre."""
    edges = list(edges)
    order = sorted(range(len(edges)), key=lambda k: (edges[k][0], k))
    pre = np.array([edges[k][0] for k in order], dtype=np.int64)
    row_ptr = np.zeros(n + 1, dtype=np.int32)
    if len(edges):
        np.add.at(row_ptr, pre + 1, 1)
    row_ptr = np.cumsum(row_ptr).astype(np.int32)
    post = np.array([edges[k][1] for k in order], dtype=np.int32)
    w = np.array([edges[k][2] for k in order], dtype=np.float32)
    d = np.array([edges[k][3] for k in order], dtype=np.int32)
    role = np.zeros(n, np.uint8) if roles is None else np.asarray(roles, np.uint8)
    ch = np.full(n, -1, np.int32) if channels is None else np.asarray(channels, np.int32)
    sg = np.ones(n, np.int8) if signs is None else np.asarray(signs, np.int8)
    return SNNGraph(row_ptr, post, w, d, sg, role, ch)


def events(triples) -> np.ndarray:
    """triples: iterable of (channel, time, amplitude)."""
    t = list(triples)
    ev = np.zeros(len(t), dtype=EVENT_DTYPE)
    for k, (c, tm, a) in enumerate(t):
        ev[k] = (0, c, tm, a)
    return ev


def random_graph(rng, n, m, max_delay, w_std=0.4, frac_input=0.3, frac_output=0.1,
                 inhib=0.2) -> SNNGraph:
    """Random directed multigraph with roles, signs and delays in [1, max_delay]."""
    pre = rng.integers(0, n, m)
    post = rng.integers(0, n, m)
    w = (rng.standard_normal(m) * w_std).astype(np.float32)
    d = rng.integers(1, max_delay + 1, m)
    edges = list(zip(pre.tolist(), post.tolist(), w.tolist(), d.tolist()))
    u = rng.random(n)
    roles = np.where(u < frac_input, 1, np.where(u > 1 - frac_output, 2, 0)).astype(np.uint8)
    channels = np.full(n, -1, np.int32)
    n_in = int((roles == 1).sum())
    n_out = int((roles == 2).sum())
    channels[roles == 1] = rng.integers(0, max(1, n_in // 2 + 1), n_in)
    channels[roles == 2] = rng.integers(0, max(1, n_out), n_out)
    signs = np.where(rng.random(n) < inhib, -1, 1).astype(np.int8)
    return make_graph(n, edges, roles, channels, signs)


def random_events(rng, g: SNNGraph, T, rate=0.2, amp=(0.2, 1.4)) -> np.ndarray:
    n_ch = g.num_input_channels
    if n_ch == 0:
        return np.zeros(0, dtype=EVENT_DTYPE)
    k = int(rate * n_ch * T)
    ev = np.zeros(k, dtype=EVENT_DTYPE)
    ev["channel_id"] = rng.integers(0, n_ch, k)
    ev["time"] = rng.integers(0, T, k)
    ev["amplitude"] = rng.uniform(amp[0], amp[1], k).astype(np.float32)
    return ev
