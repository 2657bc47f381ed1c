"""hip-pull (batched gather over halo-indexed incoming columns) through the C ABI.

Light targets (in-degree up to the plan's cap) sum their inputs in column
order, and hub targets walk their chunks' per-sample arrival lists in column
order too, so every sample's raster, counts, logits and membranes equal the
CPU oracle's bit for bit (and hip-bucket's), with 32-bit fired words (up to
32 samples) and 64-bit ones (33 to 64), with and without hubs.
"""
import numpy as np
import pytest

import oracle as O
from snc_amd import _lib as L
from snc_amd import configs as K
from snc_amd import graph as G
from snc_amd import inputs as I
from snc_amd.runtime import Backend, ForwardPlan, HipGraph, PlanConfig

pytestmark = pytest.mark.gpu



def _drives(n, T, B, seed, p0=0.01, dp=0.003):
    evs = [I.background_drive(n, T, p0 + dp * (b % 5), 1.0, seed=seed + b) for b in range(B)]
    allev = np.concatenate([e.copy() for e in evs])
    off = 0
    for b, e in enumerate(evs):
        allev["sample_id"][off:off + e.size] = b
        off += e.size
    return evs, allev


def _pull(hg, allev, T, lif, B, cap, blk=256):
    print(f"hip-pull plan: B={B} T={T} cap={cap} block={blk}", flush=True)
    return ForwardPlan(hg, allev, T, lif, config=PlanConfig(Backend.HipPull, blk, cap, False),
                       num_samples=B)


@pytest.mark.parametrize("B,blk", [(1, 256), (7, 256), (32, 256), (33, 256), (64, 256),
                                   (33, 208), (20, 64)])
def test_pull_without_hubs_is_bit_exact(B, blk):
    g = G.sheet_graph([400, 300, 20], 24, recurrent_fan_in=16, radius=3.0, sheet=20.0,
                      w_std=0.35, max_delay=9, seed=3)
    cap = int(g.in_degree().max())
    lif = K.lif_tau20()
    T = 60
    evs, allev = _drives(g.num_neurons, T, B, seed=40, p0=0.05, dp=0.01)
    hg = HipGraph(g)
    m0 = hg.memory()
    plan = _pull(hg, allev, T, lif, B, cap, blk)
    assert hg.memory() > m0  # the light stream and halos count toward the graph
    p0 = plan.memory()
    r = plan.run(capture_raster=True, capture_v=True)
    assert plan.memory() >= p0 + B * T * ((g.num_neurons + 31) // 32) * 4  # the raster
    assert r.used_backend == Backend.HipPull
    for b in range(B):
        ref = O.forward(g, evs[b], T, lif)
        assert ref["spikes"] > 0
        assert np.array_equal(r.raster[b], ref["raster"]), b
        assert np.array_equal(r.neuron_spikes[b], ref["neuron_spikes"])
        assert np.array_equal(r.logits[b], ref["logits"])
        assert np.array_equal(r.final_v[b].view(np.uint32), ref["final_v"].view(np.uint32))
        assert np.array_equal(r.step_spikes[b], ref["step_spikes"])


@pytest.mark.parametrize("B", [12, 50])
def test_pull_with_hubs_is_bit_exact(B):
    """Hub columns of up to 5000 entries (up to ten 512-entry chunks), at 32-
    and 64-bit fired words: raster, counts, logits and membranes bit for bit,
    and the same again on a second run of the plan."""
    g = G.hub_sheet_graph(60, 50, 40, hub_fraction=0.02, hub_min=40, hub_max=5000,
                          w_std=0.06, max_delay=20, seed=9)
    lif = K.lif_tau20()
    T = 90
    evs, allev = _drives(g.num_neurons, T, B, seed=20)
    hg = HipGraph(g)
    cap = 64
    deg = g.in_degree()
    assert int((deg > cap).sum()) > 20 and int(deg.max()) > 1024  # hubs, some of several chunks
    plan = _pull(hg, allev, T, lif, B, cap)
    r = plan.run(capture_raster=True, capture_v=True)
    sp, ev = plan.sample_stats()
    assert sp.sum() == r.stats.spikes and ev.sum() == r.stats.synaptic_events
    for b in range(B):
        ref = O.forward(g, evs[b], T, lif)
        assert sp[b] == ref["spikes"] and ev[b] == ref["events"]
        assert np.array_equal(r.raster[b], ref["raster"]), b
        assert np.array_equal(r.neuron_spikes[b], ref["neuron_spikes"])
        assert np.array_equal(r.logits[b], ref["logits"])
        assert np.array_equal(r.final_v[b].view(np.uint32), ref["final_v"].view(np.uint32)), b
        assert np.array_equal(r.step_spikes[b], ref["step_spikes"])
    r2 = plan.run(capture_raster=True, capture_v=True)
    assert np.array_equal(r2.raster, r.raster)
    assert np.array_equal(r2.final_v.view(np.uint32), r.final_v.view(np.uint32))


def test_pull_rejects_what_it_cannot_hold():
    g = G.sheet_graph([64, 48, 8], 12, seed=2)
    hg = HipGraph(g)
    lif = K.lif_tau20()
    _, allev = _drives(g.num_neurons, 10, 65, seed=1)
    with pytest.raises(L.HipError, match="at most 64"):
        _pull(hg, allev, 10, lif, 65, 64)
    for blk in (512, 48, 200):
        with pytest.raises(L.HipError, match="64 to 256 neurons"):
            ForwardPlan(hg, allev[allev["sample_id"] < 2], 10, lif,
                        config=PlanConfig(Backend.HipPull, blk, 64, False), num_samples=2)


def test_pull_dense_input_fills_the_largest_pool():
    """Every input neuron driven on every step in all 32 samples: 8192 staged
    injections per (step, block), the largest pool, still bit-exact."""
    g = G.sheet_graph([256, 64, 8], 12, seed=6, w_std=0.2)
    lif = K.lif_tau20()
    T, B = 12, 32
    evs = []
    for b in range(B):
        e = np.zeros(256 * T, dtype=I.EVENT_DTYPE)
        e["channel_id"] = np.tile(np.arange(256), T)
        e["time"] = np.repeat(np.arange(T), 256)
        e["amplitude"] = 0.05 + 0.01 * b
        evs.append(e)
    allev = np.concatenate([e.copy() for e in evs])
    allev["sample_id"] = np.repeat(np.arange(B), 256 * T)
    hg = HipGraph(g)
    r = _pull(hg, allev, T, lif, B, int(g.in_degree().max())).run(capture_raster=True,
                                                                   capture_v=True)
    for b in range(B):
        ref = O.forward(g, evs[b], T, lif)
        assert np.array_equal(r.raster[b], ref["raster"])
        assert np.array_equal(r.final_v[b].view(np.uint32), ref["final_v"].view(np.uint32))


def test_batched_autotuner_considers_pull():
    g = G.hub_sheet_graph(60, 50, 40, hub_fraction=0.02, hub_min=40, hub_max=1500,
                          w_std=0.06, max_delay=20, seed=9)
    lif = K.lif_tau20()
    T, B = 60, 16
    evs, allev = _drives(g.num_neurons, T, B, seed=5)
    hg = HipGraph(g)
    plan = ForwardPlan(hg, allev, T, lif, Backend.HipAuto, num_samples=B)
    rep = hg.tune_report(batched=True)
    assert any(c.variant == Backend.HipPull for c, _ in rep["candidates"])
    r = plan.run(capture_v=True)
    for b in range(B):
        ref = O.forward(g, evs[b], T, lif)
        assert np.array_equal(r.neuron_spikes[b], ref["neuron_spikes"])


@pytest.mark.parametrize("B", [9, 40])
def test_pull_halo_overflow_reads_fired_words_directly(B):
    """Random recurrent columns (no locality): each 256-target block uses more
    distinct (pre, delay) words than its LDS halo holds, so part of every
    column reads F directly; still bit-exact without hubs."""
    g = G.random_recurrent_graph(700, 1024, 20, 0.1, 0.1, 1, 31, 1.0 / np.sqrt(102.0), seed=4)
    cap = int(g.in_degree().max())
    assert cap <= 1024
    lif = K.lif_tau20()
    T = 70
    evs, allev = _drives(g.num_neurons, T, B, seed=70, p0=0.04, dp=0.01)
    hg = HipGraph(g)
    r = _pull(hg, allev, T, lif, B, cap).run(capture_raster=True, capture_v=True)
    for b in range(B):
        ref = O.forward(g, evs[b], T, lif)
        assert ref["spikes"] > 0
        assert np.array_equal(r.raster[b], ref["raster"]), b
        assert np.array_equal(r.final_v[b].view(np.uint32), ref["final_v"].view(np.uint32))


def test_batched_autotuner_searches_again_for_another_sample_count():
    """The batched search is cached per sample count: a second plan with the
    same count reuses it, a plan with another count searches again (the
    fastest strategy depends on how many samples share a step)."""
    g = G.hub_sheet_graph(60, 50, 40, hub_fraction=0.02, hub_min=40, hub_max=1500,
                          w_std=0.06, max_delay=20, seed=9)
    lif = K.lif_tau20()
    T = 40
    hg = HipGraph(g)
    _, ev4 = _drives(g.num_neurons, T, 4, seed=8)
    _, ev16 = _drives(g.num_neurons, T, 16, seed=8)
    ForwardPlan(hg, ev4, T, lif, Backend.HipAuto, num_samples=4).close()
    r1 = hg.tune_report(batched=True)
    ForwardPlan(hg, ev4, T, lif, Backend.HipAuto, num_samples=4).close()
    r2 = hg.tune_report(batched=True)
    assert [m for _, m in r2["candidates"]] == [m for _, m in r1["candidates"]]
    assert r2["search_ms"] == r1["search_ms"]
    ForwardPlan(hg, ev16, T, lif, Backend.HipAuto, num_samples=16).close()
    r3 = hg.tune_report(batched=True)
    assert r3["search_ms"] != r1["search_ms"]
    assert [m for _, m in r3["candidates"]] != [m for _, m in r1["candidates"]]


def test_batched_exact_autotuner_considers_pull():
    """hip-auto-exact on a batched plan searches hip-bucket and hip-pull (both
    bit-exact), never hip-atomic; whichever it picks matches the oracle bit
    for bit."""
    g = G.hub_sheet_graph(60, 50, 40, hub_fraction=0.02, hub_min=40, hub_max=1500,
                          w_std=0.06, max_delay=20, seed=9)
    lif = K.lif_tau20()
    T, B = 60, 16
    evs, allev = _drives(g.num_neurons, T, B, seed=6)
    hg = HipGraph(g)
    plan = ForwardPlan(hg, allev, T, lif, Backend.HipAutoExact, num_samples=B)
    rep = hg.tune_report(exact_only=True, batched=True)
    variants = {c.variant for c, _ in rep["candidates"]}
    assert Backend.HipPull in variants and Backend.HipAtomic not in variants
    assert plan.variant in (Backend.HipPull, Backend.HipBucket)
    r = plan.run(capture_raster=True, capture_v=True)
    for b in range(B):
        ref = O.forward(g, evs[b], T, lif)
        assert np.array_equal(r.raster[b], ref["raster"])
        assert np.array_equal(r.final_v[b].view(np.uint32), ref["final_v"].view(np.uint32))
