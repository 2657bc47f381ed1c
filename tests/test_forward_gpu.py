"""GPU parity of the forward path through the C ABI (libsnc_hip.so).

hip-bucket and hip-sort: rasters, totals, logits and final membrane voltages
bit-exact against the CPU oracle and the reference's golden fixtures.
hip-atomic: identical spike totals and per-neuron counts, voltages within
1e-5 relative (absolute floor 1e-6) -- float atomics reorder the sums.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, load_graph_fixture, unpack_raster
from helpers import events, make_graph, random_events, random_graph
from snc_amd import configs as K
from snc_amd import graph as G
from snc_amd.runtime import (Backend, ForwardPlan, HipGraph, LIFParams, PlanConfig, available,
                             forward)

pytestmark = pytest.mark.gpu

EXACT = [Backend.HipBucket, Backend.HipSort]
V_RTOL, V_ATOL = 1e-5, 1e-6


def test_device_available():
    assert available(), "gfx950 device / libsnc_hip.so not usable"


def _check_exact(res, ref, n):
    assert res.stats.spikes == ref["spikes"]
    assert res.stats.synaptic_events == ref["events"]
    assert np.array_equal(res.logits, ref["logits"])
    assert np.array_equal(res.neuron_spikes, ref["neuron_spikes"])
    assert np.array_equal(res.raster, ref["raster"])
    assert np.array_equal(res.final_v.view(np.uint32), ref["final_v"].view(np.uint32))


def _check_atomic(res, ref):
    assert res.stats.spikes == ref["spikes"]
    assert res.stats.synaptic_events == ref["events"]
    assert np.array_equal(res.neuron_spikes, ref["neuron_spikes"])
    np.testing.assert_allclose(res.final_v, ref["final_v"], rtol=V_RTOL, atol=V_ATOL)


@pytest.mark.parametrize("backend", EXACT)
@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "fwd_*.npz"))))
def test_exact_backends_match_reference_goldens(path, backend):
    g, lif, z = load_graph_fixture(path)
    T = int(z["T"])
    hg = HipGraph(g)
    for s in range(int(z["n_samples"])):
        plan = ForwardPlan(hg, z[f"events_{s}"], T, lif, backend)
        r = plan.run(capture_raster=True, capture_v=True)
        assert r.used_backend == backend
        assert np.array_equal(r.logits, z[f"logits_{s}"])
        assert r.stats.spikes == int(z[f"spikes_{s}"])
        assert r.stats.synaptic_events == int(z[f"synaptic_events_{s}"])
        assert np.array_equal(r.step_spikes, z[f"step_spikes_{s}"])
        ids = z[f"raster_ids_{s}"]
        want = np.unpackbits(z[f"raster_{s}"], axis=1, bitorder="little")[:, : ids.size]
        assert np.array_equal(unpack_raster(r.raster, g.num_neurons)[:, ids], want)
        ref = O.forward(g, z[f"events_{s}"], T, lif)
        _check_exact(r, ref, g.num_neurons)
        plan.close()


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "fwd_*.npz"))))
def test_atomic_matches_reference_goldens(path):
    g, lif, z = load_graph_fixture(path)
    T = int(z["T"])
    hg = HipGraph(g)
    for s in range(int(z["n_samples"])):
        r = ForwardPlan(hg, z[f"events_{s}"], T, lif, Backend.HipAtomic).run(capture_v=True)
        assert r.stats.spikes == int(z[f"spikes_{s}"])
        assert np.array_equal(r.logits, z[f"logits_{s}"])
        _check_atomic(r, O.forward(g, z[f"events_{s}"], T, lif))


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("backend", EXACT + [Backend.HipAtomic])
def test_random_graphs(seed, backend):
    rng = np.random.default_rng(100 + seed)
    g = random_graph(rng, 100 + 37 * seed, 1500 + 300 * seed, max_delay=1 + 4 * seed)
    T = 60
    ev = random_events(rng, g, T)
    lif = LIFParams(1.0, 0.9, 0.0, seed % 3, reset_mode=seed % 2)
    r = forward(g, ev, T, lif, backend, capture_raster=True, capture_v=True)
    ref = O.forward(g, ev, T, lif)
    if backend == Backend.HipAtomic:
        _check_atomic(r, ref)
    else:
        _check_exact(r, ref, g.num_neurons)


@pytest.mark.parametrize("backend", EXACT)
def test_fold_order_sensitive_sum(backend):
    g = make_graph(4, [(0, 3, 1e8, 1), (1, 3, 1.0, 1), (2, 3, -1e8, 1)], roles=[1, 1, 1, 0],
                   channels=[0, 0, 0, -1])
    r = forward(g, events([(0, 0, 1.0)]), 4, LIFParams(0.5, 1.0, 0.0, 1), backend)
    assert r.neuron_spikes.tolist() == [1, 1, 1, 0]


@pytest.mark.parametrize("backend", EXACT + [Backend.HipAtomic])
def test_heavy_targets(backend):
    # Two targets with in-degree 3000 and 700 (beyond the one-lane fold) fed by
    # 3000 inputs with mixed delays; plus ordinary traffic.
    rng = np.random.default_rng(7)
    n_in = 3000
    edges = []
    for a in range(n_in):
        edges.append((a, n_in, float(rng.normal(0, 0.05)), int(rng.integers(1, 40))))
        if a < 700:
            edges.append((a, n_in + 1, float(rng.normal(0, 0.08)), int(rng.integers(1, 5))))
        edges.append((a, n_in + 2 + a % 50, float(rng.normal(0, 0.3)), 1))
    n = n_in + 52
    roles = [1] * n_in + [0] * 52
    chans = list(range(n_in)) + [-1] * 52
    g = make_graph(n, edges, roles, chans)
    T = 80
    ev = random_events(rng, g, T, rate=0.05, amp=(0.9, 1.2))
    lif = LIFParams(1.0, 0.95, 0.0, 1)
    r = forward(g, ev, T, lif, backend, capture_raster=True, capture_v=True)
    ref = O.forward(g, ev, T, lif)
    if backend == Backend.HipAtomic:
        _check_atomic(r, ref)
    else:
        _check_exact(r, ref, n)


@pytest.mark.parametrize("backend", EXACT + [Backend.HipAtomic])
def test_edge_cases(backend):
    lif = LIFParams()
    g = make_graph(70, [(0, 1, 1.0, 1)], roles=[1] + [0] * 69, channels=[0] + [-1] * 69)
    # no events
    r = forward(g, events([]), 10, lif, backend, capture_raster=True, capture_v=True)
    assert r.stats.spikes == 0 and not r.raster.any()
    # zero steps
    r = forward(g, events([(0, 0, 2.0)]), 0, lif, backend)
    assert r.stats.spikes == 0
    # one step, events outside the window and on unknown channels are ignored
    r = forward(g, events([(0, 0, 2.0), (0, 5, 2.0), (3, 0, 2.0), (-1, 0, 2.0)]), 1, lif,
                backend, capture_raster=True)
    assert r.stats.spikes == 1 and r.stats.synaptic_events == 1
    # graph without synapses
    g0 = make_graph(5, [], roles=[1] * 5, channels=[0, 1, 2, 3, 4])
    r = forward(g0, events([(c, 2, 1.5) for c in range(5)]), 4, lif, backend)
    assert r.stats.spikes == 5 and r.stats.synaptic_events == 0


def test_invalid_graph_is_rejected():
    from snc_amd._lib import HipError
    g = make_graph(3, [(0, 1, 0.5, 1)])
    g.delays[0] = 0
    with pytest.raises(HipError):
        HipGraph(g)


@pytest.mark.parametrize("backend", EXACT + [Backend.HipAtomic])
def test_c0_workload_all_samples(backend):
    """The plain-forward workload: 64 samples, 50 steps, subtract reset."""
    w = K.c0(n_samples=64)
    hg = HipGraph(w.graph)
    for s, ev in enumerate(w.samples):
        r = ForwardPlan(hg, ev, w.num_steps, w.lif, backend).run(capture_raster=True,
                                                                capture_v=True)
        ref = O.forward(w.graph, ev, w.num_steps, w.lif)
        if backend == Backend.HipAtomic:
            _check_atomic(r, ref)
        else:
            _check_exact(r, ref, w.graph.num_neurons)


@pytest.mark.parametrize("backend", EXACT + [Backend.HipAtomic])
def test_c1_shaped_reduced(backend):
    """Hub sheet with the delivery-stress structure at a size the oracle
    finishes in seconds: 120 x 100 sheet, out-degree 60, hubs up to 3000."""
    g = G.hub_sheet_graph(120, 100, 60, hub_fraction=0.01, hub_min=60, hub_max=3000,
                          w_std=0.05, max_delay=32, seed=5)
    from snc_amd import inputs as I
    ev = I.background_drive(g.num_neurons, 300, 0.01, 1.0, seed=6)
    lif = K.lif_tau20()
    hg = HipGraph(g)
    plan = ForwardPlan(hg, ev, 300, lif, backend)
    r = plan.run(capture_raster=True, capture_v=True)
    ref = O.forward(g, ev, 300, lif)
    if backend == Backend.HipAtomic:
        _check_atomic(r, ref)
    else:
        _check_exact(r, ref, g.num_neurons)
    # rerunning the plan starts from a clean state
    r2 = plan.run(capture_raster=True, capture_v=True)
    assert r2.stats.spikes == r.stats.spikes
    if backend != Backend.HipAtomic:
        assert np.array_equal(r2.raster, r.raster)


def _heavy_graph(seed=7):
    """3000 inputs onto two heavy targets (in-degree 3000 and 700) with mixed
    delays, plus 50 light targets."""
    rng = np.random.default_rng(seed)
    n_in = 3000
    edges = []
    for a in range(n_in):
        edges.append((a, n_in, float(rng.normal(0, 0.05)), int(rng.integers(1, 40))))
        if a < 700:
            edges.append((a, n_in + 1, float(rng.normal(0, 0.08)), int(rng.integers(1, 5))))
        if a % 3 == 0:
            edges.append((a, n_in + 2 + a % 50, float(rng.normal(0, 0.3)), int(rng.integers(1, 3))))
    n = n_in + 52
    g = make_graph(n, edges, [1] * n_in + [0] * 52, list(range(n_in)) + [-1] * 52)
    return g, random_events(rng, g, 80, rate=0.05, amp=(0.9, 1.2))


@pytest.mark.parametrize("block", [128, 256, 512])
@pytest.mark.parametrize("threshold", [64, 200, 512])
def test_bucket_launch_configs_are_bit_exact(block, threshold):
    """Every block shape and hub threshold the autotuner may pick keeps the
    bucket strategy bit-exact (thresholds below the light limit move the
    70-180 in-degree targets of the reduced delivery-stress graph and the
    heavy targets onto the hub kernel)."""
    g = G.hub_sheet_graph(60, 50, 40, hub_fraction=0.02, hub_min=40, hub_max=1500,
                          w_std=0.06, max_delay=20, seed=9)
    from snc_amd import inputs as I
    ev = I.background_drive(g.num_neurons, 120, 0.02, 1.0, seed=3)
    lif = K.lif_tau20()
    cfg = PlanConfig(Backend.HipBucket, block, threshold, False)
    r = ForwardPlan(HipGraph(g), ev, 120, lif, config=cfg).run(capture_raster=True, capture_v=True)
    _check_exact(r, O.forward(g, ev, 120, lif), g.num_neurons)
    g2, ev2 = _heavy_graph()
    lif2 = LIFParams(1.0, 0.95, 0.0, 1)
    r2 = ForwardPlan(HipGraph(g2), ev2, 80, lif2, config=cfg).run(capture_raster=True,
                                                                  capture_v=True)
    _check_exact(r2, O.forward(g2, ev2, 80, lif2), g2.num_neurons)


@pytest.mark.parametrize("block", [128, 512])
def test_atomic_block_shapes(block):
    g2, ev2 = _heavy_graph(8)
    lif2 = LIFParams(1.0, 0.95, 0.0, 1)
    cfg = PlanConfig(Backend.HipAtomic, block, 512, False)
    r = ForwardPlan(HipGraph(g2), ev2, 80, lif2, config=cfg).run(capture_v=True)
    _check_atomic(r, O.forward(g2, ev2, 80, lif2))


def test_invalid_plan_config_is_rejected():
    from snc_amd._lib import HipError
    g, ev = _heavy_graph()
    hg = HipGraph(g)
    for cfg in (PlanConfig(Backend.HipBucket, 192, 512), PlanConfig(Backend.HipBucket, 256, 600),
                PlanConfig(Backend.HipBucket, 256, 0), PlanConfig(Backend.HipAuto, 256, 512)):
        with pytest.raises(HipError):
            ForwardPlan(hg, ev, 10, LIFParams(), config=cfg)


def test_autotuner_search_picks_fastest_and_stays_exact():
    g = G.hub_sheet_graph(120, 100, 60, hub_fraction=0.01, hub_min=60, hub_max=3000,
                          w_std=0.05, max_delay=32, seed=5)
    from snc_amd import inputs as I
    ev = I.background_drive(g.num_neurons, 200, 0.01, 1.0, seed=6)
    lif = K.lif_tau20()
    hg = HipGraph(g)
    ref = O.forward(g, ev, 200, lif)
    plan = ForwardPlan(hg, ev, 200, lif, Backend.HipAutoExact)
    assert plan.variant in (Backend.HipBucket, Backend.HipSort)
    _check_exact(plan.run(capture_raster=True, capture_v=True), ref, g.num_neurons)
    rep = hg.tune_report(exact_only=True)
    ms = [m for _, m in rep["candidates"]]
    assert len(ms) >= 4 and rep["trial_steps"] == 128
    assert all(c.variant != Backend.HipAtomic for c, _ in rep["candidates"])
    assert not any(c.persistent for c, _ in rep["candidates"])  # forced plans only
    assert rep["failed"] == []
    ran = [m for m in ms if m >= 0]
    assert ms[rep["chosen"]] == min(ran)
    assert plan.config == rep["candidates"][rep["chosen"]][0]
    # cached: a second plan reuses the search
    plan2 = ForwardPlan(hg, ev, 200, lif, Backend.HipAutoExact)
    assert plan2.config == plan.config
    assert hg.tune_report(exact_only=True)["search_ms"] == rep["search_ms"]
    # the unrestricted search may pick hip-atomic; its result meets that bar
    p3 = ForwardPlan(hg, ev, 200, lif, Backend.HipAuto)
    r3 = p3.run(capture_raster=True, capture_v=True)
    if p3.variant == Backend.HipAtomic:
        _check_atomic(r3, ref)
    else:
        _check_exact(r3, ref, g.num_neurons)


@pytest.mark.parametrize("backend", EXACT + [Backend.HipAtomic])
def test_mega_hub_multi_round_fold(backend):
    """A hub with in-degree 70,000 (2188 column words: two load rounds, eight
    words per thread in the first) and a second with 20,000, where every
    input fires in the same step (70,000 set bits: nine LDS passes of the
    ordered fold), with delays 1..20 so the entries spread over ring slots."""
    rng = np.random.default_rng(11)
    n_in = 70000
    edges = []
    for a in range(n_in):
        edges.append((a, n_in, float(rng.normal(0, 0.01)), int(rng.integers(1, 21))))
        if a % 7 < 2:
            edges.append((a, n_in + 1, float(rng.normal(0, 0.02)), int(rng.integers(1, 4))))
    g = make_graph(n_in + 2, edges, [1] * n_in + [0, 0], list(range(n_in)) + [-1, -1])
    ev = [(c, 0, 1.5) for c in range(n_in)]
    ev += [(int(c), 3, 1.5) for c in rng.choice(n_in, 20000, replace=False)]
    lif = LIFParams(1.0, 0.95, 0.0, 1)
    r = forward(g, events(ev), 30, lif, backend, capture_raster=True, capture_v=True)
    ref = O.forward(g, events(ev), 30, lif)
    assert ref["spikes"] >= 90000
    if backend == Backend.HipAtomic:
        _check_atomic(r, ref)
    else:
        _check_exact(r, ref, g.num_neurons)


@pytest.mark.parametrize("cfg", [PlanConfig(Backend.HipBucket, 256, 512),
                                 PlanConfig(Backend.HipBucket, 128, 64),
                                 PlanConfig(Backend.HipAtomic, 512, 512),
                                 None])
def test_batched_forward_equals_one_sample_at_a_time(cfg):
    """Five samples with different drives in one batched plan (grid.y =
    sample): each sample's raster, counts, logits and membranes equal the
    oracle's single-sample forward (atomic: counts, voltages within 1e-5);
    cfg None takes the autotuner's batched search."""
    g = G.hub_sheet_graph(60, 50, 40, hub_fraction=0.02, hub_min=40, hub_max=1500,
                          w_std=0.06, max_delay=20, seed=9)
    from snc_amd import inputs as I
    lif = K.lif_tau20()
    T, B = 90, 5
    evs = [I.background_drive(g.num_neurons, T, 0.01 + 0.004 * b, 1.0, seed=20 + b)
           for b in range(B)]
    allev = np.concatenate([e.copy() for e in evs])
    off = 0
    for b, e in enumerate(evs):
        allev["sample_id"][off:off + e.size] = b
        off += e.size
    hg = HipGraph(g)
    plan = ForwardPlan(hg, allev, T, lif, Backend.HipAuto, config=cfg, num_samples=B)
    r = plan.run(capture_raster=True, capture_v=True)
    sp, ev = plan.sample_stats()
    assert r.raster.shape == (B, T, (g.num_neurons + 31) // 32)
    assert sp.sum() == r.stats.spikes and ev.sum() == r.stats.synaptic_events
    variant = plan.variant
    for b in range(B):
        ref = O.forward(g, evs[b], T, lif)
        assert sp[b] == ref["spikes"] and ev[b] == ref["events"]
        assert np.array_equal(r.neuron_spikes[b], ref["neuron_spikes"])
        if variant == Backend.HipAtomic:
            np.testing.assert_allclose(r.final_v[b], ref["final_v"], rtol=V_RTOL, atol=V_ATOL)
        else:
            assert np.array_equal(r.raster[b], ref["raster"])
            assert np.array_equal(r.final_v[b].view(np.uint32), ref["final_v"].view(np.uint32))
            assert np.array_equal(r.logits[b], ref["logits"])
    if cfg is None:
        rep = hg.tune_report(batched=True)
        assert all(c.variant != Backend.HipSort and not c.persistent for c, _ in rep["candidates"])
