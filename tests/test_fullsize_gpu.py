"""Oracle parity of the benchmarked workloads at their own sizes.

The delivery-stress forward (200,000 neurons, 20,000,000 synapses, 1000
steps, batch 1) runs on every hip strategy and on forced persistent plans,
and through the one-shot snc_hip_forward; one full minibatch of each e-prop
workload (784-512r-10 with B = 256, T_b in [50, 200]; 700-1024r-20 with
B = 128, delays 1..31) runs on the batched and the plain paths. Each result
is compared with the CPU oracle on the same inputs:

  hip-bucket, hip-sort   raster, per-step spikes, per-neuron counts, logits
                         and final membranes bit for bit
  hip-atomic             identical spike totals and per-neuron counts;
                         membranes within 1e-5 relative (absolute 1e-6)
  e-prop                 logits and rasters bit for bit, spike and event
                         totals equal, weight deltas within 1e-5 relative
                         (absolute 6e-8, ~2 ulp of a weight: the softmax
                         uses the device expf)
"""
import numpy as np
import pytest

import oracle as O
from snc_amd import configs as K
from snc_amd.runtime import (Backend, ForwardPlan, HipGraph, PlanConfig, forward_oneshot)
from snc_amd.train import BATCHED, ELIG_F32, ELIG_F64, PLAIN, SERIAL, HipTrainSession

pytestmark = pytest.mark.gpu

V_RTOL, V_ATOL = 1e-5, 1e-6
W_RTOL, W_ATOL = 1e-5, 6e-8


def _check(r, ref, exact):
    assert r.stats.spikes == ref["spikes"]
    assert r.stats.synaptic_events == ref["events"]
    assert np.array_equal(r.neuron_spikes, ref["neuron_spikes"])
    if exact:
        assert np.array_equal(r.raster, ref["raster"])
        assert np.array_equal(r.step_spikes, ref["step_spikes"])
        assert np.array_equal(r.logits, ref["logits"])
        assert np.array_equal(r.final_v.view(np.uint32), ref["final_v"].view(np.uint32))
    else:
        np.testing.assert_allclose(r.final_v, ref["final_v"], rtol=V_RTOL, atol=V_ATOL)


@pytest.fixture(scope="module")
def c1():
    w = K.c1()
    g, ev = w.graph, w.samples[0]
    assert g.num_neurons == 200_000 and g.num_synapses == 20_000_000
    ref = O.forward(g, ev, w.num_steps, w.lif)
    assert ref["spikes"] > 1_000_000 and ref["events"] > 100_000_000
    print(f"c1 oracle: {ref['spikes']} spikes, {ref['events']} events", flush=True)
    hg = HipGraph(g)
    yield w, hg, ref
    hg.close()


@pytest.mark.parametrize("backend", [Backend.HipBucket, Backend.HipSort, Backend.HipAtomic,
                                     Backend.HipAuto, Backend.HipAutoExact])
def test_c1_full_size(c1, backend):
    w, hg, ref = c1
    print(f"c1 full size, {backend.name}", flush=True)
    plan = ForwardPlan(hg, w.samples[0], w.num_steps, w.lif, backend)
    r = plan.run(capture_raster=True, capture_v=True)
    exact = plan.variant in (Backend.HipBucket, Backend.HipSort)
    if backend == Backend.HipAutoExact:
        assert exact
    _check(r, ref, exact)
    plan.close()


def test_c1_full_size_pull_64_samples(c1):
    """64 independent drives of the 20M-synapse graph in one hip-pull plan
    (64-bit fired words, default light cap, 1947 hubs of up to 32k inputs in
    play): every checked sample has the oracle's raster, per-neuron counts,
    totals, logits and final membranes bit for bit; the rest of the batch
    shares the launches."""
    from snc_amd import inputs as I
    w, hg, ref0 = c1
    B = 64
    evs = [w.samples[0]] + [I.background_drive(w.graph.num_neurons, w.num_steps, 0.01, 1.0,
                                               seed=100 + b) for b in range(1, B)]
    allev = np.concatenate([e.copy() for e in evs])
    off = 0
    for b, e in enumerate(evs):
        allev["sample_id"][off:off + e.size] = b
        off += e.size
    print("c1 full size, hip-pull x64", flush=True)
    plan = ForwardPlan(hg, allev, w.num_steps, w.lif, Backend.HipPull, num_samples=B)
    r = plan.run(capture_raster=True, capture_v=True)
    sp, ev = plan.sample_stats()
    for b in (0, 1, 33, 63):
        ref = ref0 if b == 0 else O.forward(w.graph, evs[b], w.num_steps, w.lif)
        assert sp[b] == ref["spikes"] and ev[b] == ref["events"], b
        assert np.array_equal(r.neuron_spikes[b], ref["neuron_spikes"]), b
        assert np.array_equal(r.raster[b], ref["raster"]), b
        assert np.array_equal(r.logits[b], ref["logits"]), b
        assert np.array_equal(r.final_v[b].view(np.uint32), ref["final_v"].view(np.uint32)), b
    plan.close()


@pytest.mark.parametrize("cfg", [PlanConfig(Backend.HipBucket, 256, 512, True),
                                 PlanConfig(Backend.HipBucket, 512, 128, True),
                                 PlanConfig(Backend.HipAtomic, 512, 512, True),
                                 PlanConfig(Backend.HipAtomic, 256, 512, True)])
def test_c1_forced_persistent(c1, cfg):
    """The cooperative whole-run kernel, forced (the autotuner may pick it)."""
    w, hg, ref = c1
    print(f"c1 persistent {cfg}", flush=True)
    plan = ForwardPlan(hg, w.samples[0], w.num_steps, w.lif, config=cfg)
    assert plan.config.persistent
    r = plan.run(capture_raster=True, capture_v=True)
    _check(r, ref, cfg.variant == Backend.HipBucket)
    # a second run of the same plan starts from a clean state
    r2 = plan.run(capture_raster=True, capture_v=True)
    _check(r2, ref, cfg.variant == Backend.HipBucket)
    plan.close()


@pytest.mark.parametrize("backend", [Backend.HipBucket, Backend.HipAtomic])
def test_forced_persistent_plain_forward_all_samples(backend):
    """Persistent plans on the 64 samples of the plain-forward workload
    (784 -> 256 -> 10, delay 1, subtract reset)."""
    w = K.c0(n_samples=64)
    hg = HipGraph(w.graph)
    cfg = PlanConfig(backend, 256, 512, True)
    for ev in w.samples:
        r = ForwardPlan(hg, ev, w.num_steps, w.lif, config=cfg).run(capture_raster=True,
                                                                    capture_v=True)
        _check(r, O.forward(w.graph, ev, w.num_steps, w.lif), backend == Backend.HipBucket)
    hg.close()


@pytest.mark.parametrize("backend", [Backend.HipBucket, Backend.HipAtomic, Backend.HipSort,
                                     Backend.HipAutoExact])
def test_oneshot_forward_plain_workload(backend):
    """snc_hip_forward (upload, run, read back, free in one call) on the
    plain-forward workload's first 8 samples."""
    w = K.c0(n_samples=8)
    for ev in w.samples:
        r = forward_oneshot(w.graph, ev, w.num_steps, w.lif, backend, capture_raster=True,
                            capture_v=True)
        ref = O.forward(w.graph, ev, w.num_steps, w.lif)
        _check(r, ref, backend != Backend.HipAtomic)
        assert r.used_backend != Backend.HipAutoExact


def test_oneshot_forward_c1(c1):
    w, _, ref = c1
    r = forward_oneshot(w.graph, w.samples[0], w.num_steps, w.lif, Backend.HipBucket,
                        capture_raster=True, capture_v=True)
    assert r.used_backend == Backend.HipBucket
    _check(r, ref, True)


def _oracle_trainer(w):
    c = w.cfg
    ocfg = O.train_config(num_steps=c.num_steps, lif=c.lif, lr=c.lr, w_init=c.w_init,
                          w_max=c.w_max, gamma=c.surrogate_scale, feedback=c.feedback_scale,
                          train_hidden=c.train_hidden, readout_mode=c.readout_mode,
                          readout_psi=c.readout_psi, encoder=c.encoder, gain=c.enc_gain,
                          max_rate=c.enc_max_rate, enc_threshold=c.enc_threshold, seed=c.seed)
    tr = O.Trainer(w.graph, ocfg)
    tr.weights = w.init_weights
    return tr


def _minibatch_events(w, idx, epoch):
    """The minibatch's inputs as the oracle takes them (samples 0..B-1): the
    device's Philox Poisson spikes of dataset sample idx[b] at this epoch, or
    the workload's event streams."""
    c = w.cfg
    evs = []
    for b, i in enumerate(idx):
        if w.kind == "features":
            e = O.encode_philox(w.X[i], int(w.lengths[i]), gain=c.enc_gain,
                                max_rate=c.enc_max_rate, threshold=c.enc_threshold, seed=c.seed,
                                sample=int(i), epoch=epoch)
        else:
            e = w.events[w.events["sample_id"] == i].copy()
        e["sample_id"] = b
        evs.append(e)
    return np.concatenate(evs)


@pytest.fixture(scope="module")
def eprop_refs():
    """One full minibatch of each workload on the oracle (the training split's
    samples 5B..6B-1, epoch 1), computed once."""
    out = {}
    for name, make in (("c2", K.c2), ("c3", K.c3)):
        w = make()
        idx = np.arange(5 * w.batch, 6 * w.batch, dtype=np.int32)
        ev = _minibatch_events(w, idx, 1)
        tr = _oracle_trainer(w)
        ref = tr.minibatch(ev, w.lengths[idx], w.labels[idx], apply=True, raster=True)
        ref["weights"] = tr.weights.copy()
        print(f"{name} oracle minibatch: {ref['spikes']} spikes, {ref['events']} events",
              flush=True)
        out[name] = (w, idx, ref)
    return out


@pytest.mark.parametrize("name", ["c2", "c3"])
@pytest.mark.parametrize("mode,elig", [(BATCHED, ELIG_F64), (PLAIN, ELIG_F64),
                                       (SERIAL, ELIG_F64), (BATCHED, ELIG_F32)])
def test_eprop_full_minibatch(eprop_refs, name, mode, elig):
    w, idx, ref = eprop_refs[name]
    print(f"{name} full minibatch, mode {mode}, elig {elig}", flush=True)
    s = HipTrainSession(w.graph, w.cfg, batch=w.batch, max_steps=w.max_steps, elig=elig)
    s.set_weights(w.init_weights)
    if w.kind == "features":
        s.upload_features(w.X, w.labels, w.lengths)
    else:
        s.upload_events(w.events, w.labels.size, w.labels, w.lengths)
    st, logits, raster = s.step(idx, epoch=1, mode=mode, apply=True, want_logits=True,
                                want_raster=True)
    assert st.samples == w.batch
    assert np.array_equal(logits, ref["logits"])
    tm = int(w.lengths[idx].max())
    assert np.array_equal(raster[:, :tm], ref["raster"])
    assert st.spikes == ref["spikes"] and st.synaptic_events == ref["events"]
    w0 = w.init_weights
    np.testing.assert_allclose(s.get_weights() - w0, ref["weights"] - w0, rtol=W_RTOL,
                               atol=W_ATOL)
    s.close()
