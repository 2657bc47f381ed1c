"""GPU e-prop parity through the C ABI (snc_hip_train_*).

Against the reference's own compiled snc::Trainer (one-sample steps, both
learning rules), and against the CPU oracle's minibatch step from the same
weights and inputs:
rasters and logits bit-exact, per-neuron fires identical, and FP32 weight
deltas within 1e-5 relative (absolute floor 6e-8, ~2 ulp of the weights:
softmax uses the device expf). The plain per-sample path and the batched
persistent path compute the same sums in the same order, so they agree bit
for bit with each other.
"""
import numpy as np
import pytest

import oracle as O
from snc_amd import graph as G
from snc_amd import inputs as I
from snc_amd.runtime import LIFParams
from snc_amd.train import (BATCHED, ELIG_F32, ELIG_F64, PLAIN, READOUT_LEAKY, READOUT_SPIKING,
                           SERIAL, HipTrainSession, TrainConfig)

pytestmark = pytest.mark.gpu

W_RTOL, W_ATOL = 1e-5, 6e-8


def _cfgs(T, readout=READOUT_SPIKING, reset_mode=0, encoder=0, lr=0.2, w_init=0.4, seed=3,
          decay=0.9, max_rate=0.5, reward=False, reward_decay=0.99):
    lif = LIFParams(1.0, decay, 0.0, 1, reset_mode)
    gpu = TrainConfig(num_steps=T, lif=lif, lr=lr, w_init=w_init, w_max=4.0, surrogate_scale=0.3,
                      feedback_scale=1.0, readout_mode=readout, readout_psi=1.0, encoder=encoder,
                      enc_max_rate=max_rate, seed=seed, reward_mode=reward,
                      reward_baseline_decay=reward_decay)
    orc = O.train_config(num_steps=T, lif=lif, lr=lr, w_init=w_init, w_max=4.0, gamma=0.3,
                         feedback=1.0, readout_mode=readout, readout_psi=1.0, encoder=encoder,
                         max_rate=max_rate, seed=seed, reward_mode=reward,
                         reward_baseline_decay=reward_decay)
    return gpu, orc


def _events_for(X, lengths, enc=0, max_rate=0.5, seed=3, epoch=0, ids=None):
    evs = []
    for b in range(len(lengths)):
        sid = b if ids is None else int(ids[b])
        if enc in (0, 2):
            e = O.encode(enc, X[b], int(lengths[b]))
        else:
            e = O.encode_philox(X[b], int(lengths[b]), max_rate=max_rate, seed=seed, sample=sid,
                                epoch=epoch)
        e["sample_id"] = b
        evs.append(e)
    return np.concatenate(evs)


def _compare(sess, orc_tr, w0, idx, ev_orc, lengths, labels, mode, want_exact_w=False):
    st, logits, raster = sess.step(idx, epoch=0, mode=mode, apply=True, want_logits=True,
                                   want_raster=True)
    ref = orc_tr.minibatch(ev_orc, lengths, labels, apply=True, raster=True)
    assert np.array_equal(logits, ref["logits"])
    Tm = int(lengths.max())
    assert np.array_equal(raster[:, :Tm], ref["raster"])
    assert st.spikes == ref["spikes"] and st.synaptic_events == ref["events"]
    wg, wo = sess.get_weights(), orc_tr.weights
    if want_exact_w:
        assert np.array_equal(wg, wo)
    np.testing.assert_allclose(wg - w0, wo - w0, rtol=W_RTOL, atol=W_ATOL)
    return wg


@pytest.mark.parametrize("mode", [BATCHED, PLAIN, SERIAL])
@pytest.mark.parametrize("elig", [ELIG_F64, ELIG_F32])
def test_minibatch_step_matches_oracle(mode, elig):
    g = G.static_snc([64, 48, 8], 1500, 0.5, 1, 0.2, 4)
    B = 12
    X, y = O.ref_make_synthetic(B, 64, 8, 0.35, 3)
    lengths = np.array([20, 17, 25, 20, 9, 31, 20, 22, 14, 20, 28, 19], np.int32)
    gcfg, ocfg = _cfgs(32)
    sess = HipTrainSession(g, gcfg, batch=B, max_steps=32, elig=elig)
    orc = O.Trainer(g, ocfg)
    w0 = orc.weights
    assert np.array_equal(sess.get_weights(), w0)  # same mt19937_64 draws
    ev = _events_for(X, lengths)
    sess.upload_events(ev, B, y, lengths)
    _compare(sess, orc, w0, np.arange(B), ev, lengths, y, mode)
    syn, neur = sess.collect_stats()
    osyn, oneur = orc.stats()
    assert np.array_equal(neur, oneur) and np.array_equal(syn, osyn)


def test_plain_serial_and_batched_are_bit_identical():
    """The three modes on one minibatch: weights, logits, rasters and totals
    bit for bit (SERIAL launches the batched kernel once per sample and sums
    the same contribution rows in the same order)."""
    g = G.sheet_graph([100, 64, 10], 16, recurrent_fan_in=8, radius=3.0, sheet=10.0, w_std=0.3,
                      max_delay=5, seed=2)
    B = 16
    X, y = I.prototype_images(B, 100, 10, 0.6, 0.2, seed=4)
    Xf = X.astype(np.float32) / np.float32(255.0)
    lengths = np.random.default_rng(1).integers(10, 40, B).astype(np.int32)
    gcfg, _ = _cfgs(40, readout=READOUT_LEAKY, reset_mode=1, encoder=1, max_rate=0.4)
    out = []
    for mode in (PLAIN, BATCHED, SERIAL):
        s = HipTrainSession(g, gcfg, batch=B, max_steps=40)
        s.upload_features(Xf, y, lengths)
        st, lg, ras = s.step(np.arange(B), epoch=2, mode=mode, want_logits=True, want_raster=True)
        out.append((s.get_weights(), lg, ras, st.spikes, st.synaptic_events))
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert np.array_equal(a, b)


def test_plain_graph_replay_equals_batched_over_minibatches():
    """Without rasters the plain path replays one captured hipGraph per sample
    length; over three minibatches (lengths repeat, so graphs are reused with
    new samples and epochs) its weights stay bit-identical to the batched
    path's, and its evaluation matches."""
    g = G.sheet_graph([100, 64, 10], 16, recurrent_fan_in=8, radius=3.0, sheet=10.0, w_std=0.3,
                      max_delay=5, seed=3)
    n = 48
    X, y = I.prototype_images(n, 100, 10, 0.6, 0.2, seed=5)
    Xf = X.astype(np.float32) / np.float32(255.0)
    lengths = np.random.default_rng(2).choice([12, 20, 27], n).astype(np.int32)
    gcfg, _ = _cfgs(30, readout=READOUT_LEAKY, reset_mode=1, encoder=1, max_rate=0.4)
    res = []
    for mode in (PLAIN, BATCHED):
        s = HipTrainSession(g, gcfg, batch=16, max_steps=30)
        s.upload_features(Xf, y, lengths)
        for k, ep in enumerate((1, 1, 2)):
            s.step(np.arange(16 * k, 16 * k + 16), epoch=ep, mode=mode)
        res.append((s.get_weights(), s.evaluate(mode=mode)))
    assert np.array_equal(res[0][0], res[1][0])
    assert res[0][1] == res[1][1]


@pytest.mark.parametrize("readout,reset_mode", [(READOUT_LEAKY, 1), (READOUT_SPIKING, 1),
                                                 (READOUT_LEAKY, 0)])
def test_device_poisson_and_extensions_match_oracle(readout, reset_mode):
    """Philox encoding on the device, leaky readout, subtract reset, uneven T."""
    g = G.sheet_graph([196, 96, 10], 24, recurrent_fan_in=12, radius=2.5, sheet=14.0,
                      w_std=0.3, delay=1, seed=7)
    B = 10
    Xu, y = I.prototype_images(B, 196, 10, 0.6, 0.2, seed=9)
    X = Xu.astype(np.float32) / np.float32(255.0)
    lengths = np.array([30, 45, 12, 50, 33, 40, 27, 50, 18, 44], np.int32)
    ids = np.arange(100, 100 + B)  # dataset indices feed the Philox counters
    gcfg, ocfg = _cfgs(50, readout=readout, reset_mode=reset_mode, encoder=1, lr=0.5,
                       w_init=0.3, max_rate=0.3, decay=float(np.float32(np.exp(-1 / 20))))
    sess = HipTrainSession(g, gcfg, batch=B, max_steps=50)
    Xd = np.zeros((200, 196), np.float32)
    Xd[ids] = X
    yd = np.zeros(200, np.int32)
    yd[ids] = y
    Ld = np.full(200, 10, np.int32)
    Ld[ids] = lengths
    sess.upload_features(Xd, yd, Ld)
    orc = O.Trainer(g, ocfg)
    w0 = orc.weights
    ev = _events_for(X, lengths, enc=1, max_rate=0.3, seed=3, epoch=0, ids=ids)
    _compare(sess, orc, w0, ids, ev, lengths, y, BATCHED)


@pytest.mark.parametrize("mode", [BATCHED, PLAIN])
def test_device_latency_matches_oracle(mode):
    """Latency encoding on the device (one spike per channel at
    lround((T_b - 1)(1 - x))) against the reference encode() of each sample's
    own length, uneven T_b."""
    g = G.sheet_graph([196, 96, 10], 24, recurrent_fan_in=12, radius=2.5, sheet=14.0,
                      w_std=0.3, delay=1, seed=7)
    B = 10
    Xu, y = I.prototype_images(B, 196, 10, 0.6, 0.2, seed=11)
    X = Xu.astype(np.float32) / np.float32(255.0)
    lengths = np.array([30, 45, 12, 50, 33, 40, 27, 50, 18, 44], np.int32)
    gcfg, ocfg = _cfgs(50, encoder=2, lr=0.5, w_init=0.6)
    sess = HipTrainSession(g, gcfg, batch=B, max_steps=50)
    sess.upload_features(X, y, lengths)
    orc = O.Trainer(g, ocfg)
    w0 = orc.weights
    ev = _events_for(X, lengths, enc=2)
    assert ev.size == sum(int((X[b] > 0).sum()) for b in range(B))  # one spike per lit channel
    _compare(sess, orc, w0, np.arange(B), ev, lengths, y, mode)


@pytest.mark.parametrize("mode", [BATCHED, PLAIN, SERIAL])
@pytest.mark.parametrize("readout,elig", [(READOUT_SPIKING, ELIG_F64), (READOUT_LEAKY, ELIG_F64),
                                          (READOUT_LEAKY, ELIG_F32)])
def test_reward_mode_matches_oracle_over_minibatches(mode, readout, elig):
    """Three-factor reward rule (trainer.cpp:165-179): actions drawn in
    minibatch order from the session's action stream, the baseline carried
    across minibatches, and a step without update drawing nothing. Each
    minibatch starts both sides from the same weights; logits and rasters are
    bit-exact, weight deltas within the tolerance above."""
    g = G.static_snc([64, 48, 8], 1500, 0.5, 1, 0.2, 4)
    B, rounds = 6, 4
    X, y = O.ref_make_synthetic(B * rounds, 64, 8, 0.35, 7)
    lengths = np.random.default_rng(3).integers(12, 30, B * rounds).astype(np.int32)
    gcfg, ocfg = _cfgs(30, readout=readout, reset_mode=1, lr=0.8, w_init=0.4, reward=True,
                       reward_decay=0.8)
    sess = HipTrainSession(g, gcfg, batch=B, max_steps=30, elig=elig)
    orc = O.Trainer(g, ocfg)
    # the dataset keeps its own indices; the oracle takes each minibatch as samples 0..B-1
    sess.upload_events(_events_for(X, lengths), B * rounds, y, lengths)
    moved = 0.0
    for r in range(rounds):
        sl = slice(B * r, B * (r + 1))
        ev = _events_for(X[sl], lengths[sl])
        w0 = sess.get_weights()
        orc.weights = w0
        apply = r != 1  # minibatch 1 runs without an update: no action is drawn
        st, logits, raster = sess.step(np.arange(B * r, B * (r + 1)), mode=mode, apply=apply,
                                       want_logits=True, want_raster=True)
        ref = orc.minibatch(ev, lengths[sl], y[sl], apply=apply, raster=True)
        assert np.array_equal(logits, ref["logits"])
        Tm = int(lengths[sl].max())
        assert np.array_equal(raster[:, :Tm], ref["raster"])
        wg = sess.get_weights()
        np.testing.assert_allclose(wg - w0, orc.weights - w0, rtol=W_RTOL, atol=W_ATOL)
        if not apply:
            assert np.array_equal(wg, w0)
        moved = max(moved, float(np.abs(wg - w0).max()))
    assert moved > 1e-3  # the rule did move the weights


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not present")
@pytest.mark.parametrize("reward", [False, True])
@pytest.mark.parametrize("mode", [BATCHED, PLAIN])
def test_single_sample_steps_match_the_compiled_reference_trainer(reward, mode):
    """Batch-1 device steps against the reference's own snc::Trainer (compiled
    from its sources): Direct encoding (no random draws), spike-count readout,
    reset to value, fixed T. Each reference epoch holds one sample, so its
    order is fixed; the device runs that sample as a one-sample minibatch.
    Weights follow within the tolerance above after every sample, in the
    supervised rule and in the three-factor reward rule (action stream and
    baseline carried across samples)."""
    g = G.static_snc([40, 30, 20, 5], 900, 0.5, 2, 0.2, 3)
    X, y = O.ref_make_synthetic(8, 40, 5, 0.3, 3)
    T = 18
    gcfg, ocfg = _cfgs(T, encoder=0, lr=0.6, w_init=0.35, seed=3, decay=0.85, reward=reward,
                       reward_decay=0.8)
    sess = HipTrainSession(g, gcfg, batch=1, max_steps=T)
    sess.upload_features(X, y, np.full(8, T, np.int32))
    ref = O.RefTrainer(g, ocfg)
    start = ref.weights
    assert np.array_equal(sess.get_weights(), start)
    for i in range(8):
        w0 = ref.weights
        ref.train_epoch(X[i:i + 1], y[i:i + 1], i + 1, 5)
        sess.step(np.array([i]), epoch=i + 1, mode=mode, apply=True)
        np.testing.assert_allclose(sess.get_weights() - w0, ref.weights - w0, rtol=W_RTOL,
                                   atol=W_ATOL)
        sess.set_weights(ref.weights)  # each sample starts both sides from the same weights
    assert np.abs(ref.weights - start).max() > 1e-3


@pytest.mark.parametrize("mode", [BATCHED, SERIAL])
@pytest.mark.parametrize("delay_store", [None, "queues", "hbm"])
def test_delay_rich_recurrent_ring_outside_lds(monkeypatch, delay_store, mode):
    """Recurrent 10 % graph with delays 1..31, with the deposits in flight kept
    where the session chooses (the ring in LDS here), in per-slot arrival
    queues, or in a ring in HBM (SNC_HIP_TRAIN_DELAY forces the last two)."""
    if delay_store:
        monkeypatch.setenv("SNC_HIP_TRAIN_DELAY", delay_store)
    g = G.random_recurrent_graph(70, 400, 20, 0.1, 0.1, 1, 31, 0.25, seed=5)
    B = 6
    ev, lengths, y = I.event_streams(B, channels=70, classes=20, n_centres=20, sd=3.0,
                                     events_per_sample=300, t_min=40, t_max=60, seed=8)
    gcfg, ocfg = _cfgs(60, reset_mode=1, lr=0.3, w_init=0.25)
    sess = HipTrainSession(g, gcfg, batch=B, max_steps=60)
    sess.upload_events(ev, B, y, lengths)
    orc = O.Trainer(g, ocfg)
    w0 = orc.weights
    _compare(sess, orc, w0, np.arange(B), ev, lengths, y, mode)


def test_delay_stores_give_identical_weights(monkeypatch):
    """The ring in LDS, the arrival queues and the ring in HBM add every
    target's arrivals in the same order: after two minibatches with updates
    the weights are bit-identical."""
    g = G.random_recurrent_graph(70, 300, 20, 0.1, 0.1, 1, 31, 0.25, seed=6)
    B = 8
    ev, lengths, y = I.event_streams(B, channels=70, classes=20, n_centres=20, sd=3.0,
                                     events_per_sample=300, t_min=30, t_max=50, seed=9)
    gcfg, _ = _cfgs(50, reset_mode=1, lr=0.3, w_init=0.25)
    out = {}
    for mode in ("lds", "queues", "hbm"):
        monkeypatch.setenv("SNC_HIP_TRAIN_DELAY", mode)
        sess = HipTrainSession(g, gcfg, batch=B, max_steps=50)
        sess.upload_events(ev, B, y, lengths)
        for _ in range(2):
            sess.step(np.arange(B, dtype=np.int32), epoch=1, mode=BATCHED, apply=True)
        out[mode] = sess.get_weights()
        sess.close()
    assert np.array_equal(out["lds"], out["queues"])
    assert np.array_equal(out["lds"], out["hbm"])


def test_epochs_learn_prototype_classes():
    """A small run of the batched e-prop workload's shape (sheet 196 -> 96
    recurrent -> 10 leaky readout, prototype images, device Poisson input):
    four epochs over a 400-sample training split lift held-out accuracy far
    above chance (the same configuration reaches 0.70 on the CPU oracle)."""
    from snc_amd import configs as K
    g = G.sheet_graph([196, 96, 10], 24, recurrent_fan_in=16, dense_readout=True, radius=3.0,
                      sheet=14.0, w_std=0.3, delay=1, seed=1)
    Xu, y = I.prototype_images(600, 196, 10, 0.7, 0.2, seed=11)
    X = Xu.astype(np.float32) / np.float32(255.0)
    cfg = TrainConfig(num_steps=40, lif=K.lif_tau20(), lr=1.0, w_init=0.0, w_max=4.0,
                      surrogate_scale=0.3, feedback_scale=1.0, readout_mode=READOUT_LEAKY,
                      readout_psi=1.0, encoder=1, enc_max_rate=0.3, seed=1)
    sess = HipTrainSession(g, cfg, batch=32, max_steps=40)
    sess.set_weights(g.weights)
    sess.upload_features(X, y, np.full(600, 40, np.int32))
    tr, te = np.arange(400, dtype=np.int32), np.arange(400, 600, dtype=np.int32)
    acc0 = sess.evaluate(te)
    for ep in range(1, 5):
        st = sess.train_epoch(ep, idx=tr)
        assert st.samples == 400
    acc1 = sess.evaluate(te)
    assert acc1 >= 0.4 and acc1 >= acc0 + 0.2, (acc0, acc1)


def test_phase_profile_applies_nothing_and_accounts_for_the_launch():
    """snc_hip_train_profile: the timed build of the batched kernel leaves the
    weights as they were, every phase is non-negative, and the slowest
    workgroup's phases together fit inside the launch's device time."""
    g = G.sheet_graph([100, 64, 10], 16, recurrent_fan_in=8, radius=3.0, sheet=10.0, w_std=0.3,
                      max_delay=5, seed=2)
    B = 16
    X, y = I.prototype_images(B, 100, 10, 0.6, 0.2, seed=4)
    Xf = X.astype(np.float32) / np.float32(255.0)
    lengths = np.random.default_rng(1).integers(10, 40, B).astype(np.int32)
    gcfg, _ = _cfgs(40, readout=READOUT_LEAKY, reset_mode=1, encoder=1, max_rate=0.4)
    s = HipTrainSession(g, gcfg, batch=B, max_steps=40)
    s.upload_features(Xf, y, lengths)
    w0 = s.get_weights()
    p = s.profile(np.arange(B), epoch=1)
    assert np.array_equal(s.get_weights(), w0)
    ph = p["phases"]
    assert all(v["mean_us"] >= 0 and v["max_us"] >= v["mean_us"] for v in ph.values())
    assert ph["lif_trace_history"]["mean_us"] > 0 and ph["eligibility"]["mean_us"] > 0
    assert sum(v["mean_us"] for v in ph.values()) <= p["kernel_ms"] * 1e3 * 1.05
    s.close()
