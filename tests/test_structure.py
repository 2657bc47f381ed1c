"""Structural slow path on the CPU: the heuristic grow/prune equals the
reference Connectome's (live against the compiled reference when present),
and the learned mode's bookkeeping (grow log, labels, features) holds."""
import numpy as np
import pytest

import oracle as O
from oracle.oracle import _p, _ref_graph_take
from snc_amd.structure import GROW_HEURISTIC, Connectome, StructConfig


def _ref_connectome_graph(h):
    return _ref_graph_take(O.ref_lib().ref_connectome_to_graph(h))


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not present")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_heuristic_update_equals_reference(seed):
    layers = np.array([60, 40, 8], np.int32)
    budget, delay, w_init = 500, 1, 0.02
    con = Connectome(layers, budget, delay, w_init, seed)
    ref = O.ref_lib().ref_connectome_create(_p(layers), 3, budget, delay, w_init, seed)
    try:
        rng = np.random.default_rng(seed)
        cfg = StructConfig(grow_per_epoch=30, delay=2, w_grow_init=0.05, locality_window=0.25,
                           protect_rounds=2, seed=seed)
        for r in range(6):
            a, b = con.to_graph(), _ref_connectome_graph(ref)
            for f in ("row_ptr", "post_ids", "weights", "delays", "role", "channel"):
                assert np.array_equal(getattr(a, f), getattr(b, f)), (r, f)
            # the same trained weights and activity go into both
            w = (rng.standard_normal(a.num_synapses) * 0.3).astype(np.float32)
            con.set_edge_weights(w)
            O.ref_lib().ref_connectome_set_edge_weights(ref, _p(w), w.size)
            fires = rng.integers(0, 50, a.num_neurons).astype(np.int64)
            syn = np.zeros(a.num_synapses, np.int64)
            rep = con.structural_update(syn, fires, cfg, mode=GROW_HEURISTIC)
            pr, gr = np.zeros(1, np.int32), np.zeros(1, np.int32)
            O.ref_lib().ref_connectome_update(ref, _p(syn), syn.size, _p(fires), fires.size,
                                              cfg.grow_per_epoch, cfg.delay, cfg.w_grow_init,
                                              cfg.locality_window, cfg.protect_rounds, seed,
                                              _p(pr), _p(gr))
            assert (rep["pruned"], rep["grown"]) == (int(pr[0]), int(gr[0]))
            assert rep["synapses"] == O.ref_lib().ref_connectome_num_synapses(ref)
    finally:
        O.ref_lib().ref_connectome_free(ref)


def test_grow_log_labels_follow_first_eligible_prune():
    con = Connectome([50, 30, 6], 300, 1, 0.02, 4)
    rng = np.random.default_rng(4)
    cfg = StructConfig(grow_per_epoch=20, protect_rounds=2)
    for r in range(6):
        n = con.num_synapses
        w = (rng.standard_normal(n) * 0.3).astype(np.float32)
        con.set_edge_weights(w)
        fires = rng.integers(0, 30, 86).astype(np.int64)
        con.structural_update(np.zeros(n, np.int64), fires, cfg, log=True)
    X, y = con.grow_log()
    assert X.shape[0] > 0 and X.shape[1] == 16
    assert set(np.unique(y)) <= {-1, 0, 1}
    # logged at rounds 0..5; eligible (age 2) two updates later: rounds 0..3 decided
    assert (y >= 0).sum() > 0 and (y == -1).sum() > 0
    assert np.all(X[:, 15] == 1.0)  # bias feature


def test_candidate_features_are_finite_and_local():
    con = Connectome([50, 30, 6], 300, 1, 0.02, 4)
    fires = np.arange(86, dtype=np.int64)
    pre = np.array([0, 10, 49], np.int32)
    post = np.array([50, 60, 79], np.int32)
    F = con.candidate_features(pre, post, fires, activity_norm=100.0)
    assert np.all(np.isfinite(F))
    assert np.allclose(F[:, 0], fires[pre] / 100.0)
    assert np.allclose(F[:, 1], fires[post] / 100.0)
    assert np.all(F[:, 3] >= 0) and np.all(F[:, 3] <= 1)
