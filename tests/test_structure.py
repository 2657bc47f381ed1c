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


def test_grow_log_rank_is_the_abs_weight_percentile_at_the_deciding_prune():
    """With protect_rounds = 0 every edge is eligible at every prune, so an
    edge grown (and logged) at update 0 is decided at update 1 among all N
    edges. With distinct |w| its rank is exactly (k + 0.5) / N for its
    position k in |w| order, and it was pruned iff k < K."""
    con = Connectome([50, 30, 6], 300, 1, 0.02, 6)
    rng = np.random.default_rng(6)
    K = 20
    cfg = StructConfig(grow_per_epoch=K, protect_rounds=0)
    n0 = con.num_synapses
    con.set_edge_weights((rng.standard_normal(n0) * 0.3).astype(np.float32))
    con.structural_update(np.zeros(n0, np.int64), rng.integers(0, 30, 86).astype(np.int64), cfg,
                          log=True)
    N = con.num_synapses
    w1 = rng.permutation(N).astype(np.float32) * np.float32(0.001) + np.float32(0.0005)
    w1 *= np.where(rng.random(N) < 0.5, -1, 1).astype(np.float32)  # distinct |w|, mixed signs
    con.set_edge_weights(w1)
    con.structural_update(np.zeros(N, np.int64), rng.integers(0, 30, 86).astype(np.int64), cfg,
                          log=False)
    _, y = con.grow_log()
    _, r = con.grow_log(rank=True)
    assert y.size == r.size > 0 and np.all(y >= 0)
    k = r.astype(np.float64) * N - 0.5
    assert np.allclose(k, np.round(k), atol=1e-3)
    k = np.round(k).astype(np.int64)
    assert np.all((k >= 0) & (k < N))
    assert np.array_equal(y == 0, k < K)
    # grown edges start at w_grow_init, but w1 replaced every weight: ranks spread
    assert len(np.unique(k)) == k.size


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


# ---- sheet connectome (2-D layers, recurrent hidden layers) ----

def _sheet_xy(layers):
    """Unit-square grid positions of every neuron, as the sheet connectome lays them out."""
    xs, ys, lay = [], [], []
    for l, cnt in enumerate(layers):
        side = int(np.ceil(np.sqrt(cnt)))
        k = np.arange(cnt)
        if side > 1:
            xs.append((k % side).astype(np.float32) / np.float32(side - 1))
            ys.append((k // side).astype(np.float32) / np.float32(side - 1))
        else:
            xs.append(np.full(cnt, 0.5, np.float32))
            ys.append(np.full(cnt, 0.5, np.float32))
        lay.append(np.full(cnt, l))
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(lay)


def _dist(x, y, q, p):
    return np.sqrt((x[q].astype(np.float64) - x[p]) ** 2 + (y[q].astype(np.float64) - y[p]) ** 2)


def test_sheet_connectome_structure():
    layers = [100, 64, 10]
    con = Connectome.sheet_recurrent(layers, fan_in=12, rec_fan_in=6, radius=0.15, seed=3)
    g = con.to_graph()
    x, y, lay = _sheet_xy(layers)
    pre, post = g.pre_ids(), g.post_ids
    assert np.all(pre != post)
    # hidden posts: 12 from the previous layer, 6 from their own; outputs: 12
    same = lay[pre] == lay[post]
    assert np.all((lay[pre] == lay[post] - 1) | (same & (lay[post] == 1)))
    for p in range(100, 174):
        fw = int(((post == p) & ~same).sum())
        rc = int(((post == p) & same).sum())
        assert fw == 12 and rc == (6 if p < 164 else 0)
    # locality: partners are much closer than random pairs of the same layers
    d = _dist(x, y, pre, post)
    rng = np.random.default_rng(0)
    q = rng.integers(0, 100, 4000)
    p = rng.integers(100, 164, 4000)
    assert d[lay[post] == 1].mean() < 0.6 * _dist(x, y, q, p).mean()
    again = Connectome.sheet_recurrent(layers, fan_in=12, rec_fan_in=6, radius=0.15, seed=3).to_graph()
    assert np.array_equal(g.post_ids, again.post_ids) and np.array_equal(g.row_ptr, again.row_ptr)
    assert list(np.unique(g.role)) == [0, 1, 2]


def test_sheet_prune_takes_the_smallest_abs_weights_and_grows_in_window():
    layers = [81, 49, 9]
    con = Connectome.sheet_recurrent(layers, fan_in=10, rec_fan_in=5, radius=0.2, seed=5)
    x, y, lay = _sheet_xy(layers)
    g0 = con.to_graph()
    m = g0.num_synapses
    rng = np.random.default_rng(1)
    w = rng.standard_normal(m).astype(np.float32) * 0.3
    w[::7] = w[3]  # ties, broken by CSR position
    con.set_edge_weights(w)
    fires = rng.integers(0, 40, g0.num_neurons).astype(np.int64)
    cfg = StructConfig(grow_per_epoch=25, protect_rounds=0, locality_window=0.3)
    rep = con.structural_update(np.zeros(m, np.int64), fires, cfg, mode=GROW_HEURISTIC)
    order = np.lexsort((np.arange(m), np.abs(w)))  # |w| ascending, then position
    gone = set(order[:25].tolist())
    pre0 = g0.pre_ids()
    kept = [(int(pre0[e]), int(g0.post_ids[e])) for e in range(m) if e not in gone]
    g1 = con.to_graph()
    now = list(zip(g1.pre_ids().tolist(), g1.post_ids.tolist()))
    assert rep["pruned"] == 25
    new = [e for e in now if e not in set(kept)]
    assert len(new) == rep["grown"] > 0
    for q, p in new:  # every grown synapse: a window member of an allowed layer
        assert _dist(x, y, q, p) <= np.float64(np.float32(0.3))
        assert lay[q] == lay[p] - 1 or (lay[q] == lay[p] == 1)


def test_sheet_candidate_distance_feature():
    layers = [64, 36, 4]
    con = Connectome.sheet_recurrent(layers, fan_in=8, rec_fan_in=4, radius=0.2, seed=2)
    x, y, _ = _sheet_xy(layers)
    pre = np.array([0, 5, 63, 70], np.int32)
    post = np.array([64, 80, 99, 71], np.int32)
    F = con.candidate_features(pre, post, np.zeros(104, np.int64))
    assert np.array_equal(F[:, 3], _dist(x, y, pre, post).astype(np.float32))


def test_prune_to_keeps_the_largest_abs_weights():
    con = Connectome.sheet_recurrent([64, 36, 4], fan_in=8, rec_fan_in=4, radius=0.2, seed=5)
    m = con.num_synapses
    rng = np.random.default_rng(5)
    w = (rng.standard_normal(m) * 0.2).astype(np.float32)
    w[::7] = w[3]  # ties: the earlier CSR position goes first
    con.set_edge_weights(w)
    g0 = con.to_graph()
    con.prune_to(m)  # at the budget already: unchanged
    assert con.num_synapses == m
    con.prune_to(100)
    assert con.num_synapses == 100
    order = sorted(range(m), key=lambda e: (abs(float(w[e])), e))
    keep = np.zeros(m, bool)
    keep[order[m - 100:]] = True
    # survivors in CSR order, with their weights and posts
    assert np.array_equal(con.edge_weights(), w[keep])
    g1 = con.to_graph()
    assert np.array_equal(g1.post_ids, np.asarray(g0.post_ids)[keep])


def test_negative_grow_init_starts_grown_synapses_at_the_survivors_median():
    con = Connectome([40, 30, 6], 400, 1, 0.02, 3)
    rng = np.random.default_rng(3)
    g0 = con.to_graph()
    w = (rng.standard_normal(g0.num_synapses) * 0.3).astype(np.float32)
    con.set_edge_weights(w)
    before = {(int(i), int(j)) for i, j in zip(np.repeat(np.arange(g0.num_neurons), np.diff(g0.row_ptr)), g0.post_ids)}
    cfg = StructConfig(grow_per_epoch=25, w_grow_init=-1.0, locality_window=0.3, protect_rounds=0, seed=3)
    fires = rng.integers(0, 50, g0.num_neurons).astype(np.int64)
    rep = con.structural_update(np.zeros(g0.num_synapses, np.int64), fires, cfg, mode=GROW_HEURISTIC)
    assert rep["grown"] > 0
    assert rep["pruned"] == 25
    # the survivors are all but the 25 smallest |w| (every synapse eligible)
    survivors = np.sort(np.abs(w))[25:]
    med = survivors[(survivors.size - 1) // 2]
    g1 = con.to_graph()
    w1 = con.edge_weights()
    pre1 = np.repeat(np.arange(g1.num_neurons), np.diff(g1.row_ptr))
    new_pair = np.array([(int(i), int(j)) not in before for i, j in zip(pre1, g1.post_ids)])
    assert new_pair.any() and np.all(w1[new_pair] == med)
    # every grown synapse (a regrown pruned pair included) carries the median
    assert np.count_nonzero(w1 == med) >= rep["grown"]


def test_sheet_heuristic_grow_equals_brute_force_rule():
    """The sheet connectome's heuristic step against a brute-force
    restatement of its rule: prune the 30 smallest (|w|, CSR position); each
    freed post, in the pruned synapses' CSR order, regrows from the open
    member of its window (previous layer, and its own when hidden, within the
    locality distance, self excluded; layer then id order) that fired most,
    the first such on ties. Enough window candidates are left that no post
    goes without, so the tournament fill never draws; the resulting edges and
    the grown synapses' start weight must match exactly."""
    layers = [81, 49, 9]
    con = Connectome.sheet_recurrent(layers, fan_in=10, rec_fan_in=5, radius=0.2, seed=7)
    x, y, lay = _sheet_xy(layers)
    g0 = con.to_graph()
    n, m = g0.num_neurons, g0.num_synapses
    rng = np.random.default_rng(11)
    w = (rng.standard_normal(m) * 0.3).astype(np.float32)
    con.set_edge_weights(w)
    fires = rng.integers(0, 12, n).astype(np.int64)  # many ties
    win = 0.35
    cfg = StructConfig(grow_per_epoch=30, protect_rounds=0, locality_window=win)
    rep = con.structural_update(np.zeros(m, np.int64), fires, cfg, mode=GROW_HEURISTIC)

    order = np.lexsort((np.arange(m), np.abs(w)))
    gone = np.zeros(m, bool)
    gone[order[:30]] = True
    pre0, post0 = g0.pre_ids(), np.asarray(g0.post_ids)
    conn = set(zip(pre0[~gone].tolist(), post0[~gone].tolist()))
    limit = float(np.float32(win))
    hidden_top = len(layers) - 2
    grown = []
    for p in post0[gone].tolist():  # the freed posts, in CSR order
        lp = int(lay[p])
        best, most = -1, -1
        for q in range(n):  # ids run layer by layer, so id order is layer order
            ok_layer = lay[q] == lp - 1 or (lay[q] == lp and 1 <= lp <= hidden_top)
            if q == p or not ok_layer or (q, p) in conn:
                continue
            if _dist(x, y, q, p) <= limit and fires[q] > most:
                best, most = q, int(fires[q])
        assert best >= 0
        conn.add((best, p))
        grown.append((best, p))
    assert rep["pruned"] == 30 and rep["grown"] == 30 == len(grown)
    g1 = con.to_graph()
    got = sorted(zip(g1.pre_ids().tolist(), np.asarray(g1.post_ids).tolist()))
    expect = sorted(list(zip(pre0[~gone].tolist(), post0[~gone].tolist())) + grown)
    assert got == expect
    w1 = con.edge_weights()
    pairs = list(zip(g1.pre_ids().tolist(), np.asarray(g1.post_ids).tolist()))
    new = np.array([pr in set(grown) for pr in pairs])
    assert new.sum() == 30 and np.all(w1[new] == np.float32(cfg.w_grow_init))
