"""The sheet connectome's structural step on the GPU (conn_gpu.hip) against
its host restatement (connectome.cpp): the prune set, the window candidates,
their 16 features and the learned grow's result are identical; one small
co-training run with the device path learns."""
import dataclasses

import numpy as np
import pytest

from snc_amd.grow import GrowMLP
from snc_amd.structure import (GROW_HEURISTIC, GROW_LEARNED, GROW_LEARNED_POOL, Connectome,
                               StructConfig)

pytestmark = pytest.mark.gpu

LAYERS = [196, 100, 10]


def _pair(seed):
    mk = lambda: Connectome.sheet_recurrent(LAYERS, fan_in=16, rec_fan_in=8, radius=0.15,
                                            seed=seed)
    return mk(), mk()


def _graph_equal(a, b):
    ga, gb = a.to_graph(), b.to_graph()
    return (np.array_equal(ga.row_ptr, gb.row_ptr) and np.array_equal(ga.post_ids, gb.post_ids)
            and np.array_equal(ga.weights, gb.weights))


@pytest.mark.parametrize("mode", [GROW_HEURISTIC, GROW_LEARNED, GROW_LEARNED_POOL])
def test_device_step_equals_host_step(mode):
    host, dev = _pair(7)
    mlp = GrowMLP(seed=3)
    rng = np.random.default_rng(7)
    cfg = StructConfig(grow_per_epoch=60, protect_rounds=1, locality_window=0.25, seed=7)
    for r in range(4):
        print(f"round {r} mode {mode}", flush=True)
        m = host.num_synapses
        w = (rng.standard_normal(m) * 0.2).astype(np.float32)
        w[::5] = w[1]  # ties in |w|: the order falls back to the CSR position
        host.set_edge_weights(w)
        dev.set_edge_weights(w)
        fires = rng.integers(0, 60, sum(LAYERS)).astype(np.int64)
        syn = np.zeros(m, np.int64)
        a = host.structural_update(syn, fires, cfg, mode=mode, model=mlp, activity_norm=50.0,
                                   log=True)
        b = dev.structural_update(syn, fires, dataclasses.replace(cfg, on_device=True), mode=mode,
                                  model=mlp, activity_norm=50.0, log=True)
        assert (a["pruned"], a["grown"], a["synapses"], a["candidates_scored"]) == \
            (b["pruned"], b["grown"], b["synapses"], b["candidates_scored"])
        assert _graph_equal(host, dev), r
    Xa, ya = host.grow_log()
    Xb, yb = dev.grow_log()
    assert np.array_equal(Xa.view(np.uint32), Xb.view(np.uint32)) and np.array_equal(ya, yb)
    if mode != GROW_HEURISTIC:
        assert Xa.shape[0] > 0
    if mode == GROW_LEARNED_POOL:
        # every window candidate of every non-input post was scored
        assert a["candidates_scored"] > 20 * sum(LAYERS[1:])


def test_device_cotraining_learns():
    """784 -> 512 (recurrent) -> 10 on sheets, signed N(0, 0.2) start weights,
    learned grow with candidates generated and scored on the device: held-out
    accuracy climbs well above chance (0.1) over four structural rounds."""
    from snc_amd import configs as K
    from snc_amd.structure import cotrain
    from snc_amd.train import HipTrainSession

    w = K.c2(n_train=1024, n_test=256, seed=2)
    con = Connectome.sheet_recurrent([784, 512, 10], fan_in=32, rec_fan_in=8, radius=0.2, seed=2)
    con.set_edge_weights(
        (np.random.default_rng(4).standard_normal(con.num_synapses) * 0.2).astype(np.float32))

    def make(g):
        s = HipTrainSession(g, dataclasses.replace(w.cfg, seed=2), batch=w.batch,
                            max_steps=w.max_steps)
        s.upload_features(w.X, w.labels, w.lengths)
        s.set_weights(con.edge_weights())  # the connectome's signed weights from structural epoch 0
        return s

    scfg = StructConfig(grow_per_epoch=300, locality_window=0.1, protect_rounds=2, seed=2,
                        on_device=True)
    n = w.labels.size
    rounds = cotrain(con, make, outer=4, inner=2, scfg=scfg, mode=GROW_LEARNED,
                     model=GrowMLP(seed=5), train_idx=np.arange(w.n_train, dtype=np.int32),
                     test_idx=np.arange(w.n_train, n, dtype=np.int32),
                     activity_norm=float(w.lengths[: w.n_train].sum()))
    accs = [r["test_acc"] for r in rounds]
    print(accs, [r.get("candidates_scored") for r in rounds], flush=True)
    assert all(r["candidates_scored"] > 1000 for r in rounds[:-1])
    assert all(r["synapses"] == rounds[0]["synapses"] for r in rounds)  # equal budget
    assert accs[-1] > 0.3 and accs[-1] > accs[0] + 0.1
