"""The learned grow proposal against the heuristic it replaces, at an equal
synapse budget, on the structural co-training workload (layered local
connectome 784 -> 512 -> 10 under 3000 synapses, grow = prune = 200 per
structural epoch, 12 structural epochs of batched e-prop on the prototype
image task).

The MLP is trained on the GPU, by its own forward/backward kernels, from grow
outcomes logged by heuristic co-training runs (seeds 101-103), then compared
on seeds 1-5 it has not seen. Every part is deterministic (ordered folds and
reductions), so the comparison is reproducible bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_learned_grow_beats_heuristic_at_equal_budget():
    from snc_amd.growlab import ab_compare

    res = ab_compare(log=lambda m: print(m, flush=True))
    # equal budget: both rules end every run with the same synapse count
    assert all(a == b for a, b in res["synapses"])
    # the MLP learned something from the log
    assert res["mlp_loss_after"] < res["mlp_loss_before"]
    # and what it gains over the heuristic on held-out seeds
    print(f"heuristic {res['heuristic']} mean {res['heuristic_mean']:.4f}")
    print(f"learned   {res['learned']} mean {res['learned_mean']:.4f}")
    assert res["learned_mean"] > res["heuristic_mean"]
    assert res["wins"] >= 3
    # every structural step scored the window candidates of all freed posts
    assert min(res["candidates_scored_per_step"]) > 10000


def test_c4_learned_pool_grow_against_heuristic_reported():
    """The structural co-training workload as stated: the batched e-prop
    network (784 -> 512 recurrent -> 10 on 2-D sheets) thinned by |w| to
    3000 synapses, grow = prune = 200 per structural epoch, 12 epochs; the
    learned rule scores every candidate pair the architecture allows (the
    whole pool, about 668k pairs) each structural step and regrows the
    highest-scoring ones. Grown synapses start at the surviving synapses'
    median |w|, so their first eligible prune judges what training made of
    them. The MLP trains on the graded outcome (the grown synapse's |w|
    percentile at that prune) logged by heuristic runs (seeds 101-103, half of
    the choices uniform) plus one on-policy round. Those settings were chosen
    on validation seeds 201-203. The comparison on seeds 1-5 against the
    heuristic (at the same whole-layer window, and at its own 0.1 locality)
    is printed, not asserted: it is reported whichever way it comes out."""
    from snc_amd.growlab import ab_compare

    res = ab_compare(kind="c4", label="rank", w_grow=-1.0, onpolicy=1,
                     log=lambda m: print(m, flush=True))
    # equal budget: every rule ends every run at the 3000-synapse budget
    assert all(s == [3000, 3000, 3000] for s in res["synapses"])
    # the whole pool was scored at every structural step
    assert min(res["candidates_scored_per_step"]) >= 600000
    assert res["mlp_loss_after"] < res["mlp_loss_before"]
    print(f"heuristic (whole-layer window) {res['heuristic']} mean {res['heuristic_mean']:.4f}")
    print(f"heuristic (0.1 locality)       {res['heuristic_local']} "
          f"mean {res['heuristic_local_mean']:.4f}")
    print(f"learned pool                   {res['learned']} mean {res['learned_mean']:.4f}")
    print(f"gain {res['gain_mean_pt']:+.2f} pt ({res['wins']}/5 wins), vs local "
          f"{res['gain_vs_local_mean_pt']:+.2f} pt ({res['wins_vs_local']}/5 wins); "
          f"label mean {res['log_label_mean']:.4f}")
