"""Host-side product code on the CPU: the C ABI library loads and exports
every declared symbol; generators, encoders and validation match the
reference (live when the compiled reference is present, else its goldens)."""
import os
import re

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, ROOT
from snc_amd import _lib as L
from snc_amd import graph as G
from snc_amd import inputs as I
from helpers import make_graph


def _declared_functions():
    names = set()
    for h in ("snc_hip.h", "snc_gen.h", "snc_struct.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(snc_\w+)\s*\(", src):
            names.add(m.group(1))
    return sorted(names)


def test_library_loads_and_exports_every_declared_symbol():
    lib = L.lib()
    declared = _declared_functions()
    assert len(declared) > 20
    for name in declared:
        assert hasattr(lib, name), name
    # and the Python binding covers all of them
    assert set(declared) <= set(L.exported_symbols())


def _header_defines():
    out = {}
    for h in ("snc_hip.h", "snc_gen.h", "snc_struct.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        for m in re.finditer(r"^#define\s+(SNC_\w+)\s+\(?(-?\d+)\)?", src, flags=re.M):
            out[m.group(1)] = int(m.group(2))
    return out


def test_python_constants_mirror_the_header():
    """Every integer the Python mirror passes across the C ABI (status codes,
    strategies, reset modes, training modes, eligibility precisions, readouts,
    kernel phases) has the value the header defines."""
    from snc_amd import train as T

    d = _header_defines()
    pairs = {
        "SNC_OK": L.SNC_OK, "SNC_E_INVALID": L.SNC_E_INVALID, "SNC_E_NODEVICE": L.SNC_E_NODEVICE,
        "SNC_E_HIP": L.SNC_E_HIP, "SNC_E_NOMEM": L.SNC_E_NOMEM,
        "SNC_E_UNSUPPORTED": L.SNC_E_UNSUPPORTED,
        "SNC_VARIANT_ATOMIC": L.VARIANT_ATOMIC, "SNC_VARIANT_BUCKET": L.VARIANT_BUCKET,
        "SNC_VARIANT_SORT": L.VARIANT_SORT, "SNC_VARIANT_AUTO": L.VARIANT_AUTO,
        "SNC_VARIANT_AUTO_EXACT": L.VARIANT_AUTO_EXACT, "SNC_VARIANT_PULL": L.VARIANT_PULL,
        "SNC_RESET_TO_VALUE": L.RESET_TO_VALUE, "SNC_RESET_SUBTRACT": L.RESET_SUBTRACT,
        "SNC_TRAIN_PLAIN": T.PLAIN, "SNC_TRAIN_BATCHED": T.BATCHED, "SNC_TRAIN_SERIAL": T.SERIAL,
        "SNC_ELIG_F32": T.ELIG_F32, "SNC_ELIG_F64": T.ELIG_F64,
        "SNC_READOUT_SPIKING": T.READOUT_SPIKING, "SNC_READOUT_LEAKY": T.READOUT_LEAKY,
        "SNC_TRAIN_PHASES": T.PHASES,
    }
    for name, value in pairs.items():
        assert name in d, name
        assert d[name] == value, (name, d[name], value)


def test_encode_matches_reference_goldens():
    z = np.load(os.path.join(GOLDEN, "encode.npz"), allow_pickle=False)
    for k in range(3):
        ev = I.encode(k, z["x"], 40, gain=1.5, max_rate=0.3, threshold=0.05, seed=77)
        assert ev.tobytes() == z[f"enc_{k}"].tobytes()
    s = np.load(os.path.join(GOLDEN, "synthetic.npz"), allow_pickle=False)
    X, y = I.make_synthetic(12, 20, 3, 0.25, 42)
    assert np.array_equal(X, s["X"]) and np.array_equal(y, s["y"])


def test_static_snc_matches_reference_golden_graph():
    z = np.load(os.path.join(GOLDEN, "fwd_static_direct.npz"), allow_pickle=False)
    g = G.static_snc([64, 48, 8], 1500, 0.5, 3, 0.2, 4)
    for f in ("row_ptr", "post_ids", "weights", "delays", "sign", "role", "channel"):
        assert np.array_equal(getattr(g, f), z[f]), f


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not present")
@pytest.mark.parametrize("layers,budget,seed", [([784, 256, 10], 40000, 1),
                                                ([100, 60, 30, 10], 4000, 9)])
def test_reference_generators_identical(layers, budget, seed):
    pairs = [(G.static_snc(layers, budget, 0.5, 2, 0.2, seed),
              O.ref_static_snc(layers, budget, 0.5, 2, 0.2, seed)),
             (G.random_sparse(layers, budget, 0.5, 1, 0.3, seed),
              O.ref_random_sparse(layers, budget, 0.5, 1, 0.3, seed)),
             (G.dense(layers[:3], 0.25, 4), O.ref_dense(layers[:3], 0.25, 4))]
    for a, r in pairs:
        for f in ("row_ptr", "post_ids", "weights", "delays", "sign", "role", "channel"):
            assert np.array_equal(getattr(a, f), getattr(r, f)), f


def test_validate_rejects_what_the_reference_rejects():
    g = make_graph(3, [(0, 1, 0.5, 1), (1, 2, 0.5, 2)])
    assert g.validate()[0]
    bad = make_graph(3, [(0, 1, 0.5, 0)])
    ok, msg = bad.validate()
    assert not ok and "delay" in msg
    bad = make_graph(3, [(0, 1, 0.5, 1)])
    bad.post_ids[0] = 7
    assert not bad.validate()[0]
    if O.ref_available():
        assert not O.ref_validate(bad)


def test_sheet_generator_shape_and_locality():
    g = G.sheet_graph([784, 256, 10], 32, radius=3.0, w_std=1 / np.sqrt(32), seed=1)
    assert g.num_neurons == 1050 and g.num_synapses == 266 * 32
    assert g.validate()[0]
    indeg = g.in_degree()
    assert (indeg[784:] == 32).all() and (indeg[:784] == 0).all()
    # partners are local: mean input-to-hidden distance is a few grid units
    pre, post = g.pre_ids(), g.post_ids
    m = (pre < 784) & (post < 1040)
    px, py = pre[m] % 28, pre[m] // 28
    k = post[m] - 784
    hx, hy = (k % 16 + 0.5) * 28 / 16 - 0.5, (k // 16 + 0.5) * 28 / 16 - 0.5
    assert np.hypot(px - hx, py - hy).mean() < 6.0
    assert abs(np.std(g.weights) - 1 / np.sqrt(32)) < 0.02


def test_hub_generator_degrees():
    g = G.hub_sheet_graph(60, 50, 20, hub_fraction=0.02, hub_min=40, hub_max=800, max_delay=16,
                          seed=3)
    assert g.num_synapses == 3000 * 20 and g.validate()[0]
    assert (np.diff(g.row_ptr) == 20).all()
    indeg = g.in_degree()
    assert indeg.max() >= 40
    assert 1 <= g.delays.min() and g.delays.max() <= 16
    # rows are sorted and duplicate-free
    for i in range(0, 3000, 97):
        r = g.post_ids[g.row_ptr[i]:g.row_ptr[i + 1]]
        assert (np.diff(r) > 0).all() and i not in r


def test_background_drive_rate_and_determinism():
    a = I.background_drive(5000, 200, 0.01, 1.0, seed=4)
    b = I.background_drive(5000, 200, 0.01, 1.0, seed=4)
    assert a.tobytes() == b.tobytes()
    rate = a.size / (5000 * 200)
    assert abs(rate - 0.01) < 0.001
    assert (np.diff(a["time"]) >= 0).all()
    # Philox stream: neuron i, step t fires iff word (i % 4) of block
    # philox(i - i % 4, t, 0x5EED, 0; seed) is below p * 2^32.
    thr = int(0.01 * 2 ** 32)
    for k in range(0, a.size, max(1, a.size // 20)):
        i, t = int(a["channel_id"][k]), int(a["time"][k])
        r = O.philox4x32([i - i % 4, t, 0x5EED, 0], [4, 0])
        assert r[i % 4] < thr


def test_trainer_rejects_unknown_encoder():
    """snc_hip_train_create checks the encoder (direct, poisson, latency)
    before it looks for a device."""
    from snc_amd.train import HipTrainSession, TrainConfig

    g = G.static_snc([20, 10, 4], 100, 0.5, 1, 0.2, 1)
    with pytest.raises(L.HipError, match="unknown encoder"):
        HipTrainSession(g, TrainConfig(num_steps=10, encoder=3), batch=2, max_steps=10)
