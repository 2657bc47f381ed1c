"""The command-line drivers on the GPU: the bench harness's self-test and its
per-sample spike counts against the oracle's forward of the same graph and
encodings; short training and co-training runs."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from snc_amd import graph as G
from snc_amd import inputs as I
from snc_amd.runtime import LIFParams

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "snc_amd", "bin")


def _run(args, timeout=240):
    print(" ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    return r


def test_bench_self_test():
    r = _run([os.path.join(BIN, "snc_hip_bench"), "--self-test"])
    assert r.returncode == 0 and "self-test: PASS" in r.stdout


@pytest.mark.parametrize("backend,batched", [("hip-bucket", False), ("hip-sort", False),
                                             ("hip-atomic", True), ("hip-auto-exact", True)])
def test_bench_matches_oracle_per_sample(tmp_path, backend, batched):
    n, dim, hidden, classes, T, budget, seed = 12, 196, 96, 10, 40, 6000, 3
    out = tmp_path / "bench.json"
    args = [os.path.join(BIN, "snc_hip_bench"), "--backend", backend, "--num-samples", str(n),
            "--dim", str(dim), "--hidden", str(hidden), "--classes", str(classes),
            "--num-steps", str(T), "--synapse-budget", str(budget), "--seed", str(seed),
            "--log-json", str(out)]
    if batched:
        args.append("--batched")
    r = _run(args)
    assert r.returncode == 0
    res = json.loads(out.read_text())
    g = G.static_snc([dim, hidden, classes], budget, 0.5, 1, 0.2, seed)
    X, _ = I.make_synthetic(n, dim, classes, 0.15, seed)
    lif = LIFParams(1.0, 0.9, 0.0, 1)
    want = [O.forward(g, I.encode("poisson", X[i], T, gain=1.0, max_rate=0.5, seed=seed + i), T,
                      lif)["spikes"] for i in range(n)]
    assert res["per_sample_spikes"] == want
    assert res["synapses"] == g.num_synapses
    # observability: device memory, rates, overflow
    assert res["device_graph_bytes"] >= 9 * g.num_synapses and res["device_plan_bytes"] > 0
    assert res["overflow"] == 0 and res["events_per_s"] > 0


def test_bench_compare_backends_all_exact():
    r = _run([os.path.join(BIN, "snc_hip_bench"), "--compare-backends", "--num-samples", "6",
              "--num-steps", "30"])
    assert r.returncode == 0
    rows = [l for l in r.stdout.splitlines() if l.startswith("hip-")]
    assert len(rows) == 5 and all(l.rstrip().endswith("exact") for l in rows)


def test_train_driver_learns():
    r = _run([os.path.join(BIN, "snc_hip_train"), "--dim", "196", "--hidden", "96",
              "--num-train", "400", "--num-test", "200", "--num-steps", "40", "--min-steps", "40",
              "--batch", "32", "--epochs", "4", "--w-std", "0.3", "--max-rate", "0.3",
              "--fan-in", "24", "--rec-fan-in", "16"])
    assert r.returncode == 0
    acc = [float(l.split("test_acc=")[1].split()[0]) for l in r.stdout.splitlines()
           if l.startswith("epoch")]
    assert len(acc) == 5 and acc[-1] >= 0.4 and acc[-1] > acc[0] + 0.2


def test_train_driver_serial_mode_equals_batched():
    """--train-mode serial launches the persistent kernel once per sample; its
    weights are the batched mode's bit for bit, so every epoch's held-out
    accuracy is the same."""
    accs = []
    for mode in ("batched", "serial"):
        r = _run([os.path.join(BIN, "snc_hip_train"), "--dim", "196", "--hidden", "96",
                  "--num-train", "128", "--num-test", "64", "--num-steps", "30",
                  "--min-steps", "20", "--batch", "32", "--epochs", "2", "--w-std", "0.3",
                  "--max-rate", "0.3", "--fan-in", "24", "--rec-fan-in", "16",
                  "--train-mode", mode])
        assert r.returncode == 0, r.stderr
        accs.append([l.split("test_acc=")[1].split()[0] for l in r.stdout.splitlines()
                     if l.startswith("epoch")])
    assert len(accs[0]) == 3 and accs[0] == accs[1]
    r = _run([os.path.join(BIN, "snc_hip_train"), "--train-mode", "sideways", "--epochs", "0"])
    assert r.returncode != 0 and "unknown --train-mode" in r.stderr


def test_train_driver_reward_mode_learns():
    """--reward-mode 1 (train_main.cpp:119): the three-factor rule sees only
    right/wrong per sample, so it climbs slower than the supervised delta;
    eight epochs of this deterministic run reach about three times chance."""
    r = _run([os.path.join(BIN, "snc_hip_train"), "--dim", "196", "--hidden", "96",
              "--num-train", "400", "--num-test", "200", "--num-steps", "40", "--min-steps", "40",
              "--batch", "32", "--epochs", "8", "--w-std", "0.3", "--max-rate", "0.3",
              "--fan-in", "24", "--rec-fan-in", "16", "--reward-mode", "1"])
    assert r.returncode == 0
    acc = [float(l.split("test_acc=")[1].split()[0]) for l in r.stdout.splitlines()
           if l.startswith("epoch")]
    assert len(acc) == 9 and acc[-1] > 0.12 and acc[-1] > acc[0] + 0.05, acc


@pytest.mark.parametrize("grow", ["heuristic", "learned"])
def test_cotrain_driver_runs(grow):
    r = _run([os.path.join(BIN, "snc_hip_cotrain"), "--grow", grow, "--outer", "4",
              "--num-train", "256", "--num-test", "128", "--num-steps", "40", "--min-steps", "30",
              "--batch", "64", "--structural-budget", "1500", "--grow-n", "60", "--log-runs", "1"])
    assert r.returncode == 0 and "final: grow=" + grow in r.stdout
    assert r.stdout.count("round ") == 4


def test_cotrain_driver_reward_rules_run():
    """--reward 1 (three-factor weight rule) and --reward-struct 1 (rewiring
    rate scaled by the reward), as cotrain_main.cpp:127-128 takes them."""
    r = _run([os.path.join(BIN, "snc_hip_cotrain"), "--grow", "heuristic", "--outer", "3",
              "--num-train", "256", "--num-test", "128", "--num-steps", "40", "--min-steps", "30",
              "--batch", "64", "--structural-budget", "1500", "--grow-n", "60", "--reward", "1",
              "--reward-struct", "1"])
    assert r.returncode == 0 and "final: grow=heuristic" in r.stdout
    assert r.stdout.count("round ") == 3


def test_cotrain_driver_sheet_structure_learns():
    r = _run([os.path.join(BIN, "snc_hip_cotrain"), "--structure", "sheet", "--grow", "learned",
              "--outer", "4", "--inner", "2", "--num-train", "1024", "--num-test", "256",
              "--grow-n", "300", "--locality", "0.1", "--log-runs", "1"])
    assert r.returncode == 0 and "final: grow=learned" in r.stdout
    acc = [float(l.split("test_acc=")[1].split()[0]) for l in r.stdout.splitlines()
           if l.startswith("round")]
    assert len(acc) == 4 and acc[-1] > 0.2 and acc[-1] > acc[0] + 0.1  # ten classes
    assert all("scored=" in l and "scored=0 " not in l for l in r.stdout.splitlines()
               if l.startswith("round") and "pruned=" in l)
