"""SNCG graph files: the reference exporter's bytes (golden fixtures written by
oracle/_ref/snc_export, and the live binary when it is built), the reader's
semantics (python/snc/graph.py:11-24) and its error handling. CPU only."""
import json
import os
import shlex
import subprocess

import numpy as np
import pytest

from snc_amd import _lib as L
from snc_amd import graph as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXPORT = os.path.join(ROOT, "snc_amd", "bin", "snc_hip_export")
REF_EXPORT = os.path.join(ROOT, "oracle", "_ref", "snc_export")
CASES = json.load(open(os.path.join(GOLD, "sncg_cases.json")))


def _parse(path):
    """The reference loader's reading of the file (graph.py:11-24)."""
    raw = np.fromfile(path, dtype="<i4")
    magic, n, s, n_in, n_out = (int(x) for x in raw[:5])
    assert magic == 0x534E4347
    o = 5
    out = {"N": n, "S": s, "n_in": n_in, "n_out": n_out}
    for name, count in (("pre", s), ("post", s), ("delays", s), ("role", n), ("channel", n)):
        out[name] = raw[o:o + count]
        o += count
    assert o == raw.size
    return out


def _generate(args):
    """The exporter's graph for a command line, through the Python API."""
    a = shlex.split(args)
    opt = dict(zip(a[::2], a[1::2]))
    layers = [int(x) for x in opt.get("--layers", "784,256,10").split(",")]
    st = opt.get("--structure", "static-snc")
    budget, seed = int(opt.get("--synapse-budget", 40000)), int(opt.get("--seed", 1))
    delay, inhib = int(opt.get("--delay", 1)), float(opt.get("--inhib", 0.2))
    if st == "dense":
        g = G.dense(layers, 0.5, delay)
    elif st == "random-sparse":
        g = G.random_sparse(layers, budget, 0.5, delay, inhib, seed)
    else:
        g = G.static_snc(layers, budget, 0.5, delay, inhib, seed)
    mode = G.DELAYS_DISTANCE if opt.get("--delay-mode") == "distance" else G.DELAYS_RANDOM
    return G.spread_delays(g, mode, int(opt.get("--delay-max", 1)), layers, seed)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_cli_writes_the_reference_bytes(case, tmp_path):
    out = tmp_path / "g.bin"
    r = subprocess.run([EXPORT, *shlex.split(case["args"]), "--out", str(out)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == open(os.path.join(GOLD, case["name"] + ".bin"), "rb").read()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_api_writes_the_reference_bytes(case, tmp_path):
    g = _generate(case["args"])
    out = tmp_path / "g.bin"
    G.write_sncg(g, out)
    assert out.read_bytes() == open(os.path.join(GOLD, case["name"] + ".bin"), "rb").read()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_read_follows_the_reference_loader(case, tmp_path):
    path = os.path.join(GOLD, case["name"] + ".bin")
    ref = _parse(path)
    g, n_in, n_out = G.read_sncg(path, weight=0.25)
    assert (g.num_neurons, g.num_synapses, n_in, n_out) == (ref["N"], ref["S"], ref["n_in"],
                                                            ref["n_out"])
    assert np.array_equal(g.pre_ids(), ref["pre"])
    assert np.array_equal(g.post_ids, ref["post"]) and np.array_equal(g.delays, ref["delays"])
    assert np.array_equal(g.role, ref["role"]) and np.array_equal(g.channel, ref["channel"])
    assert np.all(g.weights == np.float32(0.25)) and np.all(g.sign == 1)
    assert (g.num_input_channels, g.num_output_channels) == (n_in, n_out)
    assert g.validate()[0]
    out = tmp_path / "again.bin"  # write(read(f)) == f
    G.write_sncg(g, out)
    assert out.read_bytes() == open(path, "rb").read()


def test_read_keeps_file_order_within_rows_of_an_unsorted_file(tmp_path):
    # pre not ascending: rows are regrouped, each row keeping file order
    pre = np.array([2, 0, 2, 1, 0], np.int32)
    post = np.array([0, 1, 1, 2, 2], np.int32)
    delays = np.array([3, 1, 2, 1, 4], np.int32)
    role = np.array([1, 0, 2], np.int32)
    channel = np.array([0, 0, 0], np.int32)
    path = tmp_path / "u.bin"
    np.concatenate([np.array([0x534E4347, 3, 5, 1, 1], np.int32), pre, post, delays, role,
                    channel]).astype("<i4").tofile(path)
    g, _, _ = G.read_sncg(path)
    assert list(g.row_ptr) == [0, 2, 3, 5]
    assert list(g.post_ids) == [1, 2, 2, 0, 1] and list(g.delays) == [1, 4, 1, 3, 2]


@pytest.mark.parametrize("damage", ["magic", "truncate", "post", "delay", "role"])
def test_read_rejects_damaged_files(damage, tmp_path):
    raw = np.fromfile(os.path.join(GOLD, "sncg_static_snc.bin"), dtype="<i4").copy()
    n, s = int(raw[1]), int(raw[2])
    if damage == "magic":
        raw[0] = 0x12345678
    elif damage == "truncate":
        raw = raw[:-1]
    elif damage == "post":
        raw[5 + s] = n
    elif damage == "delay":
        raw[5 + 2 * s] = 0
    elif damage == "role":
        raw[5 + 3 * s] = 3
    path = tmp_path / "bad.bin"
    raw.tofile(path)
    with pytest.raises(L.HipError):
        G.read_sncg(path)


def test_distance_delays_closed_form():
    # one synapse per pre onto post A + pre: span 0 -> delay 1; the exporter's
    # rule gives 1 + round(|pre - (post - A)| / (max(A, B) - 1) * (dmax - 1))
    A, B, dmax = 5, 5, 9
    rows = np.arange(A + B + 1, dtype=np.int32)
    rows[A + 1:] = A
    post = np.array([A + 4, A + 3, A + 2, A + 1, A], np.int32)  # spans 4, 2, 0, 2, 4
    g = G.SNNGraph(rows, post, np.ones(A, np.float32), np.ones(A, np.int32),
                   np.ones(A + B, np.int8), np.zeros(A + B, np.uint8), np.zeros(A + B, np.int32))
    d = G.spread_delays(g, G.DELAYS_DISTANCE, dmax, [A, B]).delays
    assert list(d) == [1 + round(k / 4 * 8) for k in (4, 2, 0, 2, 4)]


@pytest.mark.skipif(not os.path.exists(REF_EXPORT), reason="reference exporter not built")
@pytest.mark.parametrize("args", [
    "--structure static-snc --layers 784,256,10 --synapse-budget 40000 --seed 1",
    "--structure random-sparse --layers 100,80,20,5 --synapse-budget 3000 --seed 9 "
    "--delay-max 31",
    "--structure static-snc --layers 300,200 --synapse-budget 5000 --seed 4 --delay-max 16 "
    "--delay-mode distance --inhib 0.35",
])
def test_cli_equals_live_reference_exporter(args, tmp_path):
    a, b = tmp_path / "ref.bin", tmp_path / "mine.bin"
    subprocess.run([REF_EXPORT, *shlex.split(args), "--out", str(a)], check=True,
                   capture_output=True, timeout=120)
    subprocess.run([EXPORT, *shlex.split(args), "--out", str(b)], check=True,
                   capture_output=True, timeout=120)
    assert a.read_bytes() == b.read_bytes()


def test_export_cli_usage():
    r = subprocess.run([EXPORT, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "snc_hip_export" in r.stdout
    r = subprocess.run([EXPORT, "--no-such-option"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
