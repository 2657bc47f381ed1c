"""Throughput benchmark of the spiking path on one MI355X.

A "step" is one forward pass of the delivery-stress workload over its whole
window: 200k LIF neurons, 20M synapses, 1000 timesteps, batch 1 (graph and
drive resident in HBM before timing starts). value = synaptic events per
second over the timed steps, in GSynOps/s.

Prints one JSON line (rank 0). Extra objects:
  roofline      dominant kernel: algorithmic bytes per launch (byte model
                below) / its average launch duration from HIP events on the
                plan's stream, against the 8 TB/s HBM peak
  cpu_baseline  the reference's own CPU forward (oracle/_ref, compiled from
                the reference sources) or the oracle restatement, timed on
                this host on the same workload
  eprop         batched vs plain e-prop on the two training workloads, and
                (eprop.*.cpu_baseline) the oracle's trainer on one core over
                one minibatch of the same workload

Byte model (algorithmic, per step): 16 B per neuron (membrane read+write,
ring slot read+clear), 9 B per synaptic event (post id 4, weight 4, delay 1),
8 B per fired neuron (row_ptr pair). Each kernel is credited the neurons it
steps and the rows it expands: the hub kernel (hip-bucket) the heavy
targets, the step kernel the rest (everything for hip-atomic).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "GSynOps/s + e-prop samples/s, batched vs plain, 1 MI355X, equal accuracy"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_units(g, neuron_spikes, steps):
    """Work per kernel class over one run: (neurons stepped, events expanded,
    rows fired). One fused step kernel steps every neuron, hub targets
    included."""
    import numpy as np

    outdeg = np.diff(g.row_ptr).astype(np.int64)
    spk = neuron_spikes.astype(np.int64)
    ev_all, f_all = int((spk * outdeg).sum()), int(spk.sum())
    n = g.num_neurons
    return {name: (n * steps, ev_all, f_all) for name in ("k_step", "sort_step", "k_persist")}


def scatter_floor(g, cfg, events_per_step, kernel_ms):
    """The device's time floor for the fused step kernel's delivery: one
    launch of the same shape doing events_per_step scattered device-scope
    atomics of the kernel's kind over a buffer the size of its target (the
    ring, or the arrival bitmaps), with nothing else (snc_hip_measure_scatter).
    frac = floor / kernel time: how close the step is to what its atomics
    alone cost on this device."""
    import numpy as np

    from snc_amd.runtime import (SCATTER_ATOMIC_ADD, SCATTER_ATOMIC_OR, measure_scatter,
                                 parse_backend)

    n = g.num_neurons
    blocks = -(-n // cfg.block_neurons)
    L = g.max_delay + 1
    if cfg.variant == parse_backend("hip-bucket"):
        indeg = np.bincount(g.post_ids, minlength=n)
        P = int((((indeg + 31) // 32) * 32).sum())
        kind, target, what = SCATTER_ATOMIC_OR, L * P // 8, "u32 atomicOr into arrival bitmaps"
    else:
        kind, target, what = SCATTER_ATOMIC_ADD, L * n * 4, "float atomicAdd into the ring"
    ops = int(round(events_per_step))
    floor_us = measure_scatter(kind, target, ops, blocks)
    empty_us = measure_scatter(kind, target, 0, blocks)
    k_us = kernel_ms * 1e3
    return {"access": what, "ops_per_launch": ops, "target_bytes": int(target),
            "blocks": blocks, "floor_us": floor_us, "empty_launch_us": empty_us,
            "kernel_us": k_us, "frac": floor_us / k_us if k_us > 0 else None,
            "ops_per_s": ops / (k_us * 1e-6) if k_us > 0 else None,
            "floor_ops_per_s": ops / (floor_us * 1e-6) if floor_us > 0 else None}


def expansion_floor(g, hg, plan, num_steps, kernel_ms):
    """The hip-atomic step's delivery alone: the fired rows of one step (the
    step whose event count is nearest the run's mean) replayed with the step
    kernel's own expansion placement and nothing else
    (snc_hip_measure_expansion). frac = floor / kernel time: the share of the
    step that its atomics alone take; the rest is the state, injection and
    row-pointer round trips before them."""
    import numpy as np

    r = plan.run(capture_raster=True)
    n = g.num_neurons
    bits = np.unpackbits(r.raster.view(np.uint8), axis=1, bitorder="little")[:, :n]
    outdeg = np.diff(g.row_ptr).astype(np.int64)
    ev = bits.astype(np.int64) @ outdeg
    t = int(np.argmin(np.abs(ev - ev.mean())))
    fired = np.nonzero(bits[t])[0].astype(np.int32)
    us = hg.measure_expansion(fired)
    k_us = kernel_ms * 1e3
    return {"step": t, "fired_rows": int(fired.size), "events": int(ev[t]),
            "mean_events_per_step": float(ev.mean()), "floor_us": us, "kernel_us": k_us,
            "frac": us / k_us if k_us > 0 else None}


def measured_traffic(kernel_symbol, label=None):
    """HBM bytes per dispatch of a kernel from the committed counter summaries
    (profiles/*pmc_traffic*.md: calibrated FETCH_SIZE + WRITE_SIZE), or None.
    label, when given, must match the summary's workload label too."""
    import glob
    import re

    pat = re.compile(r"<!-- traffic_bytes_per_dispatch (.+) (\S+) (\d+) -->")
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*pmc_traffic*.md"))):
        for line in open(path):
            m = pat.search(line)
            if m and m.group(1) == kernel_symbol and (label is None or m.group(2) == label):
                return {"bytes_per_launch": float(m.group(3)),
                        "source": os.path.relpath(path, ROOT)}
    return None


PULL_KERNELS = ("void snc_amd::(anonymous namespace)::k_pull_step<{w}, false>(snc_amd::PullArgs, int)",
                "void snc_amd::(anonymous namespace)::k_pull_hub_chunks<{w}>(snc_amd::PullArgs, int)")


def model_bytes(neuron_steps, events, fired):
    return 16.0 * neuron_steps + 9.0 * events + 8.0 * fired


def cpu_baseline(w, repeats):
    import oracle as O

    g, ev, T, lif = w.graph, w.samples[0], w.num_steps, w.lif
    use_ref = O.ref_available() and lif.reset_mode == 0
    runs = []
    for _ in range(repeats):
        r = O.ref_forward(g, ev, T, lif) if use_ref else O.forward(g, ev, T, lif,
                                                                   capture_raster=False)
        runs.append(r)
        log(f"cpu baseline run: {r['ms']:.1f} ms, {r['events']} events")
    ms = statistics.median(r["ms"] for r in runs)
    events = runs[0]["events"]
    out = {"value": events / (ms / 1e3) / 1e9, "unit": "GSynOps/s", "cores": 1,
           "kind": "reference" if use_ref else "port",
           "sample": f"full run of the same workload ({T} steps), {repeats} repeats, median "
                     f"step-loop time (RunStats.ms semantics)",
           "ms_per_run": ms, "events_per_run": events}
    try:
        model = open("/proc/cpuinfo").read().split("model name")[1].split(":")[1].split("\n")[0]
        out["cpu_model"] = model.strip()
    except Exception:
        pass
    # OpenMP form of the same path on every host core, once.
    ncores = os.cpu_count() or 1
    t0 = time.time()
    r = O.ref_forward(g, ev, T, lif, use_omp=True) if use_ref else O.forward(
        g, ev, T, lif, use_omp=True, capture_raster=False)
    out["openmp"] = {"value": r["events"] / (r["ms"] / 1e3) / 1e9, "cores": ncores,
                     "ms_per_run": r["ms"], "wall_s": time.time() - t0}
    return out


def _train_session(w, elig):
    from snc_amd.train import HipTrainSession

    s = HipTrainSession(w.graph, w.cfg, batch=w.batch, max_steps=w.max_steps, elig=elig)
    if w.init_weights is not None:
        s.set_weights(w.init_weights)
    n = w.labels.size
    if w.kind == "features":
        s.upload_features(w.X, w.labels, w.lengths)
    else:
        s.upload_events(w.events, n, w.labels, w.lengths)
    return s


def eprop_bench(w, epochs, plain_batches, elig, cpu_samples=0):
    """Batched persistent e-prop against the plain per-sample path on one
    workload. Both start from the same weights; the plain path runs the first
    `plain_batches` minibatches of epoch 1 and its weights are compared with
    the batched path's after the same minibatches (same sums in the same
    order, so they must be bit-identical; equal weights give equal accuracy).
    samples/s are wall-clock rates of whole training steps (host loop,
    launches, synchronisation and the per-step readback included)."""
    import numpy as np
    from snc_amd.train import BATCHED, PLAIN, SERIAL

    n = w.labels.size
    tr = np.arange(w.n_train, dtype=np.int32)
    te = np.arange(w.n_train, n, dtype=np.int32)
    rng = np.random.default_rng(12345)
    order = rng.permutation(tr).astype(np.int32)
    batches = [order[i:i + w.batch] for i in range(0, order.size, w.batch)]
    nb = min(plain_batches, len(batches))

    sb = _train_session(w, elig)
    acc0 = sb.evaluate(te)
    sb.step(batches[0], epoch=0, mode=BATCHED, apply=False)   # warm-up (no update)
    synchronize_train()

    def run(sess, mode, bl, epoch):
        t0 = time.perf_counter()
        dev, samples, spikes, events, steps = 0.0, 0, 0, 0, 0
        for b in bl:
            st, _, _ = sess.step(b, epoch=epoch, mode=mode, apply=True)
            dev += st.ms
            samples += st.samples
            spikes += st.spikes
            events += st.synaptic_events
            steps += int(w.lengths[b].sum())
        return time.perf_counter() - t0, dev, samples, spikes, events, steps

    # plain path over the first nb minibatches (nb = 0: batched only)
    if nb > 0:
        sp = _train_session(w, elig)
        wall_p, dev_p, n_p, spk_p, ev_p, _ = run(sp, PLAIN, batches[:nb], 1)
        w_plain = sp.get_weights()
        acc_plain = sp.evaluate(te)
        sp.close()
        # the per-sample persistent path: the batched kernel launched once per
        # sample (a minibatch of one), same minibatches, same start weights
        ss = _train_session(w, elig)
        wall_s, dev_s, n_s, _, _, _ = run(ss, SERIAL, batches[:nb], 1)
        w_serial = ss.get_weights()
        ss.close()
    # batched path over the same minibatches, then the rest of the epochs
    wall_b, dev_b, n_b, spk_b, ev_b, st_b = run(sb, BATCHED, batches[:nb], 1)
    nmb = len(batches[:nb])
    w_batch = sb.get_weights()
    acc_same = sb.evaluate(te)
    identical = bool(np.array_equal(w_plain, w_batch)) if nb > 0 else None
    identical_serial = bool(np.array_equal(w_serial, w_batch)) if nb > 0 else None
    if nb < len(batches):
        r = run(sb, BATCHED, batches[nb:], 1)
        wall_b, dev_b, n_b = wall_b + r[0], dev_b + r[1], n_b + r[2]
        spk_b, ev_b, st_b = spk_b + r[3], ev_b + r[4], st_b + r[5]
        nmb += len(batches) - nb
    curve = [sb.evaluate(te)]
    for ep in range(2, epochs + 1):
        o = np.random.default_rng(12345 + ep).permutation(tr).astype(np.int32)
        bl = [o[i:i + w.batch] for i in range(0, o.size, w.batch)]
        r = run(sb, BATCHED, bl, ep)
        wall_b, dev_b, n_b = wall_b + r[0], dev_b + r[1], n_b + r[2]
        spk_b, ev_b, st_b = spk_b + r[3], ev_b + r[4], st_b + r[5]
        nmb += len(bl)
        curve.append(sb.evaluate(te))
    N, S = w.graph.num_neurons, w.graph.num_synapses
    phases = eprop_phases_of(sb, w, batches, elig, 2)
    sb.close()
    roofline = eprop_roofline(w, elig, dev_b, nmb, st_b, ev_b, spk_b, phases)
    bsps = n_b / wall_b
    out = {
        "workload": w.name, "neurons": w.graph.num_neurons, "synapses": w.graph.num_synapses,
        "batch": w.batch, "train_samples": int(tr.size), "test_samples": int(te.size),
        "lengths": [int(w.lengths.min()), int(w.lengths.max())],
        "eligibility": "f64" if elig == 1 else "f32",
        "batched": {"samples_per_s": bsps, "device_samples_per_s": n_b / (dev_b / 1e3),
                    "samples": n_b, "epochs": epochs, "GSynOps/s": ev_b / wall_b / 1e9,
                    "test_acc_per_epoch": curve},
        "test_acc_initial": acc0,
        "roofline": roofline,
        "phases": phases,
    }
    if nb > 0:
        psps = n_p / wall_p
        out["plain"] = {"samples_per_s": psps, "device_samples_per_s": n_p / (dev_p / 1e3),
                        "samples": n_p, "GSynOps/s": ev_p / wall_p / 1e9}
        out["speedup_batched_vs_plain"] = bsps / psps
        ssps = n_s / wall_s
        out["serial"] = {"samples_per_s": ssps, "device_samples_per_s": n_s / (dev_s / 1e3),
                         "samples": n_s,
                         "what": "one launch of the batched kernel per sample (minibatch of one)"}
        out["speedup_batched_vs_serial"] = bsps / ssps
        out["equal_accuracy_check"] = {"minibatches": nb, "weights_bit_identical": identical,
                                       "serial_weights_bit_identical": identical_serial,
                                       "test_acc_plain": acc_plain, "test_acc_batched": acc_same,
                                       "acc_diff_pt": 100.0 * (acc_same - acc_plain)}
        log(f"e-prop {w.name}: batched {bsps:.1f} samples/s, plain {psps:.1f} samples/s "
            f"(x{bsps / psps:.1f}), serial {ssps:.1f} samples/s (x{bsps / ssps:.1f}), "
            f"identical={identical}/{identical_serial}, acc {acc0:.3f} -> {curve}")
    else:
        log(f"e-prop {w.name}: batched {bsps:.1f} samples/s, acc {acc0:.3f} -> {curve}")
    out["global_T_max"] = eprop_global_t(w, epochs, elig)
    log(f"e-prop {w.name} with every sample at T_max = {w.max_steps}: "
        f"{out['global_T_max']['samples_per_s']:.1f} samples/s, "
        f"acc -> {out['global_T_max']['test_acc_per_epoch']}")
    if cpu_samples > 0:
        out["cpu_baseline"] = eprop_cpu_baseline(w, cpu_samples)
        out["speedup_batched_vs_cpu"] = bsps / out["cpu_baseline"]["value"]
    return out


def eprop_global_t(w, epochs, elig):
    """Sensitivity to the per-sample lengths: the same batched training with
    every sample run for the workload's longest length T_max (lr / T_max for
    all, the reference's global-T rule) instead of its own T_b, from the same
    start weights and minibatch orders: samples/s and held-out accuracy per
    epoch."""
    import dataclasses

    import numpy as np
    from snc_amd.train import BATCHED

    wt = dataclasses.replace(w, lengths=np.full_like(w.lengths, w.max_steps))
    s = _train_session(wt, elig)
    n = w.labels.size
    tr = np.arange(w.n_train, dtype=np.int32)
    te = np.arange(w.n_train, n, dtype=np.int32)
    wall, samples, curve = 0.0, 0, []
    for ep in range(1, epochs + 1):
        seed = 12345 if ep == 1 else 12345 + ep
        o = np.random.default_rng(seed).permutation(tr).astype(np.int32)
        t0 = time.perf_counter()
        for i in range(0, o.size, w.batch):
            st, _, _ = s.step(o[i:i + w.batch], epoch=ep, mode=BATCHED, apply=True)
            samples += st.samples
        wall += time.perf_counter() - t0
        curve.append(s.evaluate(te))
    s.close()
    return {"T": int(w.max_steps), "samples_per_s": samples / wall,
            "test_acc_per_epoch": curve}


def eprop_cpu_baseline(w, n_samples):
    """The oracle's e-prop trainer on one core: one minibatch step over the
    first n_samples training samples of the same workload (same graph, start
    weights, per-sample lengths, readout and reset; feature inputs as the
    device's Philox Poisson spikes, turned into event lists), timed."""
    import numpy as np
    import oracle as O

    c = w.cfg
    ocfg = O.train_config(num_steps=c.num_steps, lif=c.lif, lr=c.lr, w_init=c.w_init,
                          w_max=c.w_max, gamma=c.surrogate_scale, feedback=c.feedback_scale,
                          train_hidden=c.train_hidden, readout_mode=c.readout_mode,
                          readout_psi=c.readout_psi, encoder=c.encoder, gain=c.enc_gain,
                          max_rate=c.enc_max_rate, enc_threshold=c.enc_threshold, seed=c.seed)
    tr = O.Trainer(w.graph, ocfg)
    if w.init_weights is not None:
        tr.weights = w.init_weights
    lengths = np.ascontiguousarray(w.lengths[:n_samples], np.int32)
    if w.kind == "features":
        evs = []
        for i in range(n_samples):
            e = O.encode_philox(w.X[i], int(lengths[i]), gain=c.enc_gain, max_rate=c.enc_max_rate,
                                threshold=c.enc_threshold, seed=c.seed, sample=i, epoch=1)
            e["sample_id"] = i
            evs.append(e)
        ev = np.concatenate(evs)
    else:
        ev = w.events[w.events["sample_id"] < n_samples]
    t0 = time.perf_counter()
    tr.minibatch(ev, lengths, w.labels[:n_samples], apply=True)
    dt = time.perf_counter() - t0
    steps = int(lengths.sum())
    log(f"e-prop cpu baseline {w.name}: {n_samples} samples in {dt:.2f} s")
    return {"value": n_samples / dt, "unit": "samples/s", "cores": 1, "kind": "port",
            "sample": f"one minibatch step of the oracle trainer over the first {n_samples} "
                      f"training samples ({steps} sample-steps), eligibility in double",
            "seconds": dt}


FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 vector / matrix
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32
MI355X_CUS = 256


def eprop_roofline(w, elig, dev_ms, minibatches, sample_steps, events, spikes, phases):
    """The batched e-prop kernel against its ceilings. Its eligibility sums
    stay on chip (registers and LDS), so HBM is not what binds it:
      hbm            counted HBM bytes per training dispatch of k_train_batched
                     (calibrated FETCH_SIZE + WRITE_SIZE from the committed
                     counter summary) over the measured mean device time
      eligibility    2 flops per synapse per sample-step (one fused multiply-add
                     of E_s += psi_post * tr_pre, f64 or f32) against the dense
                     peak: over the whole minibatch's device time, and over the
                     CU time the sums phase itself took (phase split)
      model_bytes    the algorithmic byte model (28 N + 8 S per sample-step, ...),
                     which charges 8 bytes per synapse-step of eligibility that
                     never leaves the chip -- reported, not the bound"""
    N, S = w.graph.num_neurons, w.graph.num_synapses
    f64 = elig == 1
    peak = FP64_PEAK_TFLOPS if f64 else FP32_PEAK_TFLOPS
    ms_mb = dev_ms / max(1, minibatches)
    flops = 2.0 * S * sample_steps
    tflops = flops / (dev_ms / 1e3) / 1e12
    # the sums phase: flops of the profiled minibatches over the CU time the
    # phase took in them, scaled to the whole chip (MI355X_CUS CUs)
    sums_cu_ms = phases["phases_mean_us"]["eligibility"] * w.batch / 1e3 * phases["minibatches"]
    prof_flops = 2.0 * S * phases["sample_steps"]
    phase_tflops = (prof_flops / (sums_cu_ms / 1e3) / 1e12 * MI355X_CUS
                    if sums_cu_ms > 0 else None)
    ring = "true" if w.name.startswith("c2") else "false"
    sym = f"snc_amd::k_train_batched<{'double' if f64 else 'float'}, {ring}, false>"
    traffic = measured_traffic(sym)
    model = ((28.0 * N + 8.0 * S) * sample_steps + 9.0 * events + 8.0 * spikes
             + (4.0 * S * w.batch + 8.0 * S) * minibatches)
    out = {"kernel": sym, "device_ms_per_minibatch": ms_mb,
           "eligibility": {"unit": "TFLOP/s", "flops": flops, "achieved": tflops, "peak": peak,
                           "frac": tflops / peak, "precision": "f64" if f64 else "f32",
                           "sums_phase_cu_ms": sums_cu_ms,
                           "sums_phase_achieved": phase_tflops,
                           "sums_phase_frac": phase_tflops / peak if phase_tflops else None},
           "model_bytes": {"bytes": model, "gbs": model / (dev_ms / 1e3) / 1e9,
                           "frac": model / (dev_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                           "note": "byte model of the reference's design (eligibility in HBM); "
                                   "this kernel keeps eligibility on chip, so this is not its "
                                   "bound"}}
    if traffic:
        kern = [v for v in phases["kernel_ms"] if v > 0]
        k_ms = sum(kern) / len(kern) if kern else ms_mb
        gbs = traffic["bytes_per_launch"] / (k_ms / 1e3) / 1e9
        out["hbm"] = {"unit": "GB/s", "traffic_bytes_per_dispatch": traffic["bytes_per_launch"],
                      "source": traffic["source"], "kernel_ms": k_ms, "achieved": gbs,
                      "peak": HBM_PEAK_GBS, "frac": gbs / HBM_PEAK_GBS}
    return out


def eprop_phases_of(s, w, batches, elig, minibatches):
    """Phase split (snc_hip_train_profile) of the first minibatches on an open
    session: per phase, the mean time per sample and the largest time one
    workgroup spent in it, and each launch's device time."""
    import numpy as np

    from snc_amd.train import PHASE_NAMES

    runs = [s.profile(b, epoch=1) for b in batches[:minibatches]]
    out = {"minibatches": len(runs), "kernel_ms": [r["kernel_ms"] for r in runs],
           "sample_steps": int(sum(int(w.lengths[b].sum()) for b in batches[:minibatches])),
           "max_steps_per_minibatch": [int(w.lengths[b].max()) for b in batches[:minibatches]],
           "phases_mean_us": {k: float(np.mean([r["phases"][k]["mean_us"] for r in runs]))
                              for k in PHASE_NAMES},
           "phases_max_us": {k: float(np.mean([r["phases"][k]["max_us"] for r in runs]))
                             for k in PHASE_NAMES}}
    log(f"e-prop phases {w.name} ({'f64' if elig == 1 else 'f32'}): kernel {out['kernel_ms']} "
        "ms; mean us " + ", ".join(f"{k} {v:.0f}" for k, v in out["phases_mean_us"].items()))
    return out


def eprop_train_only(w, elig, minibatches):
    """Only training minibatches (no evaluation, no warm-up launch without
    update), for counter runs: every k_train_batched dispatch is a training
    dispatch of the workload's minibatch size."""
    import numpy as np

    from snc_amd.train import BATCHED

    order = np.random.default_rng(12345).permutation(np.arange(w.n_train, dtype=np.int32))
    s = _train_session(w, elig)
    for i in range(minibatches):
        s.step(order[i * w.batch:(i + 1) * w.batch], epoch=1, mode=BATCHED, apply=True)
    s.close()
    log(f"e-prop {w.name}: {minibatches} training minibatches")


def eprop_phases(w, elig, minibatches=3):
    """Where the batched e-prop kernel's time goes on one workload: the phase
    split (snc_hip_train_profile) of the first `minibatches` training
    minibatches of epoch 1, after one untimed warm-up launch."""
    import numpy as np

    from snc_amd.train import BATCHED

    order = np.random.default_rng(12345).permutation(np.arange(w.n_train, dtype=np.int32))
    batches = [order[i:i + w.batch].astype(np.int32) for i in range(0, order.size, w.batch)]
    s = _train_session(w, elig)
    s.step(batches[0], epoch=0, mode=BATCHED, apply=False)
    out = {"workload": w.name, "batch": w.batch, "eligibility": "f64" if elig == 1 else "f32"}
    out.update(eprop_phases_of(s, w, batches, elig, minibatches))
    s.close()
    return out


def batched_forward(w, hg, B, steps, warmup, variant):
    """The delivery-stress graph with B independent drives in one batched plan
    (one launch per step covers every sample): events/s and the step
    kernel's algorithmic bandwidth."""
    import numpy as np
    from snc_amd import inputs as I
    from snc_amd.runtime import ForwardPlan, parse_backend, synchronize

    g = w.graph
    drives = [w.samples[0]] + [I.background_drive(g.num_neurons, w.num_steps, 0.01, 1.0,
                                                  seed=100 + b) for b in range(1, B)]
    allev = np.concatenate([d.copy() for d in drives])
    off = 0
    for b, d in enumerate(drives):
        allev["sample_id"][off:off + d.size] = b
        off += d.size
    t0 = time.time()
    plan = ForwardPlan(hg, allev, w.num_steps, w.lif, parse_backend(variant), num_samples=B)
    plan_s = time.time() - t0
    r = plan.run()
    events, spikes = r.stats.synaptic_events, r.stats.spikes
    for _ in range(warmup):
        plan.launch()
    synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        plan.launch()
    synchronize()
    wall = (time.perf_counter() - t1) / max(1, steps)
    prof = plan.profile()
    units = kernel_units(g, r.neuron_spikes.sum(axis=0), w.num_steps)
    kernels = {}
    for name, (tot, k) in prof.items():
        nu, ev_u, f_u = units[name]
        by = model_bytes(B * nu, ev_u, f_u) / k  # neuron terms scale with B
        kernels[name] = {"avg_ms": tot / k, "launches": k, "bytes_per_launch": by,
                         "gbs": by / (tot / k / 1e3) / 1e9}
    step_bytes = model_bytes(B * g.num_neurons * w.num_steps, events, spikes)
    cfgp = plan.config
    out = {"samples": B, "variant": plan.variant.name,
           "plan": {"block_neurons": cfgp.block_neurons, "heavy_threshold": cfgp.heavy_threshold},
           "plan_build_s": plan_s, "ms_per_run": wall * 1e3,
           "GSynOps/s": events / wall / 1e9, "events_per_run": events,
           "whole_run": {"gbs": step_bytes / wall / 1e9,
                         "frac": step_bytes / wall / 1e9 / HBM_PEAK_GBS},
           "kernels": kernels}
    if plan.variant == parse_backend("hip-pull"):
        # counted HBM traffic of a step (the hub-chunk launch plus the step
        # launch), from the committed counter summary of this workload
        word = "unsigned long" if B > 32 else "unsigned int"
        parts = [measured_traffic(k.format(w=word), f"hip-pull_x{B}") for k in PULL_KERNELS]
        if all(parts):
            by = sum(p["bytes_per_launch"] for p in parts)
            out["counted_traffic"] = {"bytes_per_step": by, "source": parts[0]["source"],
                                      "gbs": by * w.num_steps / wall / 1e9,
                                      "frac": by * w.num_steps / wall / 1e9 / HBM_PEAK_GBS}
    log(f"batched forward x{B} ({out['variant']}): {wall * 1e3:.2f} ms/run, "
        f"{out['GSynOps/s']:.2f} GSynOps/s, whole-run {out['whole_run']['frac'] * 100:.1f}% of HBM")
    plan.close()
    return out


def synchronize_train():
    from snc_amd.runtime import synchronize

    synchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--variant", default="hip-auto")
    ap.add_argument("--block", type=int, default=0,
                    help="with an explicit --variant: neurons per step-kernel block")
    ap.add_argument("--hub-threshold", type=int, default=512,
                    help="with --block: hip-bucket hub in-degree threshold")
    ap.add_argument("--compare", default="hip-atomic,hip-bucket,hip-sort,hip-auto-exact",
                    help="other variants to time (reported under 'variants')")
    ap.add_argument("--width", type=int, default=500)
    ap.add_argument("--height", type=int, default=400)
    ap.add_argument("--timesteps", type=int, default=1000)
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--eprop-epochs", type=int, default=2)
    ap.add_argument("--plain-batches", type=int, default=8,
                    help="minibatches the plain per-sample path is timed on (0: skip it)")
    ap.add_argument("--batch-samples", default="8,32,64",
                    help="sample counts of the batched-forward lines (empty: none)")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched-forward lines (--batch-samples x --batched-variants)")
    ap.add_argument("--batched-variants", default="hip-auto,hip-auto-exact,hip-pull")
    ap.add_argument("--no-eprop", action="store_true")
    ap.add_argument("--no-grow", action="store_true",
                    help="skip the learned-vs-heuristic grow comparison")
    ap.add_argument("--eprop-only", action="store_true")
    ap.add_argument("--eprop-train-only", type=int, default=0, metavar="K",
                    help="for counter runs: K training minibatches of each training workload "
                         "(f64 eligibility), nothing else, then exit")
    ap.add_argument("--eprop-phases", action="store_true",
                    help="print the batched e-prop kernel's phase split on both training "
                         "workloads (f64 and f32 eligibility) and exit")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="for counter runs: a 256 MiB device copy (the calibration of the "
                         "traffic counters), then one run of --variant on the workload "
                         "(with --eprop-only: the e-prop measurements instead)")
    ap.add_argument("--profile-only", action="store_true",
                    help="build the workload and run K steps without extra measurements")
    ap.add_argument("--score-only", action="store_true",
                    help="time the grow MLP scoring 2^20 candidate descriptors, K calls after "
                         "W warm-up calls, and print that line alone")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the spiking path here runs on a single MI355X "
              "(multi-GPU sharding is design-only); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from snc_amd import configs as K
    from snc_amd.runtime import (ForwardPlan, HipGraph, PlanConfig, available, parse_backend,
                                 synchronize)

    if not available():
        raise SystemExit("HIP backend unavailable (no gfx950 device or library not built)")

    def eprop_all():
        res = {}
        t = time.time()
        # CPU leg: one minibatch of each workload (256 and 128 samples)
        res["c2"] = eprop_bench(K.c2(), args.eprop_epochs, args.plain_batches, 1,
                                0 if args.no_cpu else 256)
        res["c3"] = eprop_bench(K.c3(), max(3, args.eprop_epochs), args.plain_batches, 1,
                                0 if args.no_cpu else 128)
        log(f"e-prop measurements: {time.time() - t:.1f}s")
        return res

    if args.pmc_probe:
        # A plain device-to-device copy of 256 MiB through the HIP runtime
        # (its copy kernel): 256 MiB read + 256 MiB written per dispatch.
        import ctypes as C

        hip = C.CDLL("libamdhip64.so")
        hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        hip.hipFree.argtypes = [C.c_void_p]
        nbytes = 256 << 20
        src, dst = C.c_void_p(), C.c_void_p()
        if hip.hipMalloc(C.byref(src), nbytes) or hip.hipMalloc(C.byref(dst), nbytes):
            raise SystemExit("hipMalloc failed")
        for _ in range(3):
            if hip.hipMemcpy(dst, src, nbytes, 3):  # hipMemcpyDeviceToDevice
                raise SystemExit("hipMemcpy failed")
        hip.hipDeviceSynchronize()
        hip.hipFree(src)
        hip.hipFree(dst)
    if args.score_only:
        import numpy as np

        from snc_amd.grow import N_FEATURES, GrowMLP

        m = 1 << 20
        mlp = GrowMLP(seed=1)
        X = np.random.default_rng(0).standard_normal((m, N_FEATURES), dtype=np.float32)
        for _ in range(args.warmup):
            mlp.score(X)
        kms = []
        for _ in range(args.steps):
            mlp.score(X)
            kms.append(mlp.last_kernel_ms)
        k = sum(kms) / len(kms)
        nbytes = (2 * N_FEATURES + 4) * m  # bf16 features in, fp32 score out
        flops = 2 * (16 * 64 + 64 * 64 + 64) * m  # the three layers' multiply-adds
        print(json.dumps({"grow_score": {
            "candidates": m, "calls": args.steps, "kernel_ms": k,
            "candidates_per_s": m / (k * 1e-3), "GB_per_s": nbytes / (k * 1e-3) / 1e9,
            "TFLOP_per_s": flops / (k * 1e-3) / 1e12}}))
        return
    if args.eprop_only:
        print(json.dumps({"eprop": eprop_all()}))
        return
    if args.eprop_train_only:
        for mk in (K.c2, K.c3):
            eprop_train_only(mk(), 1, args.eprop_train_only)
        return
    if args.eprop_phases:
        print(json.dumps({"eprop_phases": [eprop_phases(mk(), el) for mk in (K.c2, K.c3)
                                           for el in (1, 0)]}))
        return
    t0 = time.time()
    w = K.c1(args.width, args.height, args.timesteps)
    g = w.graph
    log(f"workload {w.name}: generated in {time.time() - t0:.1f}s, "
        f"{w.samples[0].size} drive events")
    hg = HipGraph(g)

    def batched_lines():
        res = {}
        for bs in [int(x) for x in args.batch_samples.split(",") if x]:
            for vn in [v for v in args.batched_variants.split(",") if v]:
                res[f"{vn}_x{bs}"] = batched_forward(w, hg, bs, max(1, args.steps // 2), 1, vn)
        return res

    if args.batched_only:
        print(json.dumps({"batched_forward": batched_lines()}))
        return

    def make_plan(name):
        if args.block and name == args.variant and not name.startswith("hip-auto"):
            cfg = PlanConfig(parse_backend(name), args.block, args.hub_threshold, False)
            return ForwardPlan(hg, w.samples[0], w.num_steps, w.lif, config=cfg)
        return ForwardPlan(hg, w.samples[0], w.num_steps, w.lif, parse_backend(name))

    t_plan = time.time()
    plan = make_plan(args.variant)
    plan_s = time.time() - t_plan
    ref = plan.run()
    events, spikes = ref.stats.synaptic_events, ref.stats.spikes
    log(f"variant {args.variant} -> {plan.variant.name} {plan.config}: spikes={spikes} "
        f"events={events} run={ref.stats.ms:.2f} ms (plan {plan_s:.1f}s)")
    autotune = None
    if args.variant in ("hip-auto", "hip-auto-exact"):
        rep = hg.tune_report(exact_only=args.variant == "hip-auto-exact")
        autotune = {"search_ms": rep["search_ms"], "trial_steps": rep["trial_steps"],
                    "chosen": rep["chosen"],
                    "candidates": [{"variant": c.variant.name, "block_neurons": c.block_neurons,
                                    "heavy_threshold": c.heavy_threshold,
                                    "persistent": c.persistent, "ms_per_step": m}
                                   for c, m in rep["candidates"]]}

    for _ in range(args.warmup):
        plan.launch()
    synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        plan.launch()
    synchronize()
    wall = time.perf_counter() - t1
    ms_per_step = wall * 1e3 / max(1, args.steps)
    value = events * args.steps / wall / 1e9
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, {value:.3f} GSynOps/s")
    if args.profile_only:
        print(json.dumps({"metric": METRIC, "value": value, "unit": "GSynOps/s",
                          "ms_per_step": ms_per_step, "steps": args.steps}))
        return

    prof = plan.profile()
    units = kernel_units(g, ref.neuron_spikes, w.num_steps)
    kernels = {}
    for name, (tot, k) in prof.items():
        by = model_bytes(*units[name]) / k
        avg = tot / k
        kernels[name] = {"avg_ms": avg, "launches": k, "total_ms": tot,
                         "bytes_per_launch": by,
                         "gbs": (by / (avg / 1e3) / 1e9) if avg > 0 else 0.0}
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    dk = kernels[dom]
    if len(kernels) == 1:
        # One kernel per step (fused) or per run (persistent): its average
        # launch duration is the graph-launched run's HIP-event time on the
        # plan's stream over its launches. The per-kernel profile above
        # times eager launches with a synchronisation after each, which adds
        # the launch latency; it stays in kernels[...]["eager_avg_ms"].
        plan.run()
        k_avg, k_n = plan.kernel_time()
        if k_avg > 0 and k_n == dk["launches"]:
            dk["eager_avg_ms"] = dk["avg_ms"]
            dk["avg_ms"] = k_avg
            dk["total_ms"] = k_avg * k_n
            dk["gbs"] = dk["bytes_per_launch"] / (k_avg / 1e3) / 1e9
    cfgp = plan.config
    sym = (f"snc_amd::k_step<{'true' if cfgp.variant == parse_backend('hip-bucket') else 'false'}, "
           f"{cfgp.block_neurons}>" if dom == "k_step" else f"snc_amd::{dom}")
    traffic = measured_traffic(sym)
    step_bytes = 16.0 * g.num_neurons * w.num_steps + 9.0 * events + 8.0 * spikes
    roofline = {"bound": "hbm", "kernel": dom, "achieved": dk["gbs"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": dk["gbs"] / HBM_PEAK_GBS,
                "traffic": traffic["bytes_per_launch"] if traffic else None,
                "traffic_source": traffic["source"] if traffic else None,
                "bytes_per_launch": dk["bytes_per_launch"], "avg_launch_ms": dk["avg_ms"],
                "kernels": kernels,
                "whole_run": {"bytes": step_bytes, "ms": ms_per_step,
                              "gbs": step_bytes / (ms_per_step / 1e3) / 1e9,
                              "frac": step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS}}
    if dom == "k_step":
        roofline["scatter_floor"] = scatter_floor(g, cfgp, events / w.num_steps, dk["avg_ms"])
        log(f"scatter floor (random addresses): {roofline['scatter_floor']}")
        if cfgp.variant == parse_backend("hip-atomic"):
            roofline["expansion_floor"] = expansion_floor(g, hg, plan, w.num_steps, dk["avg_ms"])
            log(f"expansion floor (this step's own deposits): {roofline['expansion_floor']}")

    variants = {}
    for name in [v for v in args.compare.split(",") if v]:
        p = make_plan(name)
        r = p.run()
        p.launch()
        synchronize()
        tt = time.perf_counter()
        for _ in range(max(1, args.steps)):
            p.launch()
        synchronize()
        wv = (time.perf_counter() - tt) / max(1, args.steps)
        variants[name] = {"GSynOps/s": r.stats.synaptic_events / wv / 1e9, "ms_per_step": wv * 1e3,
                          "spikes": r.stats.spikes, "events": r.stats.synaptic_events,
                          "resolved": p.variant.name}
        log(f"variant {name}: {variants[name]}")
        p.close()

    out = {"metric": METRIC, "value": value, "unit": "GSynOps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": w.name, "neurons": g.num_neurons, "synapses": g.num_synapses,
                      "timesteps": w.num_steps, "batch": 1, "variant": plan.variant.name,
                      "plan": {"block_neurons": plan.config.block_neurons,
                               "heavy_threshold": plan.config.heavy_threshold,
                               "persistent": plan.config.persistent},
                      "spikes_per_run": spikes, "events_per_run": events,
                      "parallelism": "single-gpu"},
           "roofline": roofline, "variants": variants, "autotune": autotune,
           "memory": {"graph_bytes": hg.memory(), "plan_bytes": plan.memory(),
                      "graph_bytes_per_synapse": hg.memory() / max(1, g.num_synapses),
                      "plan_bytes_per_neuron": plan.memory() / max(1, g.num_neurons),
                      "overflow": ref.stats.overflow}}
    out["batched_forward"] = batched_lines()
    if not args.no_eprop:
        out["eprop"] = eprop_all()
    if not args.no_grow:
        from snc_amd.growlab import ab_compare

        out["grow_ab"] = ab_compare(log=log)
        out["grow_ab_sheet"] = ab_compare(log=log, kind="sheet")
        # the structural co-training workload as stated: the batched e-prop
        # network under 3000 synapses, the whole candidate pool scored
        out["grow_ab_c4"] = ab_compare(log=log, kind="c4", label="rank", w_grow=-1.0,
                                       onpolicy=1)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_repeats)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
