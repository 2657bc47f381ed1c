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

Byte model (algorithmic, per step): 16 B per neuron (membrane read+write,
ring slot read+clear), 9 B per synaptic event (post id 4, weight 4, delay 1),
8 B per fired neuron (row_ptr pair). The LIF kernel is credited the neuron
term, the expansion kernel the event and fired terms.
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


def kernel_bytes(name, n, events, spikes, launches):
    """Algorithmic bytes per launch of one kernel class over a run."""
    if launches <= 0:
        return 0.0
    if name == "k_lif":
        return 16.0 * n
    if name == "k_expand":
        return (9.0 * events + 8.0 * spikes) / launches
    if name == "k_persistent":
        return (16.0 * n * launches_steps[0] + 9.0 * events + 8.0 * spikes) / launches
    if name == "sort_step":
        return (16.0 * n * launches + 9.0 * events + 8.0 * spikes) / launches
    if name == "k_fold_heavy":
        return 0.0
    return 0.0


launches_steps = [0]


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--variant", default="hip-auto")
    ap.add_argument("--compare", default="hip-atomic,hip-bucket",
                    help="other variants to time (reported under 'variants')")
    ap.add_argument("--width", type=int, default=500)
    ap.add_argument("--height", type=int, default=400)
    ap.add_argument("--timesteps", type=int, default=1000)
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="build the workload and run K steps without extra measurements")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the spiking path here runs on a single MI355X "
              "(multi-GPU sharding is design-only); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from snc_amd import configs as K
    from snc_amd.runtime import ForwardPlan, HipGraph, available, parse_backend, synchronize

    if not available():
        raise SystemExit("HIP backend unavailable (no gfx950 device or library not built)")
    t0 = time.time()
    w = K.c1(args.width, args.height, args.timesteps)
    g = w.graph
    log(f"workload {w.name}: generated in {time.time() - t0:.1f}s, "
        f"{w.samples[0].size} drive events")
    hg = HipGraph(g)

    def make_plan(name):
        return ForwardPlan(hg, w.samples[0], w.num_steps, w.lif, parse_backend(name))

    plan = make_plan(args.variant)
    ref = plan.run()
    events, spikes = ref.stats.synaptic_events, ref.stats.spikes
    log(f"variant {args.variant} -> {plan.variant.name}: spikes={spikes} events={events} "
        f"run={ref.stats.ms:.2f} ms")

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
    launches_steps[0] = w.num_steps
    kernels = {}
    for name, (tot, k) in prof.items():
        by = kernel_bytes(name, g.num_neurons, events, spikes, k)
        avg = tot / k
        kernels[name] = {"avg_ms": avg, "launches": k, "total_ms": tot,
                         "bytes_per_launch": by,
                         "gbs": (by / (avg / 1e3) / 1e9) if avg > 0 else 0.0}
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    dk = kernels[dom]
    step_bytes = 16.0 * g.num_neurons * w.num_steps + 9.0 * events + 8.0 * spikes
    roofline = {"bound": "hbm", "kernel": dom, "achieved": dk["gbs"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": dk["gbs"] / HBM_PEAK_GBS, "traffic": None,
                "bytes_per_launch": dk["bytes_per_launch"], "avg_launch_ms": dk["avg_ms"],
                "kernels": kernels,
                "whole_run": {"bytes": step_bytes, "ms": ms_per_step,
                              "gbs": step_bytes / (ms_per_step / 1e3) / 1e9,
                              "frac": step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS}}

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
                          "spikes": r.stats.spikes, "events": r.stats.synaptic_events}
        log(f"variant {name}: {variants[name]}")
        p.close()

    out = {"metric": METRIC, "value": value, "unit": "GSynOps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": w.name, "neurons": g.num_neurons, "synapses": g.num_synapses,
                      "timesteps": w.num_steps, "batch": 1, "variant": plan.variant.name,
                      "spikes_per_step": spikes, "events_per_step": events,
                      "parallelism": "single-gpu"},
           "roofline": roofline, "variants": variants}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_repeats)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
