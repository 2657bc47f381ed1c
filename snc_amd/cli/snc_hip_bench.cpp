// snc_hip_bench -- the snc_bench harness (src/bench/snc_bench.cpp) on the
// MI355X backend: build a graph by structure, encode a synthetic dataset,
// run every sample through forward() on a hip-* backend, report spikes,
// synaptic events, throughput and argmax accuracy of the untrained pass.
//
//   snc_hip_bench --structure static-snc --encoder poisson --backend hip-auto
//                 --num-steps 50 --num-samples 64
//   snc_hip_bench --compare-backends      (every hip-* backend, parity vs hip-bucket)
//   snc_hip_bench --batched               (all samples in one batched plan)
//   snc_hip_bench --self-test             (3-neuron graph, closed-form outcome)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <map>

#include "cli_common.hpp"

using namespace snc_cli;

namespace {

struct Options {
  std::string encoder = "poisson", backend = "hip-auto", structure = "static-snc", log_json;
  std::string reset = "value";
  int num_steps = 50, num_samples = 64, hidden = 128, dim = 784, classes = 10;
  int synapse_budget = 50000, delay = 1, refractory = 1;
  float threshold = 1.0f, decay = 0.9f, inhib = 0.2f, gain = 1.0f, max_rate = 0.5f;
  uint64_t seed = 1;
  bool compare = false, self_test = false, batched = false;
};

void usage() {
  std::printf(
      "snc_hip_bench -- SNN forward benchmark on the MI355X backend\n"
      "  --encoder    direct|poisson|latency     (default poisson)\n"
      "  --backend    hip-atomic|hip-bucket|hip-sort|hip-pull|hip-auto|hip-auto-exact\n"
      "  --structure  dense|random-sparse|static-snc\n"
      "  --num-steps N  --num-samples N  --hidden N  --dim N  --classes N\n"
      "  --synapse-budget S  --delay D  --threshold T  --decay D  --refractory R\n"
      "  --reset value|subtract  --inhib F  --gain G  --max-rate P  --seed S\n"
      "  --batched            run all samples in one batched plan\n"
      "  --compare-backends   run every hip-* backend, report parity and speedup\n"
      "  --self-test          3-neuron graph with a closed-form outcome\n"
      "  --log-json FILE\n");
}

Options parse(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      usage();
      std::exit(0);
    }
    if (a == "--compare-backends") { o.compare = true; continue; }
    if (a == "--self-test") { o.self_test = true; continue; }
    if (a == "--batched") { o.batched = true; continue; }
    if (take(i, argc, argv, "--encoder", o.encoder) || take(i, argc, argv, "--backend", o.backend) ||
        take(i, argc, argv, "--structure", o.structure) ||
        take(i, argc, argv, "--log-json", o.log_json) || take(i, argc, argv, "--reset", o.reset) ||
        take(i, argc, argv, "--num-steps", o.num_steps) ||
        take(i, argc, argv, "--num-samples", o.num_samples) ||
        take(i, argc, argv, "--hidden", o.hidden) || take(i, argc, argv, "--dim", o.dim) ||
        take(i, argc, argv, "--classes", o.classes) ||
        take(i, argc, argv, "--synapse-budget", o.synapse_budget) ||
        take(i, argc, argv, "--delay", o.delay) ||
        take(i, argc, argv, "--refractory", o.refractory) ||
        take(i, argc, argv, "--threshold", o.threshold) ||
        take(i, argc, argv, "--decay", o.decay) || take(i, argc, argv, "--inhib", o.inhib) ||
        take(i, argc, argv, "--gain", o.gain) || take(i, argc, argv, "--max-rate", o.max_rate) ||
        take(i, argc, argv, "--seed", o.seed))
      continue;
    std::fprintf(stderr, "unknown option: %s\n", a.c_str());
    usage();
    std::exit(2);
  }
  return o;
}

Graph build_graph(const Options& o) {
  const int32_t layers[3] = {o.dim, o.hidden, o.classes};
  snc_graph_buf* b = nullptr;
  if (o.structure == "dense")
    check(snc_gen_dense(layers, 3, 0.5f, o.delay, &b), "dense graph");
  else if (o.structure == "random-sparse")
    check(snc_gen_random_sparse(layers, 3, o.synapse_budget, 0.5f, o.delay, o.inhib, o.seed, &b),
          "random-sparse graph");
  else if (o.structure == "static-snc")
    check(snc_gen_static_snc(layers, 3, o.synapse_budget, 0.5f, o.delay, o.inhib, o.seed, &b),
          "static-snc graph");
  else {
    std::fprintf(stderr, "unknown structure: %s\n", o.structure.c_str());
    std::exit(2);
  }
  return Graph(b);
}

struct Eval {
  int64_t spikes = 0, events = 0;
  double ms = 0.0;
  int correct = 0, total = 0;
  int used = -1;
  int64_t overflow = 0;
  int64_t graph_bytes = 0, plan_bytes = 0;  // device memory (largest plan)
  std::vector<int64_t> per_sample_spikes;
};

int encoder_id(const std::string& e) {
  if (e == "direct") return 0;
  if (e == "poisson") return 1;
  if (e == "latency") return 2;
  std::fprintf(stderr, "unknown encoder: %s\n", e.c_str());
  std::exit(2);
}

// One forward() per sample (snc_bench.cpp:185-211): encode with seed + i,
// argmax over the output-channel spike counts.
Eval evaluate(const Options& o, const Graph& g, DeviceGraph& dg, const std::vector<float>& X,
              const std::vector<int32_t>& y, int variant) {
  const snc_lif lif{o.threshold, o.decay, 0.0f, o.refractory,
                    o.reset == "subtract" ? SNC_RESET_SUBTRACT : SNC_RESET_TO_VALUE};
  const int C = g.num_output_channels();
  const int enc = encoder_id(o.encoder);
  Eval r;
  r.total = o.num_samples;
  std::vector<snc_event> all;
  std::vector<std::vector<snc_event>> per(o.num_samples);
  for (int i = 0; i < o.num_samples; ++i) {
    snc_event_buf* eb = nullptr;
    check(snc_encode(enc, X.data() + (size_t)i * o.dim, o.dim, o.num_steps, o.gain, o.max_rate,
                     0.0f, o.seed + i, o.batched ? i : 0, &eb),
          "encode");
    Events ev(eb);
    per[i].assign(ev.data, ev.data + ev.n);
  }
  std::vector<float> logits;
  if (o.batched) {
    for (auto& p : per) all.insert(all.end(), p.begin(), p.end());
    snc_hip_plan* p = nullptr;
    check(snc_hip_plan_create_batch(dg.h, all.data(), (int64_t)all.size(), o.num_samples,
                                    o.num_steps, &lif, variant, nullptr, &p),
          "batched plan");
    logits.assign((size_t)o.num_samples * std::max(C, 1), 0.0f);
    snc_forward_capture cap{logits.data(), nullptr, nullptr, nullptr, nullptr};
    snc_forward_stats st{};
    check(snc_hip_plan_run(p, &cap, &st), "forward");
    r.per_sample_spikes.assign(o.num_samples, 0);
    std::vector<int64_t> ev(o.num_samples);
    check(snc_hip_plan_sample_stats(p, r.per_sample_spikes.data(), ev.data()), "sample stats");
    r.spikes = st.spikes;
    r.events = st.synaptic_events;
    r.ms = st.ms;
    r.used = st.used_variant;
    r.overflow = st.overflow;
    check(snc_hip_plan_memory(p, &r.plan_bytes), "plan memory");
    snc_hip_plan_free(p);
  } else {
    logits.assign((size_t)o.num_samples * std::max(C, 1), 0.0f);
    for (int i = 0; i < o.num_samples; ++i) {
      snc_hip_plan* p = nullptr;
      check(snc_hip_plan_create(dg.h, per[i].data(), (int64_t)per[i].size(), o.num_steps, &lif,
                                variant, &p),
            "plan");
      snc_forward_capture cap{logits.data() + (size_t)i * std::max(C, 1), nullptr, nullptr,
                              nullptr, nullptr};
      snc_forward_stats st{};
      check(snc_hip_plan_run(p, &cap, &st), "forward");
      r.spikes += st.spikes;
      r.events += st.synaptic_events;
      r.ms += st.ms;
      r.used = st.used_variant;
      r.overflow += st.overflow;
      r.per_sample_spikes.push_back(st.spikes);
      int64_t pb = 0;
      check(snc_hip_plan_memory(p, &pb), "plan memory");
      r.plan_bytes = std::max(r.plan_bytes, pb);
      snc_hip_plan_free(p);
    }
  }
  check(snc_hip_graph_memory(dg.h, &r.graph_bytes), "graph memory");
  for (int i = 0; i < o.num_samples; ++i) {
    const float* l = logits.data() + (size_t)i * std::max(C, 1);
    int pred = 0;
    float best = -1.0f;
    for (int c = 0; c < C; ++c)
      if (l[c] > best) {
        best = l[c];
        pred = c;
      }
    if (pred == y[i]) ++r.correct;
  }
  return r;
}

// a -> b (0.3, delay 2), b -> c (0.7, delay 4, b inhibitory), a -> c (0.5,
// delay 1); a is input channel 0, c output channel 0. Drive a with 2.0 at
// step 0: a fires at 0, c receives 0.5 at step 1, b receives 0.3 at step 2,
// neither crosses 1.0: one spike, two synaptic events, c ends at
// 0.5 * 0.9^(T-2) after T steps.
int self_test() {
  const int32_t row_ptr[4] = {0, 2, 3, 3}, post[3] = {1, 2, 2}, delays[3] = {2, 1, 4};
  const float w[3] = {0.3f, 0.5f, 0.7f};
  const int8_t sign[3] = {1, -1, 1};
  const uint8_t role[3] = {SNC_ROLE_INPUT, SNC_ROLE_INTERNAL, SNC_ROLE_OUTPUT};
  const int32_t channel[3] = {0, -1, 0};
  const snc_graph_view v{3, 3, row_ptr, post, w, delays, sign, role, channel};
  bool ok = snc_graph_validate(&v) == 1;  // 1: valid
  DeviceGraph dg(v);
  const snc_event ev{0, 0, 0, 2.0f};
  const snc_lif lif{1.0f, 0.9f, 0.0f, 1, SNC_RESET_TO_VALUE};
  const int T = 6;
  bool all = true;
  for (int variant : {SNC_VARIANT_ATOMIC, SNC_VARIANT_BUCKET, SNC_VARIANT_SORT}) {
    snc_hip_plan* p = nullptr;
    check(snc_hip_plan_create(dg.h, &ev, 1, T, &lif, variant, &p), "plan");
    float logits[1] = {0.0f}, v_out[3] = {0.0f, 0.0f, 0.0f};
    int32_t nspk[3] = {0, 0, 0};
    snc_forward_capture cap{logits, nspk, nullptr, v_out, nullptr};
    snc_forward_stats st{};
    check(snc_hip_plan_run(p, &cap, &st), "forward");
    snc_hip_plan_free(p);
    // c: 0.5 arrives at step 1, then decays four times
    float vc = 0.0f;
    for (int t = 0; t < T; ++t) {
      vc = vc * 0.9f;
      if (t == 1) vc = vc + 0.5f;
    }
    const bool good = st.spikes == 1 && st.synaptic_events == 2 && nspk[0] == 1 &&
                      nspk[1] == 0 && nspk[2] == 0 && logits[0] == 0.0f &&
                      std::fabs(v_out[2] - vc) <= 1e-6f * std::fabs(vc);
    std::printf("self-test %-10s spikes=%lld events=%lld v_c=%.7f (want %.7f): %s\n",
                variant_name(variant), (long long)st.spikes, (long long)st.synaptic_events,
                v_out[2], vc, good ? "ok" : "WRONG");
    all = all && good;
  }
  ok = ok && all;
  std::printf("self-test: %s\n", ok ? "PASS" : "FAIL");
  return ok ? 0 : 1;
}

void report(const Options& o, const snc_graph_view& g, const Eval& r) {
  const double secs = r.ms / 1000.0;
  std::printf("\n=== snc_hip_bench ===\n");
  std::printf("structure=%s encoder=%s backend=%s(->%s) steps=%d samples=%d%s\n",
              o.structure.c_str(), o.encoder.c_str(), o.backend.c_str(), variant_name(r.used),
              o.num_steps, r.total, o.batched ? " (batched plan)" : "");
  std::printf("graph: neurons=%d synapses=%d\n", g.num_neurons, g.num_synapses);
  std::printf("sim:   spikes=%lld synaptic_events=%lld time=%.3f ms\n", (long long)r.spikes,
              (long long)r.events, r.ms);
  std::printf("rate:  %.3g spikes/s  %.3g events/s\n", secs > 0 ? r.spikes / secs : 0.0,
              secs > 0 ? r.events / secs : 0.0);
  std::printf("ml:    argmax-acc=%.3f (%d/%d, untrained forward pass)\n",
              r.total ? (double)r.correct / r.total : 0.0, r.correct, r.total);
  std::printf("mem:   graph=%.2f MB (%.1f B/synapse)  plan=%.2f MB (%.1f B/neuron)  "
              "overflow=%lld\n",
              r.graph_bytes / 1e6, g.num_synapses ? (double)r.graph_bytes / g.num_synapses : 0.0,
              r.plan_bytes / 1e6, g.num_neurons ? (double)r.plan_bytes / g.num_neurons : 0.0,
              (long long)r.overflow);
}

void write_json(const Options& o, const snc_graph_view& g, const Eval& r) {
  std::ofstream f(o.log_json);
  if (!f) {
    std::fprintf(stderr, "could not open %s\n", o.log_json.c_str());
    return;
  }
  f << "{\n  \"structure\": \"" << o.structure << "\",\n  \"encoder\": \"" << o.encoder
    << "\",\n  \"backend\": \"" << variant_name(r.used) << "\",\n  \"num_steps\": "
    << o.num_steps << ",\n  \"num_samples\": " << r.total << ",\n  \"neurons\": "
    << g.num_neurons << ",\n  \"synapses\": " << g.num_synapses << ",\n  \"spikes\": "
    << r.spikes << ",\n  \"synaptic_events\": " << r.events << ",\n  \"sim_ms\": " << r.ms
    << ",\n  \"argmax_acc\": " << (r.total ? (double)r.correct / r.total : 0.0)
    << ",\n  \"spikes_per_s\": " << (r.ms > 0 ? r.spikes / (r.ms / 1e3) : 0.0)
    << ",\n  \"events_per_s\": " << (r.ms > 0 ? r.events / (r.ms / 1e3) : 0.0)
    << ",\n  \"overflow\": " << r.overflow << ",\n  \"device_graph_bytes\": " << r.graph_bytes
    << ",\n  \"device_plan_bytes\": " << r.plan_bytes << ",\n  \"bytes_per_synapse\": "
    << (g.num_synapses ? (double)r.graph_bytes / g.num_synapses : 0.0)
    << ",\n  \"bytes_per_neuron\": "
    << (g.num_neurons ? (double)r.plan_bytes / g.num_neurons : 0.0)
    << ",\n  \"per_sample_spikes\": [";
  for (size_t i = 0; i < r.per_sample_spikes.size(); ++i)
    f << (i ? ", " : "") << r.per_sample_spikes[i];
  f << "]\n}\n";
  std::printf("wrote %s\n", o.log_json.c_str());
}

}  // namespace

int main(int argc, char** argv) {
  Options o = parse(argc, argv);
  if (!snc_hip_available()) {
    std::fprintf(stderr, "no usable gfx950 device: %s\n", snc_hip_last_error());
    return 3;
  }
  if (o.self_test) return self_test();
  std::vector<float> X((size_t)o.num_samples * o.dim);
  std::vector<int32_t> y(o.num_samples);
  snc_make_synthetic(o.num_samples, o.dim, o.classes, 0.15f, o.seed, X.data(), y.data());
  Graph g = build_graph(o);
  if (snc_graph_validate(&g.view) != 1) {
    std::fprintf(stderr, "graph validation failed: %s\n", snc_hip_last_error());
    return 1;
  }
  DeviceGraph dg(g.view);

  if (o.compare) {
    std::printf("graph: neurons=%d synapses=%d\n", g.view.num_neurons, g.view.num_synapses);
    std::printf("\nbackend           spikes      events     time(ms)  speedup  match\n");
    Eval ref;
    bool have = false;
    for (const char* bn : {"hip-bucket", "hip-sort", "hip-atomic", "hip-auto-exact", "hip-auto"}) {
      const Eval r = evaluate(o, g, dg, X, y, parse_variant(bn));
      if (!have) {
        ref = r;
        have = true;
      }
      const bool same = r.per_sample_spikes == ref.per_sample_spikes && r.events == ref.events;
      std::printf("%-16s %10lld %11lld %10.3f %8.2f  %s\n", bn, (long long)r.spikes,
                  (long long)r.events, r.ms, r.ms > 0 ? ref.ms / r.ms : 0.0,
                  same ? "exact" : "DIFF");
    }
    return 0;
  }
  const int variant = parse_variant(o.backend);
  if (variant < 0) {
    std::fprintf(stderr, "unknown backend: %s\n", o.backend.c_str());
    return 2;
  }
  const Eval r = evaluate(o, g, dg, X, y, variant);
  report(o, g.view, r);
  if (!o.log_json.empty()) write_json(o, g.view, r);
  return 0;
}
