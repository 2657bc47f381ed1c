// snc_hip_cotrain -- the two-timescale co-training driver
// (src/training/cotrain_main.cpp:236-285) on the MI355X backend: per outer
// round, compile the connectome to a frozen graph, train e-prop on the GPU,
// carry the weights back, and run one structural update (prune the weakest,
// grow) from the round's activity. --grow learned scores the grow candidates
// with the MLP, trained first on the GPU from grow outcomes logged by
// heuristic runs (--log-runs of them, exploring a share --explore of grows).
//
// --structure sheet runs the 2-D recurrent sheet connectome instead of the
// layered one (--fan-in, --rec-fan-in, --radius; signed N(0, --w-std) start
// weights), with the structural step on the GPU unless --host-step.
//
//   snc_hip_cotrain --grow heuristic
//   snc_hip_cotrain --grow learned --log-runs 3 --explore 0.5
//   snc_hip_cotrain --structure sheet --grow-n 300 --locality 0.1
#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <random>

#include "cli_common.hpp"

using namespace snc_cli;

namespace {

struct Options {
  std::string grow = "heuristic", structure = "layered";
  int fan_in = 32, rec_fan_in = 8, host_step = 0, reward = 0, reward_struct = 0;
  float radius = 0.2f, w_std = 0.2f;
  int outer = 12, inner = 1, batch = 256, num_steps = 200, min_steps = 50, num_train = 2048,
      num_test = 512, hidden = 512, dim = 784, classes = 10, budget = 3000, grow_n = 200,
      protect = 2, log_runs = 3, mlp_epochs = 40;
  float lr = 1.0f, w_init = 0.3f, w_grow = 0.02f, locality = 0.2f, noise = 0.2f,
        explore = 0.5f, mlp_lr = 2e-3f;
  uint64_t seed = 1;
};

void usage() {
  std::printf(
      "snc_hip_cotrain -- structural co-training on the MI355X backend\n"
      "  --grow heuristic|learned  --outer N  --inner N  --structural-budget S  --grow-n K\n"
      "  --protect R  --locality F  --w-grow F  --batch N  --num-steps T  --min-steps T\n"
      "  --num-train N  --num-test N  --hidden N  --dim N  --classes N  --lr F  --w-init F\n"
      "  --noise F  --seed S\n"
      "  learned: --log-runs N  --explore F  --mlp-epochs N  --mlp-lr F\n"
      "  --structure layered|sheet; sheet: --fan-in N  --rec-fan-in N  --radius F  --w-std F\n"
      "  --host-step  --reward 0|1 (three-factor weight rule)  --reward-struct 0|1\n");
}

Options parse(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      usage();
      std::exit(0);
    }
    if (a == "--host-step") {
      o.host_step = 1;
      continue;
    }
    if (take(i, argc, argv, "--grow", o.grow) || take(i, argc, argv, "--outer", o.outer) ||
        take(i, argc, argv, "--inner", o.inner) || take(i, argc, argv, "--batch", o.batch) ||
        take(i, argc, argv, "--num-steps", o.num_steps) ||
        take(i, argc, argv, "--min-steps", o.min_steps) ||
        take(i, argc, argv, "--num-train", o.num_train) ||
        take(i, argc, argv, "--num-test", o.num_test) || take(i, argc, argv, "--hidden", o.hidden) ||
        take(i, argc, argv, "--dim", o.dim) || take(i, argc, argv, "--classes", o.classes) ||
        take(i, argc, argv, "--structural-budget", o.budget) ||
        take(i, argc, argv, "--grow-n", o.grow_n) || take(i, argc, argv, "--protect", o.protect) ||
        take(i, argc, argv, "--log-runs", o.log_runs) ||
        take(i, argc, argv, "--mlp-epochs", o.mlp_epochs) || take(i, argc, argv, "--lr", o.lr) ||
        take(i, argc, argv, "--w-init", o.w_init) || take(i, argc, argv, "--w-grow", o.w_grow) ||
        take(i, argc, argv, "--locality", o.locality) || take(i, argc, argv, "--noise", o.noise) ||
        take(i, argc, argv, "--explore", o.explore) || take(i, argc, argv, "--mlp-lr", o.mlp_lr) ||
        take(i, argc, argv, "--seed", o.seed) ||
        take(i, argc, argv, "--structure", o.structure) ||
        take(i, argc, argv, "--fan-in", o.fan_in) ||
        take(i, argc, argv, "--rec-fan-in", o.rec_fan_in) ||
        take(i, argc, argv, "--radius", o.radius) || take(i, argc, argv, "--w-std", o.w_std) ||
        take(i, argc, argv, "--reward", o.reward) ||
        take(i, argc, argv, "--reward-struct", o.reward_struct))
      continue;
    std::fprintf(stderr, "unknown option: %s\n", a.c_str());
    usage();
    std::exit(2);
  }
  if (o.structure != "layered" && o.structure != "sheet") {
    std::fprintf(stderr, "unknown structure: %s\n", o.structure.c_str());
    std::exit(2);
  }
  return o;
}

struct Data {
  std::vector<float> X;
  std::vector<int32_t> y, lengths, tr, te;
  double train_steps = 0.0;
};

Data make_data(const Options& o) {
  Data d;
  const int n = o.num_train + o.num_test;
  std::vector<uint8_t> U((size_t)n * o.dim);
  d.y.resize(n);
  snc_gen_prototype_images(n, o.dim, o.classes, 0.7f, o.noise, o.seed + 10, U.data(), d.y.data());
  d.X.resize(U.size());
  for (size_t k = 0; k < U.size(); ++k) d.X[k] = (float)U[k] / 255.0f;
  std::mt19937_64 rng(o.seed * 0x9E3779B97F4A7C15ULL + 11);
  std::uniform_int_distribution<int> len(std::min(o.min_steps, o.num_steps), o.num_steps);
  d.lengths.resize(n);
  for (int i = 0; i < n; ++i) d.lengths[i] = len(rng);
  d.tr.resize(o.num_train);
  d.te.resize(o.num_test);
  std::iota(d.tr.begin(), d.tr.end(), 0);
  std::iota(d.te.begin(), d.te.end(), o.num_train);
  for (int i = 0; i < o.num_train; ++i) d.train_steps += d.lengths[i];
  return d;
}

struct RunResult {
  double test_acc = 0.0;
  int synapses = 0;
};

// One co-training run (cotrain_main.cpp:247-285 with GPU sessions).
RunResult run(const Options& o, const Data& d, uint64_t seed, int grow_mode, snc_hip_mlp* model,
              bool log, float explore, snc_connectome** keep, bool verbose) {
  const int32_t layers[3] = {o.dim, o.hidden, o.classes};
  snc_connectome* c = nullptr;
  const bool sheet = o.structure == "sheet";
  if (sheet) {
    check(snc_conn_sheet_recurrent(layers, 3, o.fan_in, o.rec_fan_in, o.radius, 1, o.w_grow, seed,
                                   &c),
          "sheet connectome");
    std::vector<float> w0(snc_conn_num_synapses(c));
    std::mt19937_64 wr(seed + 1000);
    std::normal_distribution<float> nd(0.0f, o.w_std);
    for (float& x : w0) x = nd(wr);
    check(snc_conn_set_edge_weights(c, w0.data(), (int64_t)w0.size()), "start weights");
  } else {
    check(snc_conn_layered_local(layers, 3, o.budget, 1, 0.02f, seed, &c), "connectome");
  }
  // reward_chance is chance accuracy, as cotrain_main.cpp:217-218 sets it.
  const snc_struct_config sc{o.grow_n, 1, o.w_grow, o.locality, o.protect, 8,
                             o.reward_struct != 0 ? 1 : 0,
                             o.classes ? 1.0f / (float)o.classes : 0.1f, 0.1f, seed, explore,
                             sheet && !o.host_step ? 1 : 0};
  const snc_train_config cfg{o.num_steps, {1.0f, 0.95122942f, 0.0f, 1, SNC_RESET_SUBTRACT},
                             o.lr, o.w_init, 4.0f, 0.3f, 1.0f, 1, SNC_READOUT_LEAKY, 1.0f,
                             SNC_ENC_POISSON, 1.0f, 0.2f, 0.0f, seed, o.reward != 0 ? 1 : 0,
                             0.99f};
  RunResult res;
  int epoch = 0;
  for (int r = 0; r < o.outer; ++r) {
    snc_graph_buf* gb = nullptr;
    check(snc_conn_to_graph(c, &gb), "compile graph");
    Graph g(gb);
    snc_hip_train* s = nullptr;
    check(snc_hip_train_create(&g.view, &cfg, o.batch, o.num_steps, SNC_ELIG_F64, &s), "session");
    std::vector<float> w(g.view.num_synapses);
    if (r > 0 || sheet) {  // the sheet's signed start weights from structural epoch 0
      check(snc_conn_edge_weights(c, w.data(), (int64_t)w.size()), "edge weights");
      check(snc_hip_train_set_weights(s, w.data(), (int64_t)w.size()), "set weights");
    }
    const int n = o.num_train + o.num_test;
    check(snc_hip_train_upload_features(s, d.X.data(), n, o.dim, d.y.data(), d.lengths.data()),
          "upload");
    check(snc_hip_train_reset_stats(s), "reset stats");
    snc_train_stats st{};
    for (int k = 0; k < o.inner; ++k)
      check(snc_hip_train_epoch(s, d.tr.data(), o.num_train, ++epoch, SNC_TRAIN_BATCHED, &st),
            "train epoch");
    check(snc_hip_train_get_weights(s, w.data(), (int64_t)w.size()), "get weights");
    check(snc_conn_set_edge_weights(c, w.data(), (int64_t)w.size()), "carry weights");
    std::vector<int64_t> syn(g.view.num_synapses), neur(g.view.num_neurons);
    check(snc_hip_train_collect_stats(s, syn.data(), neur.data()), "collect stats");
    snc_train_stats et{};
    check(snc_hip_train_evaluate(s, d.te.data(), o.num_test, SNC_TRAIN_BATCHED, &et), "evaluate");
    snc_hip_train_destroy(s);
    res.test_acc = (double)et.correct / std::max<int64_t>(1, et.samples);
    const double train_acc = (double)st.correct / std::max<int64_t>(1, st.samples);
    res.synapses = snc_conn_num_synapses(c);
    if (r < o.outer - 1) {
      snc_struct_report rep{};
      check(snc_conn_structural_update(c, syn.data(), (int64_t)syn.size(), neur.data(),
                                       (int64_t)neur.size(), &sc, (float)train_acc, grow_mode,
                                       model, d.train_steps, log ? 1 : 0, &rep),
            "structural update");
      if (verbose)
        std::printf("round %2d  synapses=%d train_acc=%.4f test_acc=%.4f  pruned=%d grown=%d "
                    "scored=%lld (%.3f ms)\n",
                    r, res.synapses, train_acc, res.test_acc, rep.pruned, rep.grown,
                    (long long)rep.candidates_scored, rep.score_ms);
    } else if (verbose) {
      std::printf("round %2d  synapses=%d train_acc=%.4f test_acc=%.4f\n", r, res.synapses,
                  train_acc, res.test_acc);
    }
  }
  if (keep)
    *keep = c;
  else
    snc_conn_free(c);
  return res;
}

}  // namespace

int main(int argc, char** argv) {
  Options o = parse(argc, argv);
  if (!snc_hip_available()) {
    std::fprintf(stderr, "no usable gfx950 device: %s\n", snc_hip_last_error());
    return 3;
  }
  const Data d = make_data(o);
  snc_hip_mlp* model = nullptr;
  int mode = SNC_GROW_HEURISTIC;
  if (o.grow == "learned") {
    // Grow outcomes from heuristic runs with exploration (seeds 101, 102, ...).
    std::vector<float> F;
    std::vector<float> Y;
    for (int k = 0; k < o.log_runs; ++k) {
      snc_connectome* c = nullptr;
      run(o, d, 101 + k, SNC_GROW_HEURISTIC, nullptr, true, o.explore, &c, false);
      const int64_t m = snc_conn_grow_log_size(c);
      std::vector<float> f((size_t)m * 16);
      std::vector<int32_t> y(m);
      check(snc_conn_grow_log(c, f.data(), y.data(), m), "grow log");
      for (int64_t i = 0; i < m; ++i)
        if (y[i] >= 0) {
          F.insert(F.end(), f.begin() + i * 16, f.begin() + (i + 1) * 16);
          Y.push_back((float)y[i]);
        }
      snc_conn_free(c);
    }
    check(snc_hip_mlp_create(11, &model), "MLP");
    double loss = 0.0;
    check(snc_hip_mlp_train(model, F.data(), Y.data(), (int64_t)Y.size(), o.mlp_epochs, 256,
                            o.mlp_lr, 5, &loss),
          "MLP training");
    std::printf("grow log: %zu labelled events, MLP trained on the GPU, BCE %.4f\n", Y.size(),
                loss);
    mode = SNC_GROW_LEARNED;
  } else if (o.grow != "heuristic") {
    std::fprintf(stderr, "unknown grow mode: %s\n", o.grow.c_str());
    return 2;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const RunResult r = run(o, d, o.seed, mode, model, false, 0.0f, nullptr, true);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("final: grow=%s synapses=%d test_acc=%.4f  (%.1f s)\n", o.grow.c_str(), r.synapses,
              r.test_acc, secs);
  snc_hip_mlp_destroy(model);
  return 0;
}
