// snc_hip_train -- the snc_train driver (src/training/train_main.cpp) on the
// MI355X backend: build a graph and a synthetic dataset, train e-prop on the
// GPU (batched persistent kernel or the plain per-sample path), print one
// line per epoch.
//
//   snc_hip_train --dataset prototype --structure sheet --epochs 3 --batch 256
//   snc_hip_train --dataset events --structure recurrent --lr 3 --epochs 5
//   snc_hip_train --train-mode plain ...      (same math, one sample at a time)
//   snc_hip_train --train-mode serial ...     (the persistent kernel once per sample)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <numeric>
#include <random>

#include "cli_common.hpp"

using namespace snc_cli;

namespace {

struct Options {
  std::string dataset = "prototype", structure = "sheet", train_mode = "batched";
  std::string readout = "leaky", reset = "subtract", encoder = "poisson", elig = "f64", log_csv;
  int epochs = 3, batch = 256, num_steps = 200, min_steps = 50, num_train = 2048,
      num_test = 512, hidden = 512, dim = 784, classes = 10, refractory = 1, fan_in = 32,
      rec_fan_in = 56, reward_mode = 0;
  float lr = 1.0f, w_init = 0.0f, w_max = 4.0f, gamma = 0.3f, feedback = 1.0f, threshold = 1.0f,
        decay = 0.95122942f, gain = 1.0f, max_rate = 0.2f, noise = 0.2f, w_std = 0.2f;
  uint64_t seed = 1;
};

void usage() {
  std::printf(
      "snc_hip_train -- e-prop training on the MI355X backend\n"
      "  --dataset prototype|synthetic|events   --structure sheet|static-snc|recurrent\n"
      "  --train-mode batched|plain|serial  --epochs N  --batch N  --num-steps T (max)  --min-steps T\n"
      "  --num-train N  --num-test N  --hidden N  --dim N  --classes N\n"
      "  --fan-in K  --rec-fan-in K   (sheet: input and recurrent partners per neuron)\n"
      "  --lr F  --w-init F (0: the graph's weights)  --w-max F  --w-std F  --gamma F\n"
      "  --feedback F  --threshold F  --decay F  --refractory R  --reset value|subtract\n"
      "  --readout spiking|leaky  --encoder direct|poisson  --gain F  --max-rate F\n"
      "  --noise F  --elig f64|f32  --seed S  --log-csv FILE\n"
      "  --reward-mode 0|1   (three-factor reward rule instead of the supervised delta)\n");
}

Options parse(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      usage();
      std::exit(0);
    }
    if (take(i, argc, argv, "--dataset", o.dataset) ||
        take(i, argc, argv, "--structure", o.structure) ||
        take(i, argc, argv, "--train-mode", o.train_mode) ||
        take(i, argc, argv, "--readout", o.readout) || take(i, argc, argv, "--reset", o.reset) ||
        take(i, argc, argv, "--encoder", o.encoder) || take(i, argc, argv, "--elig", o.elig) ||
        take(i, argc, argv, "--log-csv", o.log_csv) || take(i, argc, argv, "--epochs", o.epochs) ||
        take(i, argc, argv, "--batch", o.batch) || take(i, argc, argv, "--num-steps", o.num_steps) ||
        take(i, argc, argv, "--min-steps", o.min_steps) ||
        take(i, argc, argv, "--num-train", o.num_train) ||
        take(i, argc, argv, "--num-test", o.num_test) || take(i, argc, argv, "--hidden", o.hidden) ||
        take(i, argc, argv, "--dim", o.dim) || take(i, argc, argv, "--classes", o.classes) ||
        take(i, argc, argv, "--refractory", o.refractory) || take(i, argc, argv, "--lr", o.lr) ||
        take(i, argc, argv, "--fan-in", o.fan_in) ||
        take(i, argc, argv, "--rec-fan-in", o.rec_fan_in) ||
        take(i, argc, argv, "--w-init", o.w_init) || take(i, argc, argv, "--w-max", o.w_max) ||
        take(i, argc, argv, "--w-std", o.w_std) || take(i, argc, argv, "--gamma", o.gamma) ||
        take(i, argc, argv, "--feedback", o.feedback) ||
        take(i, argc, argv, "--threshold", o.threshold) ||
        take(i, argc, argv, "--decay", o.decay) || take(i, argc, argv, "--gain", o.gain) ||
        take(i, argc, argv, "--max-rate", o.max_rate) || take(i, argc, argv, "--noise", o.noise) ||
        take(i, argc, argv, "--seed", o.seed) ||
        take(i, argc, argv, "--reward-mode", o.reward_mode))
      continue;
    std::fprintf(stderr, "unknown option: %s\n", a.c_str());
    usage();
    std::exit(2);
  }
  return o;
}

Graph build_graph(const Options& o, int n_in) {
  snc_graph_buf* b = nullptr;
  const int32_t layers[3] = {n_in, o.hidden, o.classes};
  if (o.structure == "sheet") {
    const float side = std::sqrt((float)n_in);
    snc_sheet_params p{3, layers, o.fan_in, o.rec_fan_in, 1, 3.0f, side, o.w_std, 1, 1, o.seed};
    check(snc_gen_sheet(&p, &b), "sheet graph");
  } else if (o.structure == "static-snc") {
    check(snc_gen_static_snc(layers, 3, 50000, 0.5f, 1, 0.2f, o.seed, &b), "static-snc graph");
  } else if (o.structure == "recurrent") {
    check(snc_gen_random_recurrent(n_in, o.hidden, o.classes, 0.1f, 0.1f, 1, 31,
                                   1.0f / std::sqrt(0.1f * o.hidden), o.seed, &b),
          "recurrent graph");
  } else {
    std::fprintf(stderr, "unknown structure: %s\n", o.structure.c_str());
    std::exit(2);
  }
  return Graph(b);
}

}  // namespace

int main(int argc, char** argv) {
  Options o = parse(argc, argv);
  if (!snc_hip_available()) {
    std::fprintf(stderr, "no usable gfx950 device: %s\n", snc_hip_last_error());
    return 3;
  }
  const int n = o.num_train + o.num_test;
  std::mt19937_64 rng(o.seed * 0x9E3779B97F4A7C15ULL + 7);
  std::vector<int32_t> lengths(n), labels(n);
  std::vector<float> X;
  Events ev;
  int n_in = o.dim;
  if (o.dataset == "events") {
    n_in = 700;
    snc_event_buf* eb = nullptr;
    check(snc_gen_event_streams(n, n_in, o.classes, 20, 15.0f, 5000, std::min(o.min_steps, o.num_steps),
                                o.num_steps, o.seed + 30, lengths.data(), labels.data(), &eb),
          "event streams");
    ev = Events(eb);
  } else {
    X.resize((size_t)n * o.dim);
    if (o.dataset == "prototype") {
      std::vector<uint8_t> U((size_t)n * o.dim);
      snc_gen_prototype_images(n, o.dim, o.classes, 0.7f, o.noise, o.seed + 10, U.data(),
                               labels.data());
      for (size_t k = 0; k < U.size(); ++k) X[k] = (float)U[k] / 255.0f;
    } else {
      snc_make_synthetic(n, o.dim, o.classes, o.noise, o.seed, X.data(), labels.data());
    }
    std::uniform_int_distribution<int> len(std::min(o.min_steps, o.num_steps), o.num_steps);
    for (int i = 0; i < n; ++i) lengths[i] = len(rng);
  }
  Graph g = build_graph(o, n_in);
  const snc_train_config cfg{o.num_steps,
                             {o.threshold, o.decay, 0.0f, o.refractory,
                              o.reset == "subtract" ? SNC_RESET_SUBTRACT : SNC_RESET_TO_VALUE},
                             o.lr, o.w_init, o.w_max, o.gamma, o.feedback, 1,
                             o.readout == "leaky" ? SNC_READOUT_LEAKY : SNC_READOUT_SPIKING,
                             1.0f, o.encoder == "direct" ? SNC_ENC_DIRECT : SNC_ENC_POISSON,
                             o.gain, o.max_rate, 0.0f, o.seed, o.reward_mode != 0 ? 1 : 0,
                             0.99f};
  snc_hip_train* s = nullptr;
  check(snc_hip_train_create(&g.view, &cfg, o.batch, o.num_steps,
                             o.elig == "f32" ? SNC_ELIG_F32 : SNC_ELIG_F64, &s),
        "train session");
  if (o.w_init == 0.0f) {  // start from the graph's signed weights
    check(snc_hip_train_set_weights(s, g.view.weights, g.view.num_synapses), "set weights");
  }
  if (o.dataset == "events")
    check(snc_hip_train_upload_events(s, ev.data, ev.n, n, labels.data(), lengths.data()),
          "upload events");
  else
    check(snc_hip_train_upload_features(s, X.data(), n, o.dim, labels.data(), lengths.data()),
          "upload features");
  std::vector<int32_t> tr(o.num_train), te(o.num_test);
  std::iota(tr.begin(), tr.end(), 0);
  std::iota(te.begin(), te.end(), o.num_train);
  const int mode = o.train_mode == "plain"    ? SNC_TRAIN_PLAIN
                   : o.train_mode == "serial" ? SNC_TRAIN_SERIAL
                                              : SNC_TRAIN_BATCHED;
  if (mode == SNC_TRAIN_BATCHED && o.train_mode != "batched") {
    std::fprintf(stderr, "unknown --train-mode: %s\n", o.train_mode.c_str());
    return 2;
  }
  std::printf("graph: neurons=%d synapses=%d  dataset=%s train=%d test=%d  mode=%s batch=%d\n",
              g.view.num_neurons, g.view.num_synapses, o.dataset.c_str(), o.num_train,
              o.num_test, o.train_mode.c_str(), o.batch);
  std::ofstream csv;
  if (!o.log_csv.empty()) {
    csv.open(o.log_csv);
    csv << "epoch,train_loss,train_acc,test_acc,samples_per_s,spikes,synaptic_events\n";
  }
  snc_train_stats e0{};
  check(snc_hip_train_evaluate(s, te.data(), o.num_test, mode, &e0), "evaluate");
  std::printf("epoch  0  test_acc=%.4f\n", (double)e0.correct / std::max<int64_t>(1, e0.samples));
  for (int ep = 1; ep <= o.epochs; ++ep) {
    snc_train_stats st{};
    const auto t0 = std::chrono::steady_clock::now();
    check(snc_hip_train_epoch(s, tr.data(), o.num_train, ep, mode, &st), "train epoch");
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    snc_train_stats et{};
    check(snc_hip_train_evaluate(s, te.data(), o.num_test, mode, &et), "evaluate");
    const double tacc = (double)et.correct / std::max<int64_t>(1, et.samples);
    std::printf("epoch %2d  loss=%.4f train_acc=%.4f test_acc=%.4f  %.0f samples/s  "
                "spikes=%lld events=%lld\n",
                ep, st.loss_sum / std::max<int64_t>(1, st.samples),
                (double)st.correct / std::max<int64_t>(1, st.samples), tacc, st.samples / secs,
                (long long)st.spikes, (long long)st.synaptic_events);
    if (csv)
      csv << ep << ',' << st.loss_sum / std::max<int64_t>(1, st.samples) << ','
          << (double)st.correct / std::max<int64_t>(1, st.samples) << ',' << tacc << ','
          << st.samples / secs << ',' << st.spikes << ',' << st.synaptic_events << '\n';
  }
  snc_hip_train_destroy(s);
  return 0;
}
