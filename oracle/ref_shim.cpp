// C-ABI driver over the reference's own CPU implementation -- TEST
// INFRASTRUCTURE ONLY. Compiled by oracle/Makefile together with the
// reference's CPU sources, where they lie under the reference tree, into
// oracle/_ref/libsnc_ref.so. It includes only the reference's public headers
// (include/snc/*.hpp) and calls only its public API:
//   snc::forward            include/snc/runtime.hpp:57-58
//   snc::encode             include/snc/encoders.hpp:42-45
//   snc::make_synthetic     include/snc/dataset.hpp:27
//   snc::make_static_snc    include/snc/snn_graph.hpp:128-130
//   snc::Trainer            include/snc/trainer.hpp:59-107
//   snc::Connectome         include/snc/connectome.hpp:56-97
// Nothing here restates the reference's algorithms; it only marshals arrays.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "snc/connectome.hpp"
#include "snc/dataset.hpp"
#include "snc/encoders.hpp"
#include "snc/runtime.hpp"
#include "snc/snn_graph.hpp"
#include "snc/trainer.hpp"

namespace {

snc::SNNGraph graph_from_arrays(int n, int m, const int32_t* row_ptr, const int32_t* post,
                                const float* w, const int32_t* d, const int8_t* sign,
                                const uint8_t* role, const int32_t* channel) {
  snc::SNNGraph g;
  g.num_neurons = n;
  g.row_ptr.assign(row_ptr, row_ptr + n + 1);
  g.post_ids.assign(post, post + m);
  g.weights.assign(w, w + m);
  g.delays.assign(d, d + m);
  g.branch_ids.assign(m, 0);
  g.sign.assign(sign, sign + n);
  g.role.resize(n);
  for (int i = 0; i < n; ++i) g.role[i] = static_cast<snc::GraphRole>(role[i]);
  g.channel.assign(channel, channel + n);
  g.rebuild_io_index();
  return g;
}

struct RefGraph {
  snc::SNNGraph g;
};

}  // namespace

extern "C" {

// ---- graphs ---------------------------------------------------------------
void* ref_make_static_snc(const int32_t* layers, int32_t n_layers, int32_t budget,
                          float weight, int32_t delay, float inhib, uint64_t seed) {
  std::vector<int> L(layers, layers + n_layers);
  auto* r = new RefGraph{snc::make_static_snc(L, budget, weight, delay, inhib, seed)};
  return r;
}
void* ref_make_random_sparse(const int32_t* layers, int32_t n_layers, int32_t budget,
                             float weight, int32_t delay, float inhib, uint64_t seed) {
  std::vector<int> L(layers, layers + n_layers);
  return new RefGraph{snc::make_random_sparse(L, budget, weight, delay, inhib, seed)};
}
void* ref_make_dense(const int32_t* layers, int32_t n_layers, float weight, int32_t delay) {
  std::vector<int> L(layers, layers + n_layers);
  return new RefGraph{snc::make_dense(L, weight, delay)};
}
void ref_graph_sizes(const void* h, int32_t* n, int32_t* m) {
  const auto& g = static_cast<const RefGraph*>(h)->g;
  *n = g.num_neurons;
  *m = g.num_synapses();
}
void ref_graph_copy(const void* h, int32_t* row_ptr, int32_t* post, float* w, int32_t* d,
                    int8_t* sign, uint8_t* role, int32_t* channel) {
  const auto& g = static_cast<const RefGraph*>(h)->g;
  std::copy(g.row_ptr.begin(), g.row_ptr.end(), row_ptr);
  std::copy(g.post_ids.begin(), g.post_ids.end(), post);
  std::copy(g.weights.begin(), g.weights.end(), w);
  std::copy(g.delays.begin(), g.delays.end(), d);
  std::copy(g.sign.begin(), g.sign.end(), sign);
  for (int i = 0; i < g.num_neurons; ++i) role[i] = static_cast<uint8_t>(g.role[i]);
  std::copy(g.channel.begin(), g.channel.end(), channel);
}
int ref_graph_validate(int32_t n, int32_t m, const int32_t* row_ptr, const int32_t* post,
                       const float* w, const int32_t* d, const int8_t* sign,
                       const uint8_t* role, const int32_t* channel) {
  std::string err;
  return graph_from_arrays(n, m, row_ptr, post, w, d, sign, role, channel).validate(err) ? 1 : 0;
}
void ref_graph_free(void* h) { delete static_cast<RefGraph*>(h); }

// ---- forward --------------------------------------------------------------
// backend: 0 cpu, 1 openmp. logits must hold num_output_channels floats.
int ref_forward(int32_t n, int32_t m, const int32_t* row_ptr, const int32_t* post,
                const float* w, const int32_t* d, const int8_t* sign, const uint8_t* role,
                const int32_t* channel, const int32_t* ev4, int64_t n_ev, int32_t steps,
                float thr, float decay, float reset, int32_t refractory, int32_t backend,
                float* logits, int64_t* spikes, int64_t* events, double* ms) {
  snc::SNNGraph g = graph_from_arrays(n, m, row_ptr, post, w, d, sign, role, channel);
  std::vector<snc::InputEvent> ev((size_t)n_ev);
  for (int64_t k = 0; k < n_ev; ++k) {
    ev[k].sample_id = ev4[4 * k + 0];
    ev[k].channel_id = ev4[4 * k + 1];
    ev[k].time = ev4[4 * k + 2];
    std::memcpy(&ev[k].amplitude, &ev4[4 * k + 3], sizeof(float));
  }
  snc::LIFParams lif;
  lif.threshold = thr;
  lif.decay = decay;
  lif.reset = reset;
  lif.refractory = refractory;
  snc::ForwardResult r =
      snc::forward(g, ev, steps, lif, backend ? snc::Backend::OpenMP : snc::Backend::Cpu);
  if (logits) std::copy(r.logits.begin(), r.logits.end(), logits);
  *spikes = r.stats.spikes;
  *events = r.stats.synaptic_events;
  if (ms) *ms = r.stats.ms;
  return (int)r.logits.size();
}

// ---- encoders / data ------------------------------------------------------
int64_t ref_encode(int32_t enc, const float* x, int32_t n_ch, int32_t steps, float gain,
                   float max_rate, float thr, uint64_t seed, int32_t sample_id, int32_t* out4,
                   int64_t cap) {
  snc::EncoderParams p;
  p.gain = gain;
  p.max_rate = max_rate;
  p.threshold = thr;
  auto ev = snc::encode(static_cast<snc::Encoder>(enc), x, n_ch, steps, p, seed, sample_id);
  const int64_t k = std::min<int64_t>(cap, (int64_t)ev.size());
  for (int64_t i = 0; i < k; ++i) {
    out4[4 * i + 0] = ev[i].sample_id;
    out4[4 * i + 1] = ev[i].channel_id;
    out4[4 * i + 2] = ev[i].time;
    std::memcpy(&out4[4 * i + 3], &ev[i].amplitude, sizeof(float));
  }
  return (int64_t)ev.size();
}

void ref_make_synthetic(int32_t n, int32_t dim, int32_t classes, float noise, uint64_t seed,
                        float* X, int32_t* y) {
  snc::Dataset d = snc::make_synthetic(n, dim, classes, noise, seed);
  for (int i = 0; i < n; ++i) {
    std::copy(d.x[i].begin(), d.x[i].end(), X + (size_t)i * dim);
    y[i] = d.y[i];
  }
}

// ---- trainer --------------------------------------------------------------
struct RefTrainer {
  snc::SNNGraph g;
  snc::Trainer* tr;
};

// enc: 0 direct, 1 poisson, 2 latency.
void* ref_trainer_create(int32_t n, int32_t m, const int32_t* row_ptr, const int32_t* post,
                         const float* w, const int32_t* d, const int8_t* sign,
                         const uint8_t* role, const int32_t* channel, int32_t num_steps,
                         float thr, float decay, float reset, int32_t refractory,
                         int32_t enc, float gain, float max_rate, float enc_thr, float lr,
                         float w_init, float w_max, float gamma, float feedback,
                         int32_t train_hidden, uint64_t seed, int32_t reward_mode,
                         float reward_baseline_decay) {
  auto* r = new RefTrainer{graph_from_arrays(n, m, row_ptr, post, w, d, sign, role, channel),
                           nullptr};
  snc::TrainConfig cfg;
  cfg.num_steps = num_steps;
  cfg.lif.threshold = thr;
  cfg.lif.decay = decay;
  cfg.lif.reset = reset;
  cfg.lif.refractory = refractory;
  cfg.encoder = static_cast<snc::Encoder>(enc);
  cfg.enc.gain = gain;
  cfg.enc.max_rate = max_rate;
  cfg.enc.threshold = enc_thr;
  cfg.lr = lr;
  cfg.w_init = w_init;
  cfg.w_max = w_max;
  cfg.surrogate_scale = gamma;
  cfg.feedback_scale = feedback;
  cfg.train_hidden = train_hidden != 0;
  cfg.seed = seed;
  cfg.reward_mode = reward_mode != 0;
  cfg.reward_baseline_decay = reward_baseline_decay;
  r->tr = new snc::Trainer(r->g, cfg);
  return r;
}
void ref_trainer_free(void* h) {
  auto* r = static_cast<RefTrainer*>(h);
  delete r->tr;
  delete r;
}
void ref_trainer_get_weights(void* h, float* w) {
  const auto& v = static_cast<RefTrainer*>(h)->tr->weights();
  std::copy(v.begin(), v.end(), w);
}
void ref_trainer_set_weights(void* h, const float* w, int64_t m) {
  static_cast<RefTrainer*>(h)->tr->set_weights(std::vector<float>(w, w + m));
}
void ref_trainer_reset_stats(void* h) { static_cast<RefTrainer*>(h)->tr->reset_stats(); }
void ref_trainer_stats(void* h, int64_t* syn_deliv, int64_t* neuron_fires) {
  auto* r = static_cast<RefTrainer*>(h);
  const auto& a = r->tr->syn_deliveries();
  const auto& b = r->tr->neuron_fires();
  std::copy(a.begin(), a.end(), syn_deliv);
  std::copy(b.begin(), b.end(), neuron_fires);
}
// One epoch over (X, y) with an empty test set; returns train accuracy.
double ref_trainer_train_epoch(void* h, const float* X, const int32_t* y, int32_t n_samples,
                               int32_t dim, int32_t classes, int32_t epoch, double* loss,
                               int64_t* spikes, int64_t* events) {
  auto* r = static_cast<RefTrainer*>(h);
  snc::Dataset tr, te;
  tr.dim = te.dim = dim;
  tr.classes = te.classes = classes;
  for (int i = 0; i < n_samples; ++i) {
    tr.x.emplace_back(X + (size_t)i * dim, X + (size_t)(i + 1) * dim);
    tr.y.push_back(y[i]);
  }
  snc::EpochStats st = r->tr->train_epoch(tr, te, epoch);
  if (loss) *loss = st.loss;
  if (spikes) *spikes = st.spikes;
  if (events) *events = st.synaptic_events;
  return st.train_acc;
}
double ref_trainer_evaluate(void* h, const float* X, const int32_t* y, int32_t n_samples,
                            int32_t dim, int32_t classes) {
  auto* r = static_cast<RefTrainer*>(h);
  snc::Dataset d;
  d.dim = dim;
  d.classes = classes;
  for (int i = 0; i < n_samples; ++i) {
    d.x.emplace_back(X + (size_t)i * dim, X + (size_t)(i + 1) * dim);
    d.y.push_back(y[i]);
  }
  return r->tr->evaluate(d);
}

// ---- connectome (the grow heuristic the learned proposal is compared with) -
void* ref_connectome_create(const int32_t* layers, int32_t n_layers, int32_t budget,
                            int32_t delay, float w_init, uint64_t seed) {
  std::vector<int> L(layers, layers + n_layers);
  return new snc::Connectome(snc::Connectome::layered_local(L, budget, delay, w_init, seed));
}
void ref_connectome_free(void* h) { delete static_cast<snc::Connectome*>(h); }
void* ref_connectome_to_graph(void* h) {
  return new RefGraph{static_cast<snc::Connectome*>(h)->to_graph()};
}
int32_t ref_connectome_num_synapses(void* h) {
  return static_cast<snc::Connectome*>(h)->num_synapses();
}
void ref_connectome_edge_weights(void* h, float* w) {
  auto v = static_cast<snc::Connectome*>(h)->edge_weights();
  std::copy(v.begin(), v.end(), w);
}
void ref_connectome_set_edge_weights(void* h, const float* w, int64_t m) {
  static_cast<snc::Connectome*>(h)->set_edge_weights(std::vector<float>(w, w + m));
}
void ref_connectome_update(void* h, const int64_t* syn_deliv, int64_t m,
                           const int64_t* neuron_fires, int64_t n, int32_t grow, int32_t delay,
                           float w_grow, float locality, int32_t protect, uint64_t seed,
                           int32_t* pruned, int32_t* grown) {
  snc::ActivityStats as;
  as.syn_deliveries.assign(syn_deliv, syn_deliv + m);
  as.neuron_fires.assign(neuron_fires, neuron_fires + n);
  snc::StructConfig sc;
  sc.grow_per_epoch = grow;
  sc.delay = delay;
  sc.w_grow_init = w_grow;
  sc.locality_window = locality;
  sc.protect_rounds = protect;
  sc.seed = seed;
  auto rep = static_cast<snc::Connectome*>(h)->structural_update(as, sc, 1.0f);
  *pruned = rep.pruned;
  *grown = rep.grown;
}

}  // extern "C"
