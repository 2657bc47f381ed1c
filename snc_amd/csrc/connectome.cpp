// Structural slow path (see include/snc_struct.h).
//
// The heuristic follows the reference's Connectome step for step, with the
// same libstdc++ draws and selection (paths relative to the reference tree):
//   layered_local       src/structure/connectome.cpp:15-64
//   to_graph            src/structure/connectome.cpp:78-102
//   structural_update   src/structure/connectome.cpp:121-219
// so its result equals the reference's for the same statistics. The learned
// mode and the grow log are this project's additions.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "common.hpp"
#include "snc_gen.h"
#include "snc_hip.h"
#include "snc_struct.h"

namespace {

constexpr int kAgeInit = 1 << 20;  // initial edges: prune-eligible at once
constexpr int kFeat = 16;

struct Edge {
  int post;
  int delay;
  float weight;
  int age;
  int64_t log_id;  // grow-log entry, -1 if none
};

float layer_pos(int k, int cnt) { return cnt > 1 ? (float)k / (cnt - 1) : 0.5f; }

}  // namespace

struct snc_connectome {
  int n = 0, num_layers = 0;
  std::vector<int> layer, layer_begin, layer_count;
  std::vector<float> pos;
  std::vector<uint8_t> role;
  std::vector<int> channel;
  std::vector<std::vector<Edge>> adj;
  std::mt19937_64 rng;
  std::mt19937_64 explore_rng;  // exploration draws of logging runs only
  std::vector<float> log_feat;
  std::vector<int32_t> log_label;

  bool connected(int pre, int post) const {
    for (const Edge& e : adj[pre])
      if (e.post == post) return true;
    return false;
  }
  int num_synapses() const {
    int m = 0;
    for (const auto& a : adj) m += (int)a.size();
    return m;
  }
  // Window of candidate pres for post p (connectome.cpp:171-176).
  bool window(int p, float w, int& b0, int& klo, int& khi) const {
    const int pre_l = layer[p] - 1;
    if (pre_l < 0) return false;
    b0 = layer_begin[pre_l];
    const int bn = layer_count[pre_l];
    klo = std::clamp((int)std::ceil((pos[p] - w) * (bn - 1)), 0, bn - 1);
    khi = std::clamp((int)std::floor((pos[p] + w) * (bn - 1)), 0, bn - 1);
    return true;
  }
};

namespace {

struct FeatureCtx {
  const snc_connectome* c;
  const int64_t* fires;
  int64_t n_fires;
  double norm;
  int delay;
  std::vector<int> indeg, outdeg;
  std::vector<double> in_w, out_w;  // mean |w| of incoming / outgoing edges
  std::vector<double> layer_mean_fire;
  double mean_in = 1.0, mean_out = 1.0;

  FeatureCtx(const snc_connectome* cc, const int64_t* f, int64_t nf, double nm, int d)
      : c(cc), fires(f), n_fires(nf), norm(nm > 0 ? nm : 1.0), delay(d) {
    const int n = c->n;
    indeg.assign(n, 0);
    outdeg.assign(n, 0);
    in_w.assign(n, 0.0);
    out_w.assign(n, 0.0);
    for (int i = 0; i < n; ++i)
      for (const Edge& e : c->adj[i]) {
        outdeg[i]++;
        indeg[e.post]++;
        out_w[i] += std::fabs(e.weight);
        in_w[e.post] += std::fabs(e.weight);
      }
    int64_t tot = 0;
    for (int i = 0; i < n; ++i) {
      if (outdeg[i]) out_w[i] /= outdeg[i];
      if (indeg[i]) in_w[i] /= indeg[i];
      tot += outdeg[i];
    }
    mean_in = mean_out = n ? std::max(1.0, (double)tot / n) : 1.0;
    layer_mean_fire.assign(c->num_layers, 0.0);
    for (int l = 0; l < c->num_layers; ++l) {
      double s = 0.0;
      for (int k = 0; k < c->layer_count[l]; ++k) s += fire(c->layer_begin[l] + k);
      layer_mean_fire[l] = c->layer_count[l] ? s / c->layer_count[l] : 0.0;
    }
  }
  double fire(int id) const { return id < n_fires ? (double)fires[id] : 0.0; }

  // Features of candidate q -> p; window statistics are of p's window.
  void features(int q, int p, double win_max, double win_rank, float* f) const {
    const double rq = fire(q) / norm, rp = fire(p) / norm;
    f[0] = (float)rq;
    f[1] = (float)rp;
    f[2] = (float)(rq * rp * 100.0);
    f[3] = std::fabs(c->pos[q] - c->pos[p]);
    f[4] = (float)delay / 32.0f;
    f[5] = (float)(indeg[p] / mean_in);
    f[6] = (float)(outdeg[q] / mean_out);
    f[7] = (float)in_w[p];
    f[8] = (float)out_w[q];
    f[9] = c->num_layers > 1 ? (float)c->layer[p] / (c->num_layers - 1) : 0.0f;
    f[10] = (float)std::log1p(fire(q));
    f[11] = (float)std::log1p(fire(p));
    f[12] = win_max > 0 ? (float)(fire(q) / win_max) : 0.0f;
    f[13] = (float)win_rank;
    const double lm = layer_mean_fire[c->layer[p]];
    f[14] = lm > 0 ? (float)(fire(p) / lm) : 0.0f;
    f[15] = 1.0f;
  }
};

}  // namespace

using snc_amd::fail;

extern "C" {

int snc_conn_layered_local(const int32_t* layers, int32_t n_layers, int32_t budget, int32_t delay,
                           float w_init, uint64_t seed, snc_connectome** out) {
  if (!layers || n_layers < 1 || !out) return fail(SNC_E_INVALID, "bad layers");
  std::unique_ptr<snc_connectome> c(new snc_connectome());
  c->num_layers = n_layers;
  for (int l = 0; l < n_layers; ++l) c->n += layers[l];
  c->layer.resize(c->n);
  c->pos.resize(c->n);
  c->role.assign(c->n, SNC_ROLE_INTERNAL);
  c->channel.assign(c->n, -1);
  c->adj.assign(c->n, {});
  c->layer_begin.resize(n_layers);
  c->layer_count.resize(n_layers);
  c->rng.seed(seed);
  c->explore_rng.seed(seed ^ 0xA5A5A5A5DEADBEEFULL);
  int idx = 0;
  for (int l = 0; l < n_layers; ++l) {
    c->layer_begin[l] = idx;
    c->layer_count[l] = layers[l];
    for (int k = 0; k < layers[l]; ++k, ++idx) {
      c->layer[idx] = l;
      c->pos[idx] = layer_pos(k, layers[l]);
      if (l == 0) {
        c->role[idx] = SNC_ROLE_INPUT;
        c->channel[idx] = k;
      } else if (l == n_layers - 1) {
        c->role[idx] = SNC_ROLE_OUTPUT;
        c->channel[idx] = k;
      }
    }
  }
  long long fwd = 0;
  for (int l = 0; l + 1 < n_layers; ++l) fwd += layers[l];
  const int fanout = fwd > 0 ? std::max(1, (int)(budget / fwd)) : 0;
  std::uniform_real_distribution<float> jitter(-0.5f, 0.5f);
  for (int l = 0; l + 1 < n_layers; ++l) {
    const int from0 = c->layer_begin[l], from_n = c->layer_count[l];
    const int to0 = c->layer_begin[l + 1], to_n = c->layer_count[l + 1];
    const int k = std::min(fanout, to_n);
    for (int a = 0; a < from_n; ++a) {
      const float pa = layer_pos(a, from_n) + jitter(c->rng) / std::max(1, to_n);
      int centre = (int)std::lround(pa * (to_n - 1));
      centre = std::clamp(centre, 0, std::max(0, to_n - 1));
      const int lo = std::clamp(centre - k / 2, 0, std::max(0, to_n - k));
      for (int j = 0; j < k; ++j) c->adj[from0 + a].push_back({to0 + lo + j, delay, w_init, kAgeInit, -1});
    }
  }
  *out = c.release();
  return SNC_OK;
}

void snc_conn_free(snc_connectome* c) { delete c; }

int32_t snc_conn_num_synapses(const snc_connectome* c) { return c ? c->num_synapses() : 0; }

int snc_conn_to_graph(const snc_connectome* c, snc_graph_buf** out) {
  if (!c || !out) return fail(SNC_E_INVALID, "null argument");
  auto* b = new snc_graph_buf();
  snc_amd::GraphArrays& g = b->g;
  g.n = c->n;
  g.role = c->role;
  g.channel.assign(c->channel.begin(), c->channel.end());
  g.sign.assign(c->n, 1);  // the trainer uses signed weights
  g.row_ptr.assign((size_t)c->n + 1, 0);
  for (int i = 0; i < c->n; ++i) g.row_ptr[i + 1] = g.row_ptr[i] + (int)c->adj[i].size();
  const int m = g.row_ptr[c->n];
  g.post.resize(m);
  g.weight.resize(m);
  g.delay.resize(m);
  int s = 0;
  for (int i = 0; i < c->n; ++i)
    for (const Edge& e : c->adj[i]) {
      g.post[s] = e.post;
      g.weight[s] = std::fabs(e.weight);
      g.delay[s] = e.delay;
      ++s;
    }
  *out = b;
  return SNC_OK;
}

int snc_conn_edge_weights(const snc_connectome* c, float* w, int64_t n) {
  if (!c || !w || n != c->num_synapses()) return fail(SNC_E_INVALID, "weight count mismatch");
  int64_t k = 0;
  for (const auto& a : c->adj)
    for (const Edge& e : a) w[k++] = e.weight;
  return SNC_OK;
}

int snc_conn_set_edge_weights(snc_connectome* c, const float* w, int64_t n) {
  if (!c || !w) return fail(SNC_E_INVALID, "null argument");
  int64_t k = 0;
  for (auto& a : c->adj)
    for (Edge& e : a) {
      if (k >= n) return SNC_OK;  // as the reference: stop when the input runs out
      e.weight = w[k++];
    }
  return SNC_OK;
}

int snc_conn_structural_update(snc_connectome* c, const int64_t* syn_deliveries, int64_t n_syn,
                               const int64_t* neuron_fires, int64_t n_neur,
                               const snc_struct_config* cfg, float reward, int32_t grow_mode,
                               snc_hip_mlp* model, double activity_norm, int32_t log_grown,
                               snc_struct_report* rep) {
  (void)syn_deliveries;
  (void)n_syn;
  if (!c || !cfg || !rep) return fail(SNC_E_INVALID, "null argument");
  std::memset(rep, 0, sizeof(*rep));
  int grow_target = cfg->grow_per_epoch;
  if (cfg->reward_modulated) {
    const float denom = std::max(1e-3f, 1.0f - cfg->reward_chance);
    const float comp = std::clamp((reward - cfg->reward_chance) / denom, 0.0f, 1.0f);
    const float frac = std::clamp(1.0f - comp, cfg->reward_floor, 1.0f);
    grow_target = std::max(0, (int)std::lround(cfg->grow_per_epoch * frac));
  }

  // --- prune the K weakest |w| among prune-eligible edges ---
  struct Cand {
    float score;
    int i, k;
  };
  std::vector<Cand> cand;
  for (int i = 0; i < c->n; ++i)
    for (int k = 0; k < (int)c->adj[i].size(); ++k)
      if (c->adj[i][k].age >= cfg->protect_rounds) cand.push_back({std::fabs(c->adj[i][k].weight), i, k});
  const int K = std::min(grow_target, (int)cand.size());
  if (K > 0)
    std::nth_element(cand.begin(), cand.begin() + K, cand.end(),
                     [](const Cand& a, const Cand& b) { return a.score < b.score; });
  std::vector<std::vector<char>> rm(c->n);
  for (int i = 0; i < c->n; ++i) rm[i].assign(c->adj[i].size(), 0);
  for (int j = 0; j < K; ++j) rm[cand[j].i][cand[j].k] = 1;
  // Grow-log labels: an edge decides its outcome at its first eligible prune.
  for (const Cand& x : cand) {
    const Edge& e = c->adj[x.i][x.k];
    if (e.log_id >= 0 && c->log_label[e.log_id] < 0)
      c->log_label[e.log_id] = rm[x.i][x.k] ? 0 : 1;
  }
  int pruned = 0;
  std::vector<int> freed;
  freed.reserve(K);
  for (int i = 0; i < c->n; ++i) {
    std::vector<Edge> keep;
    keep.reserve(c->adj[i].size());
    for (int k = 0; k < (int)c->adj[i].size(); ++k) {
      if (rm[i][k]) {
        ++pruned;
        freed.push_back(c->adj[i][k].post);
      } else {
        keep.push_back(c->adj[i][k]);
      }
    }
    c->adj[i].swap(keep);
  }

  // --- grow ---
  auto fire_of = [&](int id) -> long long { return id < n_neur ? neuron_fires[id] : 0LL; };
  auto grow_local = [&](int p) -> int {
    int b0, klo, khi;
    if (!c->window(p, cfg->locality_window, b0, klo, khi)) return -1;
    int best = -1;
    long long best_fire = -1;
    for (int kk = klo; kk <= khi; ++kk) {
      const int q = b0 + kk;
      if (c->connected(q, p)) continue;
      if (fire_of(q) > best_fire) {
        best_fire = fire_of(q);
        best = q;
      }
    }
    return best;
  };
  FeatureCtx fx(c, neuron_fires, n_neur, activity_norm, cfg->delay);
  auto window_features = [&](int p, std::vector<int>& qs, std::vector<float>& feats) {
    qs.clear();
    feats.clear();
    int b0, klo, khi;
    if (!c->window(p, cfg->locality_window, b0, klo, khi)) return;
    double mx = 0.0;
    std::vector<std::pair<double, int>> order;
    for (int kk = klo; kk <= khi; ++kk) {
      const int q = b0 + kk;
      mx = std::max(mx, fx.fire(q));
      order.push_back({fx.fire(q), q});
    }
    std::sort(order.begin(), order.end(), [](const auto& a, const auto& b) {
      return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    std::vector<double> rank(c->n, 0.0);
    for (size_t r = 0; r < order.size(); ++r)
      rank[order[r].second] = order.size() > 1 ? (double)r / (order.size() - 1) : 0.0;
    for (int kk = klo; kk <= khi; ++kk) {
      const int q = b0 + kk;
      qs.push_back(q);
      feats.resize(feats.size() + kFeat);
      fx.features(q, p, mx, rank[q], feats.data() + feats.size() - kFeat);
    }
  };
  auto log_edge = [&](int q, int p, const float* f) -> int64_t {
    if (!log_grown) return -1;
    const int64_t id = (int64_t)c->log_label.size();
    c->log_feat.insert(c->log_feat.end(), f, f + kFeat);
    c->log_label.push_back(-1);
    (void)q;
    (void)p;
    return id;
  };

  int grown = 0;
  const bool learned = grow_mode == SNC_GROW_LEARNED && model != nullptr;
  if (learned) {
    // Score every window candidate of every freed post in one call, then
    // regrow each freed post from its best-scoring still-unconnected pre.
    std::vector<int> offs(freed.size() + 1, 0), all_q;
    std::vector<float> all_f, qs_feat;
    std::vector<int> qs;
    for (size_t u = 0; u < freed.size(); ++u) {
      window_features(freed[u], qs, qs_feat);
      all_q.insert(all_q.end(), qs.begin(), qs.end());
      all_f.insert(all_f.end(), qs_feat.begin(), qs_feat.end());
      offs[u + 1] = (int)all_q.size();
    }
    std::vector<float> score(all_q.size());
    double ms = 0.0;
    if (!all_q.empty()) {
      const int rc = snc_hip_mlp_score(model, all_f.data(), (int64_t)all_q.size(), score.data(), &ms);
      if (rc) return rc;
    }
    rep->candidates_scored = (int64_t)all_q.size();
    rep->score_ms = ms;
    for (size_t u = 0; u < freed.size() && grown < pruned; ++u) {
      const int p = freed[u];
      int best = -1;
      float bs = -INFINITY;
      for (int k = offs[u]; k < offs[u + 1]; ++k) {
        const int q = all_q[k];
        if (c->connected(q, p)) continue;
        if (score[k] > bs) {
          bs = score[k];
          best = k;
        }
      }
      if (best >= 0) {
        const int q = all_q[best];
        c->adj[q].push_back({p, cfg->delay, cfg->w_grow_init, 0, log_edge(q, p, &all_f[(size_t)best * kFeat])});
        ++grown;
      }
    }
  } else {
    std::vector<int> qs;
    std::vector<float> qf;
    const bool explore = log_grown && cfg->explore > 0.0f;
    std::uniform_real_distribution<float> u01(0.0f, 1.0f);
    for (int p : freed) {
      if (grown >= pruned) break;
      int q = grow_local(p);
      if (q >= 0 && explore && u01(c->explore_rng) < cfg->explore) {
        int b0, klo, khi;
        std::vector<int> open;
        if (c->window(p, cfg->locality_window, b0, klo, khi))
          for (int kk = klo; kk <= khi; ++kk)
            if (!c->connected(b0 + kk, p)) open.push_back(b0 + kk);
        if (!open.empty())
          q = open[std::uniform_int_distribution<size_t>(0, open.size() - 1)(c->explore_rng)];
      }
      if (q >= 0) {
        int64_t lid = -1;
        if (log_grown) {
          window_features(p, qs, qf);
          for (size_t k = 0; k < qs.size(); ++k)
            if (qs[k] == q) lid = log_edge(q, p, &qf[k * kFeat]);
        }
        c->adj[q].push_back({p, cfg->delay, cfg->w_grow_init, 0, lid});
        ++grown;
      }
    }
  }
  // Fallback onto under-driven posts (4-way tournament), heuristic choice.
  std::uniform_int_distribution<int> any_post(c->layer_begin.size() > 1 ? c->layer_begin[1] : 0,
                                              std::max(0, c->n - 1));
  int attempts = 0;
  const int max_attempts = cfg->max_attempts_mult * std::max(1, pruned);
  while (grown < pruned && attempts < max_attempts) {
    ++attempts;
    int p = -1;
    for (int t = 0; t < 4; ++t) {
      const int q = any_post(c->rng);
      if (c->layer[q] == 0) continue;
      if (p < 0 || fire_of(q) < fire_of(p)) p = q;
    }
    if (p < 0) continue;
    const int q = grow_local(p);
    if (q >= 0) {
      c->adj[q].push_back({p, cfg->delay, cfg->w_grow_init, 0, -1});
      ++grown;
    }
  }
  for (auto& a : c->adj)
    for (Edge& e : a)
      if (e.age < kAgeInit) ++e.age;
  rep->pruned = pruned;
  rep->grown = grown;
  rep->synapses = c->num_synapses();
  return SNC_OK;
}

int64_t snc_conn_grow_log_size(const snc_connectome* c) {
  return c ? (int64_t)c->log_label.size() : 0;
}

int snc_conn_grow_log(const snc_connectome* c, float* features, int32_t* labels, int64_t n) {
  if (!c || n != (int64_t)c->log_label.size()) return fail(SNC_E_INVALID, "log size mismatch");
  if (features) std::copy(c->log_feat.begin(), c->log_feat.end(), features);
  if (labels) std::copy(c->log_label.begin(), c->log_label.end(), labels);
  return SNC_OK;
}

int snc_conn_candidate_features(const snc_connectome* c, const int32_t* pre, const int32_t* post,
                                int64_t n, const int64_t* neuron_fires, int64_t n_neur,
                                double activity_norm, int32_t delay, float* features) {
  if (!c || (n > 0 && (!pre || !post || !features))) return fail(SNC_E_INVALID, "bad arguments");
  FeatureCtx fx(c, neuron_fires, n_neur, activity_norm, delay);
  for (int64_t k = 0; k < n; ++k) {
    if (pre[k] < 0 || pre[k] >= c->n || post[k] < 0 || post[k] >= c->n)
      return fail(SNC_E_INVALID, "neuron id out of range");
    fx.features(pre[k], post[k], 0.0, 0.0, features + k * kFeat);
  }
  return SNC_OK;
}

}  // extern "C"
