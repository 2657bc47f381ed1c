// Structural slow path (see include/snc_struct.h).
//
// The heuristic follows the reference's Connectome step for step, with the
// same libstdc++ draws and selection (paths relative to the reference tree):
//   layered_local       src/structure/connectome.cpp:15-64
//   to_graph            src/structure/connectome.cpp:78-102
//   structural_update   src/structure/connectome.cpp:121-219
// so its result equals the reference's for the same statistics. The learned
// mode and the grow log are this project's additions, and so is the sheet
// connectome (snc_conn_sheet_recurrent): layers on 2-D grids, recurrent
// hidden layers, disc-shaped grow windows over the previous and (hidden) the
// own layer, and a total prune order -- (|w|, synapse position in CSR order)
// ascending -- that the device path (conn_gpu.hip) reproduces exactly.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "common.hpp"
#include "conn_gpu.hpp"
#include "rng.hpp"
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
  // Sheet connectome: every layer a side x side grid over the unit square.
  bool sheet = false;
  std::vector<int> side;       // per layer
  std::vector<float> px, py;   // per neuron
  std::vector<uint8_t> role;
  std::vector<int> channel;
  std::vector<std::vector<Edge>> adj;
  std::mt19937_64 rng;
  std::mt19937_64 explore_rng;  // exploration draws of logging runs only
  std::vector<float> log_feat;
  std::vector<int32_t> log_label;
  std::vector<float> log_rank;  // |w| percentile among the eligible edges at that prune

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
  // 2-D distance of two sheet neurons, in double (the device computes the same).
  double dist2d(int q, int p) const {
    const double dx = (double)px[q] - (double)px[p], dy = (double)py[q] - (double)py[p];
    return std::sqrt(dx * dx + dy * dy);
  }
  // Candidate pres of post p in order: the layered window range, or on a
  // sheet every neuron of layer l-1 and (hidden p) of layer l, p excluded,
  // within distance w, by layer then id.
  void window_list(int p, float w, std::vector<int>& out) const {
    out.clear();
    if (!sheet) {
      int b0, klo, khi;
      if (window(p, w, b0, klo, khi))
        for (int k = klo; k <= khi; ++k) out.push_back(b0 + k);
      return;
    }
    const int l = layer[p];
    if (l < 1) return;
    const bool hidden = l < num_layers - 1;
    for (int ll = l - 1; ll <= (hidden ? l : l - 1); ++ll)
      for (int k = 0; k < layer_count[ll]; ++k) {
        const int q = layer_begin[ll] + k;
        if (q != p && dist2d(q, p) <= (double)w) out.push_back(q);
      }
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
    f[3] = c->sheet ? (float)c->dist2d(q, p) : std::fabs(c->pos[q] - c->pos[p]);
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

// The device path's inputs: the sheet layout and the per-neuron columns of
// the features, computed here as FeatureCtx computes them.
snc_amd::ConnSheetView sheet_view(const snc_connectome* c, const FeatureCtx& fx) {
  snc_amd::ConnSheetView v;
  const int n = c->n;
  v.n = n;
  v.num_layers = c->num_layers;
  v.layer.assign(c->layer.begin(), c->layer.end());
  v.layer_begin.assign(c->layer_begin.begin(), c->layer_begin.end());
  v.layer_count.assign(c->layer_count.begin(), c->layer_count.end());
  v.side.assign(c->side.begin(), c->side.end());
  v.px = c->px;
  v.py = c->py;
  v.fire.resize(n);
  v.rate.resize(n);
  v.f5.resize(n);
  v.f6.resize(n);
  v.f7.resize(n);
  v.f8.resize(n);
  v.f9.resize(n);
  v.f10.resize(n);
  v.f14.resize(n);
  for (int i = 0; i < n; ++i) {
    v.fire[i] = fx.fire(i);
    v.rate[i] = fx.fire(i) / fx.norm;
    v.f5[i] = (float)(fx.indeg[i] / fx.mean_in);
    v.f6[i] = (float)(fx.outdeg[i] / fx.mean_out);
    v.f7[i] = (float)fx.in_w[i];
    v.f8[i] = (float)fx.out_w[i];
    v.f9[i] = c->num_layers > 1 ? (float)c->layer[i] / (c->num_layers - 1) : 0.0f;
    v.f10[i] = (float)std::log1p(fx.fire(i));
    const double lm = fx.layer_mean_fire[c->layer[i]];
    v.f14[i] = lm > 0 ? (float)(fx.fire(i) / lm) : 0.0f;
  }
  v.f4 = (float)fx.delay / 32.0f;
  return v;
}

}  // namespace

using snc_amd::conn_gpu_candidates;
using snc_amd::conn_gpu_prune_select;
using snc_amd::ConnSheetView;
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

int snc_conn_sheet_recurrent(const int32_t* layers, int32_t n_layers, int32_t fan_in,
                             int32_t rec_fan_in, float radius, int32_t delay, float w_init,
                             uint64_t seed, snc_connectome** out) {
  if (!layers || n_layers < 2 || !out || fan_in < 0 || rec_fan_in < 0 || !(radius > 0.0f) ||
      delay < 1)
    return fail(SNC_E_INVALID, "bad sheet connectome arguments");
  for (int l = 0; l < n_layers; ++l)
    if (layers[l] < 1) return fail(SNC_E_INVALID, "empty layer");
  std::unique_ptr<snc_connectome> c(new snc_connectome());
  c->sheet = true;
  c->num_layers = n_layers;
  for (int l = 0; l < n_layers; ++l) c->n += layers[l];
  const int n = c->n;
  c->layer.resize(n);
  c->pos.resize(n);
  c->px.resize(n);
  c->py.resize(n);
  c->role.assign(n, SNC_ROLE_INTERNAL);
  c->channel.assign(n, -1);
  c->adj.assign(n, {});
  c->layer_begin.resize(n_layers);
  c->layer_count.resize(n_layers);
  c->side.resize(n_layers);
  c->rng.seed(seed);
  c->explore_rng.seed(seed ^ 0xA5A5A5A5DEADBEEFULL);
  int idx = 0;
  for (int l = 0; l < n_layers; ++l) {
    const int cnt = layers[l];
    const int side = (int)std::ceil(std::sqrt((double)cnt));
    c->layer_begin[l] = idx;
    c->layer_count[l] = cnt;
    c->side[l] = side;
    for (int k = 0; k < cnt; ++k, ++idx) {
      c->layer[idx] = l;
      c->pos[idx] = layer_pos(k, cnt);
      c->px[idx] = side > 1 ? (float)(k % side) / (float)(side - 1) : 0.5f;
      c->py[idx] = side > 1 ? (float)(k / side) / (float)(side - 1) : 0.5f;
      if (l == 0) {
        c->role[idx] = SNC_ROLE_INPUT;
        c->channel[idx] = k;
      } else if (l == n_layers - 1) {
        c->role[idx] = SNC_ROLE_OUTPUT;
        c->channel[idx] = k;
      }
    }
  }
  // Partners of each post: the k smallest exponential race times
  // -log(u) * exp(d / radius) among the candidates (a draw without
  // replacement with probability proportional to exp(-d / radius)).
  snc_amd::SplitMix64 rng(seed ^ 0x5EE7C0DEULL);
  std::vector<std::pair<double, int>> race;
  std::vector<std::vector<int>> pres(n);
  for (int p = c->layer_begin[1]; p < n; ++p) {
    const int l = c->layer[p];
    for (int pass = 0; pass < 2; ++pass) {
      const int ll = pass == 0 ? l - 1 : l;
      const int k = pass == 0 ? fan_in : (l < n_layers - 1 ? rec_fan_in : 0);
      if (k <= 0) continue;
      race.clear();
      for (int kk = 0; kk < c->layer_count[ll]; ++kk) {
        const int q = c->layer_begin[ll] + kk;
        if (q == p) continue;
        race.push_back({-std::log(rng.uniform_pos()) * std::exp(c->dist2d(q, p) / radius), q});
      }
      const int take = std::min(k, (int)race.size());
      std::partial_sort(race.begin(), race.begin() + take, race.end());
      for (int t = 0; t < take; ++t) pres[p].push_back(race[t].second);
    }
    std::sort(pres[p].begin(), pres[p].end());
    for (int q : pres[p]) c->adj[q].push_back({p, delay, w_init, kAgeInit, -1});
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
  std::vector<std::vector<char>> rm(c->n);
  for (int i = 0; i < c->n; ++i) rm[i].assign(c->adj[i].size(), 0);
  if (c->sheet && cfg->on_device) {
    // The K smallest (|w|, CSR position) on the device.
    std::vector<float> w;
    std::vector<int32_t> age;
    for (const auto& a : c->adj)
      for (const Edge& e : a) {
        w.push_back(e.weight);
        age.push_back(e.age);
      }
    std::vector<uint8_t> mask(w.size(), 0);
    const int rc = conn_gpu_prune_select(w.data(), age.data(), (int64_t)w.size(),
                                         cfg->protect_rounds, K, mask.data());
    if (rc) return rc;
    int64_t q = 0;
    for (int i = 0; i < c->n; ++i)
      for (size_t k = 0; k < c->adj[i].size(); ++k) rm[i][k] = (char)mask[q++];
  } else if (c->sheet) {
    // Total order (|w|, CSR position): cand is built in CSR order.
    if (K > 0)
      std::nth_element(cand.begin(), cand.begin() + K, cand.end(), [&](const Cand& a, const Cand& b) {
        return a.score != b.score ? a.score < b.score : (a.i != b.i ? a.i < b.i : a.k < b.k);
      });
    for (int j = 0; j < K; ++j) rm[cand[j].i][cand[j].k] = 1;
  } else {
    if (K > 0)
      std::nth_element(cand.begin(), cand.begin() + K, cand.end(),
                       [](const Cand& a, const Cand& b) { return a.score < b.score; });
    for (int j = 0; j < K; ++j) rm[cand[j].i][cand[j].k] = 1;
  }
  // Grow-log labels: an edge decides its outcome at its first eligible prune:
  // whether it survived, and where its |w| ranks among the eligible edges
  // (fraction below it, ties counted half).
  std::vector<float> sorted_scores;
  for (const Cand& x : cand) {
    const Edge& e = c->adj[x.i][x.k];
    if (e.log_id < 0 || c->log_label[e.log_id] >= 0) continue;
    c->log_label[e.log_id] = rm[x.i][x.k] ? 0 : 1;
    if (sorted_scores.empty()) {
      sorted_scores.reserve(cand.size());
      for (const Cand& y : cand) sorted_scores.push_back(y.score);
      std::sort(sorted_scores.begin(), sorted_scores.end());
    }
    const auto lo = std::lower_bound(sorted_scores.begin(), sorted_scores.end(), x.score);
    const auto hi = std::upper_bound(lo, sorted_scores.end(), x.score);
    const double below = (double)(lo - sorted_scores.begin()) + 0.5 * (double)(hi - lo);
    c->log_rank[e.log_id] = (float)(below / (double)sorted_scores.size());
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
  std::vector<int> wl;
  auto grow_local = [&](int p) -> int {
    c->window_list(p, cfg->locality_window, wl);
    int best = -1;
    long long best_fire = -1;
    for (int q : wl) {
      if (c->connected(q, p)) continue;
      if (fire_of(q) > best_fire) {
        best_fire = fire_of(q);
        best = q;
      }
    }
    return best;
  };
  FeatureCtx fx(c, neuron_fires, n_neur, activity_norm, cfg->delay);
  std::vector<int> wf;
  auto window_features = [&](int p, std::vector<int>& qs, std::vector<float>& feats) {
    qs.clear();
    feats.clear();
    c->window_list(p, cfg->locality_window, wf);
    if (wf.empty()) return;
    double mx = 0.0;
    std::vector<std::pair<double, int>> order;
    for (int q : wf) {
      mx = std::max(mx, fx.fire(q));
      order.push_back({fx.fire(q), q});
    }
    std::sort(order.begin(), order.end(), [](const auto& a, const auto& b) {
      return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    std::vector<double> rank(c->n, 0.0);
    for (size_t r = 0; r < order.size(); ++r)
      rank[order[r].second] = order.size() > 1 ? (double)r / (order.size() - 1) : 0.0;
    for (int q : wf) {
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
    c->log_rank.push_back(-1.0f);
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
    std::vector<float> score;
    double ms = 0.0;
    if (c->sheet && cfg->on_device) {
      // Candidates, features and scores on the device; back come the
      // candidates' pres, scores and features (the features for the log).
      const ConnSheetView sv = sheet_view(c, fx);
      const int rc = conn_gpu_candidates(sv, freed, cfg->locality_window, model, offs, all_q,
                                         all_f, score, &ms);
      if (rc) return rc;
    } else {
      for (size_t u = 0; u < freed.size(); ++u) {
        window_features(freed[u], qs, qs_feat);
        all_q.insert(all_q.end(), qs.begin(), qs.end());
        all_f.insert(all_f.end(), qs_feat.begin(), qs_feat.end());
        offs[u + 1] = (int)all_q.size();
      }
      score.assign(all_q.size(), 0.0f);
      if (!all_q.empty()) {
        const int rc =
            snc_hip_mlp_score(model, all_f.data(), (int64_t)all_q.size(), score.data(), &ms);
        if (rc) return rc;
      }
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
        std::vector<int> open;
        c->window_list(p, cfg->locality_window, wl);
        for (int cq : wl)
          if (!c->connected(cq, p)) open.push_back(cq);
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

int snc_conn_grow_log_rank(const snc_connectome* c, float* rank, int64_t n) {
  if (!c || !rank || n != (int64_t)c->log_rank.size()) return fail(SNC_E_INVALID, "log size mismatch");
  std::copy(c->log_rank.begin(), c->log_rank.end(), rank);
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
