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
#include <numeric>
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

// Position of the k-th of cnt neurons of a layer on [0, 1] (0.5 for one).
float unit_coord(int k, int cnt) { return cnt > 1 ? (float)k / (cnt - 1) : 0.5f; }

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
  // Candidate pres of post p on a layered connectome: the neurons of layer
  // l-1 whose positions lie within w of p's, as an index range [first,
  // last] of that layer (connectome.cpp:171-176).
  bool layered_window(int p, float w, int& layer0, int& first, int& last) const {
    const int src = layer[p] - 1;
    if (src < 0) return false;
    layer0 = layer_begin[src];
    const int top = layer_count[src] - 1;
    first = std::clamp((int)std::ceil((pos[p] - w) * top), 0, top);
    last = std::clamp((int)std::floor((pos[p] + w) * top), 0, top);
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
      int layer0, first, last;
      if (layered_window(p, w, layer0, first, last))
        for (int k = first; k <= last; ++k) out.push_back(layer0 + k);
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

namespace {

// Synapses regrown per step: grow_per_epoch, or with reward modulation a
// share of it that falls as the reward rises above chance, never below
// reward_floor (connectome.cpp:124-130).
int regrowth_quota(const snc_struct_config& cfg, float reward) {
  if (!cfg.reward_modulated) return cfg.grow_per_epoch;
  const float above_chance =
      std::clamp((reward - cfg.reward_chance) / std::max(1e-3f, 1.0f - cfg.reward_chance), 0.0f,
                 1.0f);
  const float share = std::clamp(1.0f - above_chance, cfg.reward_floor, 1.0f);
  return std::max(0, (int)std::lround(cfg.grow_per_epoch * share));
}

// One structural step on a connectome: prune, regrow (heuristic or learned),
// fill up by tournament, age. Edges are addressed by their CSR position,
// edge_base[pre] + slot, fixed between the prune's selection and its removal.
class StructuralStep {
 public:
  StructuralStep(snc_connectome* c, const snc_struct_config& cfg, const int64_t* fires,
                 int64_t n_fires, double activity_norm, bool log)
      : c_(c), cfg_(cfg), fires_(fires), n_fires_(n_fires), norm_(activity_norm), log_(log) {}

  int pruned() const { return (int)freed_.size(); }
  int grown() const { return grown_; }

  // Removes the `quota` prune-eligible synapses (age >= protect_rounds) of
  // smallest |w| (connectome.cpp:132-160) and records each one's post, in
  // CSR order, as a post to regrow. The layered connectome selects with the
  // reference's partial ordering by |w| alone (ties as its nth_element
  // leaves them); the sheet connectome with the total order (|w|, CSR
  // position), on the host or on the device.
  int prune(int quota) {
    std::vector<int64_t> base((size_t)c_->n + 1, 0);
    for (int i = 0; i < c_->n; ++i) base[i + 1] = base[i] + (int64_t)c_->adj[i].size();
    struct Eligible {
      float mag;
      int64_t at;  // CSR position
    };
    std::vector<Eligible> pool;
    for (int i = 0; i < c_->n; ++i)
      for (size_t k = 0; k < c_->adj[i].size(); ++k)
        if (c_->adj[i][k].age >= cfg_.protect_rounds)
          pool.push_back({std::fabs(c_->adj[i][k].weight), base[i] + (int64_t)k});
    const int take = std::min(quota, (int)pool.size());
    std::vector<uint8_t> drop((size_t)base[c_->n], 0);
    if (c_->sheet && cfg_.on_device) {
      std::vector<float> w;
      std::vector<int32_t> age;
      w.reserve(drop.size());
      age.reserve(drop.size());
      for (const auto& row : c_->adj)
        for (const Edge& e : row) {
          w.push_back(e.weight);
          age.push_back(e.age);
        }
      const int rc = conn_gpu_prune_select(w.data(), age.data(), (int64_t)w.size(),
                                           cfg_.protect_rounds, take, drop.data());
      if (rc) return rc;
    } else {
      if (take > 0) {
        if (c_->sheet)
          std::nth_element(pool.begin(), pool.begin() + take, pool.end(),
                           [](const Eligible& x, const Eligible& y) {
                             return x.mag != y.mag ? x.mag < y.mag : x.at < y.at;
                           });
        else
          std::nth_element(pool.begin(), pool.begin() + take, pool.end(),
                           [](const Eligible& x, const Eligible& y) { return x.mag < y.mag; });
      }
      for (int j = 0; j < take; ++j) drop[(size_t)pool[j].at] = 1;
    }
    // Logged synapses meeting their first eligible prune get their outcome:
    // survived or not, and the share of eligible |w| below theirs (ties
    // counted half).
    std::vector<float> mags;
    for (const Eligible& x : pool) {
      const int pre = (int)(std::upper_bound(base.begin(), base.end(), x.at) - base.begin()) - 1;
      const Edge& e = c_->adj[pre][(size_t)(x.at - base[pre])];
      if (e.log_id < 0 || c_->log_label[e.log_id] >= 0) continue;
      c_->log_label[e.log_id] = drop[(size_t)x.at] ? 0 : 1;
      if (mags.empty()) {
        for (const Eligible& y : pool) mags.push_back(y.mag);
        std::sort(mags.begin(), mags.end());
      }
      const auto lo = std::lower_bound(mags.begin(), mags.end(), x.mag);
      const auto hi = std::upper_bound(lo, mags.end(), x.mag);
      const double below = (double)(lo - mags.begin()) + 0.5 * (double)(hi - lo);
      c_->log_rank[e.log_id] = (float)(below / (double)mags.size());
    }
    for (int i = 0; i < c_->n; ++i) {
      std::vector<Edge>& row = c_->adj[i];
      size_t kept = 0;
      for (size_t k = 0; k < row.size(); ++k) {
        if (drop[(size_t)(base[i] + (int64_t)k)])
          freed_.push_back(row[k].post);
        else
          row[kept++] = row[k];
      }
      row.resize(kept);
    }
    // the features see the connectome as the prune left it
    fx_.reset(new FeatureCtx(c_, fires_, n_fires_, norm_, cfg_.delay));
    // a negative w_grow_init starts grown synapses at the surviving synapses'
    // median |w| (the lower middle one of an even count)
    w_start_ = cfg_.w_grow_init;
    if (w_start_ < 0.0f) {
      std::vector<float> m;
      for (const auto& row : c_->adj)
        for (const Edge& e : row) m.push_back(std::fabs(e.weight));
      w_start_ = 0.0f;
      if (!m.empty()) {
        const size_t mid = (m.size() - 1) / 2;
        std::nth_element(m.begin(), m.begin() + (ptrdiff_t)mid, m.end());
        w_start_ = m[mid];
      }
    }
    return SNC_OK;
  }

  // The reference's regrowth (connectome.cpp:162-190): each freed post, in
  // prune order, regrows from the unconnected pre of its window that fired
  // most (the first such in window order). Logging runs may swap the choice
  // for a uniformly drawn open window member (cfg.explore).
  void grow_heuristic() {
    const bool explore = log_ && cfg_.explore > 0.0f;
    std::uniform_real_distribution<float> coin(0.0f, 1.0f);
    for (int p : freed_) {
      if (grown_ >= pruned()) break;
      int q = strongest_open_pre(p);
      if (q >= 0 && explore && coin(c_->explore_rng) < cfg_.explore) {
        std::vector<int> open;
        c_->window_list(p, cfg_.locality_window, scan_);
        for (int cand : scan_)
          if (!c_->connected(cand, p)) open.push_back(cand);
        if (!open.empty())
          q = open[std::uniform_int_distribution<size_t>(0, open.size() - 1)(c_->explore_rng)];
      }
      if (q < 0) continue;
      int64_t id = -1;
      if (log_) {
        window_features(p, wq_, wf_);
        for (size_t k = 0; k < wq_.size(); ++k)
          if (wq_[k] == q) id = log_features(&wf_[k * kFeat]);
      }
      add_edge(q, p, id);
    }
  }

  // Learned regrowth. Per post: each freed post regrows from the open
  // window candidate the MLP scores highest (the first such in window
  // order). Pool: every candidate of every non-input post's window is
  // scored, and the `pruned` highest-scoring open pairs grow (ties: post
  // order, then window order), wherever their posts are.
  int grow_learned(snc_hip_mlp* model, bool pool, snc_struct_report* rep) {
    std::vector<int> posts;
    if (pool) {
      for (int p = 0; p < c_->n; ++p)
        if (c_->layer[p] >= 1) posts.push_back(p);
    } else {
      posts = freed_;
    }
    std::vector<int> offs(posts.size() + 1, 0), cand;
    std::vector<float> feats, score, rank;
    std::vector<double> wmax;
    double ms = 0.0;
    const bool device = c_->sheet && cfg_.on_device;
    if (device) {
      // Candidates, features and scores on the device; back come the pres,
      // scores and the two window statistics, from which the log's float
      // features of the grown candidates are rebuilt here.
      const ConnSheetView sv = sheet_view(c_, *fx_);
      const int rc = conn_gpu_candidates(sv, posts, cfg_.locality_window, model, offs, cand, rank,
                                         wmax, score, &ms);
      if (rc) return rc;
    } else {
      for (size_t u = 0; u < posts.size(); ++u) {
        window_features(posts[u], wq_, wf_);
        cand.insert(cand.end(), wq_.begin(), wq_.end());
        feats.insert(feats.end(), wf_.begin(), wf_.end());
        offs[u + 1] = (int)cand.size();
      }
      score.assign(cand.size(), 0.0f);
      if (!cand.empty()) {
        const int rc = snc_hip_mlp_score(model, feats.data(), (int64_t)cand.size(), score.data(), &ms);
        if (rc) return rc;
      }
    }
    rep->candidates_scored = (int64_t)cand.size();
    rep->score_ms = ms;
    float f[kFeat];
    auto grow_candidate = [&](size_t u, int k) {
      const int q = cand[k], p = posts[u];
      int64_t id = -1;
      if (log_) {
        if (device)
          fx_->features(q, p, wmax[u], (double)rank[k], f);
        else
          std::memcpy(f, &feats[(size_t)k * kFeat], sizeof(f));
        id = log_features(f);
      }
      add_edge(q, p, id);
    };
    if (!pool) {
      for (size_t u = 0; u < posts.size() && grown_ < pruned(); ++u) {
        int pick = -1;
        float top = -INFINITY;
        for (int k = offs[u]; k < offs[u + 1]; ++k)
          if (!c_->connected(cand[k], posts[u]) && score[k] > top) {
            top = score[k];
            pick = k;
          }
        if (pick >= 0) grow_candidate(u, pick);
      }
      return SNC_OK;
    }
    std::vector<int> order(cand.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) {
      return score[x] != score[y] ? score[x] > score[y] : x < y;
    });
    std::vector<size_t> post_of(cand.size());
    for (size_t u = 0; u < posts.size(); ++u)
      for (int k = offs[u]; k < offs[u + 1]; ++k) post_of[k] = u;
    for (int k : order) {
      if (grown_ >= pruned()) break;
      if (!c_->connected(cand[k], posts[post_of[k]])) grow_candidate(post_of[k], k);
    }
    return SNC_OK;
  }

  // Until as many synapses grew as were pruned (or the attempts run out):
  // draw four non-input posts uniformly, take the one that fired least (the
  // first such), regrow it from its strongest open pre (connectome.cpp:192-211).
  void grow_tournament() {
    const int lo = c_->layer_begin.size() > 1 ? c_->layer_begin[1] : 0;
    std::uniform_int_distribution<int> draw(lo, std::max(0, c_->n - 1));
    const int budget = cfg_.max_attempts_mult * std::max(1, pruned());
    for (int tries = 0; grown_ < pruned() && tries < budget; ++tries) {
      int p = -1;
      for (int t = 0; t < 4; ++t) {
        const int x = draw(c_->rng);
        if (c_->layer[x] != 0 && (p < 0 || fires(x) < fires(p))) p = x;
      }
      if (p < 0) continue;
      const int q = strongest_open_pre(p);
      if (q >= 0) add_edge(q, p, -1);
    }
  }

  // Every synapse younger than the initial ones ages by one step.
  void age_edges() {
    for (auto& row : c_->adj)
      for (Edge& e : row)
        if (e.age < kAgeInit) ++e.age;
  }

 private:
  long long fires(int id) const { return id < n_fires_ ? fires_[id] : 0LL; }

  void add_edge(int pre, int post, int64_t log_id) {
    c_->adj[pre].push_back({post, cfg_.delay, w_start_, 0, log_id});
    ++grown_;
  }

  int strongest_open_pre(int p) {
    c_->window_list(p, cfg_.locality_window, scan_);
    int pick = -1;
    long long most = -1;
    for (int q : scan_)
      if (!c_->connected(q, p) && fires(q) > most) {
        most = fires(q);
        pick = q;
      }
    return pick;
  }

  // The window members of post p in window order and their 16 features.
  void window_features(int p, std::vector<int>& qs, std::vector<float>& feats) {
    qs.clear();
    feats.clear();
    c_->window_list(p, cfg_.locality_window, qs);
    if (qs.empty()) return;
    // window maximum, and each member's rank: fire descending, id ascending
    double top = 0.0;
    std::vector<std::pair<double, int>> by_fire;
    by_fire.reserve(qs.size());
    for (int q : qs) {
      top = std::max(top, fx_->fire(q));
      by_fire.push_back({-fx_->fire(q), q});
    }
    std::sort(by_fire.begin(), by_fire.end());
    const double last = by_fire.size() > 1 ? (double)(by_fire.size() - 1) : 0.0;
    rank_of_.assign((size_t)c_->n, 0.0);
    for (size_t r = 0; r < by_fire.size(); ++r)
      rank_of_[by_fire[r].second] = last > 0 ? (double)r / last : 0.0;
    feats.resize(qs.size() * kFeat);
    for (size_t k = 0; k < qs.size(); ++k)
      fx_->features(qs[k], p, top, rank_of_[qs[k]], &feats[k * kFeat]);
  }

  int64_t log_features(const float* f) {
    const int64_t id = (int64_t)c_->log_label.size();
    c_->log_feat.insert(c_->log_feat.end(), f, f + kFeat);
    c_->log_label.push_back(-1);
    c_->log_rank.push_back(-1.0f);
    return id;
  }

  snc_connectome* c_;
  const snc_struct_config& cfg_;
  const int64_t* fires_;
  int64_t n_fires_;
  double norm_;
  bool log_;
  std::unique_ptr<FeatureCtx> fx_;  // built by prune()
  float w_start_ = 0.0f;            // weight of a grown synapse (set by prune())
  std::vector<int> freed_;
  int grown_ = 0;
  std::vector<int> scan_, wq_;
  std::vector<float> wf_;
  std::vector<double> rank_of_;
};

}  // namespace

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
      c->pos[idx] = unit_coord(k, layers[l]);
      if (l == 0) {
        c->role[idx] = SNC_ROLE_INPUT;
        c->channel[idx] = k;
      } else if (l == n_layers - 1) {
        c->role[idx] = SNC_ROLE_OUTPUT;
        c->channel[idx] = k;
      }
    }
  }
  // Every neuron of a non-final layer projects to a run of `span` adjacent
  // neurons of the next layer around its own position there, jittered by up
  // to half a target spacing; span spreads the budget evenly over all
  // projecting neurons (connectome.cpp:38-62, same draws).
  int64_t sources = 0;
  for (int l = 0; l + 1 < n_layers; ++l) sources += layers[l];
  const int per_source = sources > 0 ? std::max(1, (int)(budget / sources)) : 0;
  std::uniform_real_distribution<float> jitter(-0.5f, 0.5f);
  for (int l = 0; l + 1 < n_layers; ++l) {
    const int next = c->layer_count[l + 1];
    const int span = std::min(per_source, next);
    const int last_start = std::max(0, next - span);
    for (int s = 0; s < c->layer_count[l]; ++s) {
      const float at = unit_coord(s, c->layer_count[l]) + jitter(c->rng) / std::max(1, next);
      const int mid = std::clamp((int)std::lround(at * (next - 1)), 0, std::max(0, next - 1));
      const int start = std::clamp(mid - span / 2, 0, last_start);
      std::vector<Edge>& row = c->adj[c->layer_begin[l] + s];
      for (int j = start; j < start + span; ++j)
        row.push_back({c->layer_begin[l + 1] + j, delay, w_init, kAgeInit, -1});
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
      c->pos[idx] = unit_coord(k, cnt);
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

int snc_conn_prune_to(snc_connectome* c, int32_t budget) {
  if (!c || budget < 0) return fail(SNC_E_INVALID, "bad budget");
  const int64_t m = c->num_synapses();
  if (m <= budget) return SNC_OK;
  struct Keyed {
    float mag;
    int64_t at;
  };
  std::vector<Keyed> all;
  all.reserve((size_t)m);
  for (const auto& row : c->adj)
    for (const Edge& e : row) all.push_back({std::fabs(e.weight), (int64_t)all.size()});
  const int64_t cut = m - budget;
  std::nth_element(all.begin(), all.begin() + cut, all.end(), [](const Keyed& x, const Keyed& y) {
    return x.mag != y.mag ? x.mag < y.mag : x.at < y.at;
  });
  std::vector<uint8_t> drop((size_t)m, 0);
  for (int64_t j = 0; j < cut; ++j) drop[(size_t)all[j].at] = 1;
  int64_t at = 0;
  for (auto& row : c->adj) {
    size_t kept = 0;
    for (size_t k = 0; k < row.size(); ++k, ++at)
      if (!drop[(size_t)at]) row[kept++] = row[k];
    row.resize(kept);
  }
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
  if (grow_mode != SNC_GROW_HEURISTIC && grow_mode != SNC_GROW_LEARNED &&
      grow_mode != SNC_GROW_LEARNED_POOL)
    return fail(SNC_E_INVALID, "unknown grow mode");
  std::memset(rep, 0, sizeof(*rep));
  StructuralStep step(c, *cfg, neuron_fires, n_neur, activity_norm, log_grown != 0);
  int rc = step.prune(regrowth_quota(*cfg, reward));
  if (rc) return rc;
  const bool learned = grow_mode != SNC_GROW_HEURISTIC && model != nullptr;
  if (learned)
    rc = step.grow_learned(model, grow_mode == SNC_GROW_LEARNED_POOL, rep);
  else
    step.grow_heuristic();
  if (rc) return rc;
  step.grow_tournament();
  step.age_edges();
  rep->pruned = step.pruned();
  rep->grown = step.grown();
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
