// Host generators and encoders (see include/snc_gen.h).
//
// The workload generators (sheet, hub sheet, random recurrent, drive, images,
// event streams) are this project's own: SplitMix64 / Philox draws only, so
// any host regenerates identical arrays. The reference generators
// (static-snc, random-sparse, dense), encode() and make_synthetic() follow the
// reference's algorithms and libstdc++ draws line for line in effect, because
// their outputs must equal the reference's:
//   make_static_snc     src/graph/snn_graph.cpp:185-236
//   make_random_sparse  src/graph/snn_graph.cpp:288-326
//   make_dense          src/graph/snn_graph.cpp:267-286
//   layered roles       src/graph/snn_graph.cpp:134-152
//   encode              src/training/encoders.cpp:25-71
//   make_synthetic      src/training/datasets.cpp:11-30

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "common.hpp"
#include "rng.hpp"
#include "snc_gen.h"

struct snc_event_buf {
  std::vector<snc_event> ev;
};

namespace snc_amd {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

void pack_csr(int32_t n, std::vector<EdgeRec>& edges, bool sort_rows, GraphArrays& g) {
  if (sort_rows)
    std::stable_sort(edges.begin(), edges.end(), [](const EdgeRec& a, const EdgeRec& b) {
      return a.pre != b.pre ? a.pre < b.pre : a.post < b.post;
    });
  else
    std::stable_sort(edges.begin(), edges.end(),
                     [](const EdgeRec& a, const EdgeRec& b) { return a.pre < b.pre; });
  g.n = n;
  g.row_ptr.assign((size_t)n + 1, 0);
  for (const EdgeRec& e : edges) g.row_ptr[e.pre + 1]++;
  for (int i = 0; i < n; ++i) g.row_ptr[i + 1] += g.row_ptr[i];
  const size_t m = edges.size();
  g.post.resize(m);
  g.weight.resize(m);
  g.delay.resize(m);
  for (size_t k = 0; k < m; ++k) {
    g.post[k] = edges[k].post;
    g.weight[k] = edges[k].w;
    g.delay[k] = edges[k].d;
  }
}

std::string validate_view(const snc_graph_view& g) {
  if (g.num_neurons < 0) return "negative num_neurons";
  if (g.num_neurons > 0 && (!g.row_ptr || !g.sign || !g.role || !g.channel))
    return "per-neuron metadata missing";
  if (g.num_synapses > 0 && (!g.post_ids || !g.weights || !g.delays))
    return "CSR column arrays missing";
  if (g.num_neurons > 0 && g.row_ptr[0] != 0) return "row_ptr[0] != 0";
  if (g.row_ptr && g.row_ptr[g.num_neurons] != g.num_synapses)
    return "row_ptr.back() != num_synapses";
  for (int i = 0; i < g.num_neurons; ++i)
    if (g.row_ptr[i] > g.row_ptr[i + 1]) return "row_ptr not monotone";
  for (int s = 0; s < g.num_synapses; ++s) {
    if (g.post_ids[s] < 0 || g.post_ids[s] >= g.num_neurons) return "post id out of range";
    if (g.delays[s] < 1) return "delay < 1";
  }
  return "";
}

namespace {

double normal(SplitMix64& r) {
  // Box-Muller on two (0,1] uniforms; one value per call keeps streams simple.
  const double u1 = r.uniform_pos(), u2 = r.uniform();
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

struct Pos {
  float x, y;
};

// Near-square grid of `cnt` points stretched over [0, sheet) x [0, sheet)
// with cell centres, so every layer covers the same extent.
std::vector<Pos> layer_positions(int cnt, float sheet) {
  const int cols = std::max(1, (int)std::ceil(std::sqrt((double)cnt)));
  const int rows = std::max(1, (cnt + cols - 1) / cols);
  std::vector<Pos> p(cnt);
  for (int k = 0; k < cnt; ++k) {
    const int cx = k % cols, cy = k / cols;
    p[k].x = (cx + 0.5f) * sheet / cols - 0.5f;
    p[k].y = (cy + 0.5f) * sheet / rows - 0.5f;
  }
  return p;
}

// Draw `k` distinct indices from candidates with weights w (Efraimidis-Spirakis
// A-Res keys ln(u)/w, largest k win). Returns ascending indices.
std::vector<int> weighted_sample(const std::vector<double>& w, int k, SplitMix64& r) {
  std::vector<std::pair<double, int>> key;
  key.reserve(w.size());
  for (int i = 0; i < (int)w.size(); ++i) {
    const double u = r.uniform_pos();
    if (w[i] > 0) key.push_back({std::log(u) / w[i], i});
  }
  k = std::min<int>(k, (int)key.size());
  std::partial_sort(key.begin(), key.begin() + k, key.end(),
                    [](const auto& a, const auto& b) {
                      return a.first != b.first ? a.first > b.first : a.second < b.second;
                    });
  std::vector<int> out(k);
  for (int i = 0; i < k; ++i) out[i] = key[i].second;
  std::sort(out.begin(), out.end());
  return out;
}

void layered_roles(const std::vector<int>& layers, GraphArrays& g) {
  const int L = (int)layers.size();
  int n = 0;
  for (int c : layers) n += c;
  g.role.assign(n, SNC_ROLE_INTERNAL);
  g.channel.assign(n, -1);
  g.sign.assign(n, 1);
  int idx = 0;
  for (int l = 0; l < L; ++l)
    for (int k = 0; k < layers[l]; ++k, ++idx) {
      if (l == 0) {
        g.role[idx] = SNC_ROLE_INPUT;
        g.channel[idx] = k;
      } else if (l == L - 1) {
        g.role[idx] = SNC_ROLE_OUTPUT;
        g.channel[idx] = k;
      }
    }
}

}  // namespace
}  // namespace snc_amd

using namespace snc_amd;

extern "C" {

int snc_graph_buf_view(const snc_graph_buf* b, snc_graph_view* out) {
  if (!b || !out) return fail(SNC_E_INVALID, "null graph buffer");
  *out = b->g.view();
  return SNC_OK;
}
void snc_graph_buf_free(snc_graph_buf* b) { delete b; }

int snc_graph_buf_from_view(const snc_graph_view* g, snc_graph_buf** out) {
  if (!g || !out) return fail(SNC_E_INVALID, "null argument");
  const std::string err = validate_view(*g);
  if (!err.empty()) return fail(SNC_E_INVALID, "graph validation failed: " + err);
  auto* b = new snc_graph_buf();
  const int n = g->num_neurons, m = g->num_synapses;
  b->g.n = n;
  b->g.row_ptr.assign(g->row_ptr, g->row_ptr + n + 1);
  b->g.post.assign(g->post_ids, g->post_ids + m);
  b->g.weight.assign(g->weights, g->weights + m);
  b->g.delay.assign(g->delays, g->delays + m);
  b->g.sign.assign(g->sign, g->sign + n);
  b->g.role.assign(g->role, g->role + n);
  b->g.channel.assign(g->channel, g->channel + n);
  *out = b;
  return SNC_OK;
}

int snc_event_buf_data(const snc_event_buf* b, const snc_event** data, int64_t* n) {
  if (!b || !data || !n) return fail(SNC_E_INVALID, "null event buffer");
  *data = b->ev.data();
  *n = (int64_t)b->ev.size();
  return SNC_OK;
}
void snc_event_buf_free(snc_event_buf* b) { delete b; }

int snc_graph_validate(const snc_graph_view* g) {
  if (!g) return fail(0, "null graph");
  const std::string err = validate_view(*g);
  if (!err.empty()) {
    set_error(err);
    return 0;
  }
  return 1;
}

int snc_gen_sheet(const snc_sheet_params* p, snc_graph_buf** out) {
  if (!p || !out || p->n_layers < 2 || !p->layers) return fail(SNC_E_INVALID, "bad sheet params");
  std::vector<int> layers(p->layers, p->layers + p->n_layers);
  for (int c : layers)
    if (c <= 0) return fail(SNC_E_INVALID, "empty layer");
  auto* b = new snc_graph_buf();
  GraphArrays& g = b->g;
  layered_roles(layers, g);
  const int L = (int)layers.size();
  std::vector<int> base(L, 0);
  for (int l = 1; l < L; ++l) base[l] = base[l - 1] + layers[l - 1];
  std::vector<std::vector<Pos>> pos(L);
  for (int l = 0; l < L; ++l) pos[l] = layer_positions(layers[l], p->sheet);
  const double dmax = std::sqrt(2.0) * p->sheet;
  SplitMix64 rng(p->seed);
  std::vector<EdgeRec> edges;
  auto delay_of = [&](double d) {
    if (p->max_delay <= 0) return std::max(1, (int)p->delay);
    return 1 + (int)std::lround((p->max_delay - 1) * std::min(1.0, d / dmax));
  };
  auto dist = [](const Pos& a, const Pos& b) {
    return std::sqrt((double)(a.x - b.x) * (a.x - b.x) + (double)(a.y - b.y) * (a.y - b.y));
  };
  for (int l = 1; l < L; ++l) {
    const bool last = l == L - 1;
    for (int k = 0; k < layers[l]; ++k) {
      const int post = base[l] + k;
      const Pos& pp = pos[l][k];
      if (last && p->dense_readout) {
        for (int a = 0; a < layers[l - 1]; ++a) {
          const double d = dist(pos[l - 1][a], pp);
          edges.push_back({base[l - 1] + a, post, (float)(normal(rng) * p->w_std), delay_of(d)});
        }
        continue;
      }
      std::vector<double> w(layers[l - 1]);
      for (int a = 0; a < layers[l - 1]; ++a) w[a] = std::exp(-dist(pos[l - 1][a], pp) / p->radius);
      for (int a : weighted_sample(w, p->fan_in, rng)) {
        const double d = dist(pos[l - 1][a], pp);
        edges.push_back({base[l - 1] + a, post, (float)(normal(rng) * p->w_std), delay_of(d)});
      }
      if (!last && p->recurrent_fan_in > 0) {
        std::vector<double> wr(layers[l]);
        for (int a = 0; a < layers[l]; ++a)
          wr[a] = a == k ? 0.0 : std::exp(-dist(pos[l][a], pp) / p->radius);
        for (int a : weighted_sample(wr, p->recurrent_fan_in, rng)) {
          const double d = dist(pos[l][a], pp);
          edges.push_back({base[l] + a, post, (float)(normal(rng) * p->w_std), delay_of(d)});
        }
      }
    }
  }
  int n = 0;
  for (int c : layers) n += c;
  pack_csr(n, edges, true, g);
  *out = b;
  return SNC_OK;
}

int snc_gen_hub_sheet(const snc_hub_params* p, snc_graph_buf** out) {
  if (!p || !out || p->width <= 0 || p->height <= 0 || p->out_degree <= 0)
    return fail(SNC_E_INVALID, "bad hub params");
  const int W = p->width, H = p->height;
  const int64_t n64 = (int64_t)W * H;
  if (n64 > (1 << 30)) return fail(SNC_E_INVALID, "sheet too large");
  const int n = (int)n64;
  const int K = p->out_degree;
  if ((int64_t)n * K > 2000000000LL) return fail(SNC_E_INVALID, "too many synapses");
  SplitMix64 rng(p->seed);

  // Hubs: a random hub_fraction of neurons, each with a power-law in-degree
  // k = k_min (1 - u (1 - (k_max/k_min)^(1-a)))^(1/(1-a)) (inverse CDF).
  const int n_hubs = std::max(0, (int)std::lround(p->hub_fraction * n));
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  for (int i = 0; i < n_hubs; ++i) std::swap(perm[i], perm[i + (int)rng.below(n - i)]);
  std::vector<int> hubs(perm.begin(), perm.begin() + n_hubs);
  std::sort(hubs.begin(), hubs.end());
  // Per-source list of hub targets.
  std::vector<std::vector<int>> hub_out(n);
  const double a = p->hub_alpha, kmin = std::max(1, p->hub_min), kmax = std::max(p->hub_min, p->hub_max);
  std::vector<uint32_t> mark(n, 0);
  uint32_t stamp = 0;
  for (int h : hubs) {
    const double u = rng.uniform();
    int kh = (int)std::floor(kmin * std::pow(1.0 - u * (1.0 - std::pow(kmax / kmin, 1.0 - a)), 1.0 / (1.0 - a)));
    kh = std::clamp(kh, (int)kmin, (int)kmax);
    kh = std::min(kh, n - 1);
    ++stamp;
    int got = 0, tries = 0;
    while (got < kh && tries < 64 * kh) {
      ++tries;
      const int s = (int)rng.below(n);
      if (s == h || mark[s] == stamp || (int)hub_out[s].size() >= K) continue;
      mark[s] = stamp;
      hub_out[s].push_back(h);
      ++got;
    }
  }

  const double dmax = std::sqrt((double)(W - 1) * (W - 1) + (double)(H - 1) * (H - 1));
  auto delay_of = [&](int s, int t) {
    const double dx = (s % W) - (t % W), dy = (s / W) - (t / W);
    const double d = std::sqrt(dx * dx + dy * dy);
    if (p->max_delay <= 1) return 1;
    return 1 + (int)std::lround((p->max_delay - 1) * std::min(1.0, d / std::max(1.0, dmax)));
  };

  auto* b = new snc_graph_buf();
  GraphArrays& g = b->g;
  g.n = n;
  g.row_ptr.resize((size_t)n + 1);
  g.post.resize((size_t)n * K);
  g.weight.resize((size_t)n * K);
  g.delay.resize((size_t)n * K);
  std::vector<int> row;
  row.reserve(K);
  for (int s = 0; s < n; ++s) {
    g.row_ptr[s] = s * K;
    row.assign(hub_out[s].begin(), hub_out[s].end());
    const int sx = s % W, sy = s / W;
    int tries = 0;
    // Local targets: radial distance ~ Gamma(2, radius), i.e. 2-D density
    // proportional to exp(-d / radius); uniform angle; reject off-sheet,
    // self and duplicates.
    while ((int)row.size() < K) {
      ++tries;
      double d = -p->radius * std::log(rng.uniform_pos() * rng.uniform_pos());
      if (tries > 64 * K) d = rng.uniform() * dmax;  // degenerate sheets
      const double th = 6.283185307179586 * rng.uniform();
      const int tx = sx + (int)std::lround(d * std::cos(th));
      const int ty = sy + (int)std::lround(d * std::sin(th));
      if (tx < 0 || tx >= W || ty < 0 || ty >= H) continue;
      const int t = ty * W + tx;
      if (t == s || std::find(row.begin(), row.end(), t) != row.end()) continue;
      row.push_back(t);
    }
    std::sort(row.begin(), row.end());
    for (int j = 0; j < K; ++j) {
      const size_t e = (size_t)s * K + j;
      g.post[e] = row[j];
      g.weight[e] = (float)(normal(rng) * p->w_std);
      g.delay[e] = delay_of(s, row[j]);
    }
  }
  g.row_ptr[n] = n * K;
  g.sign.assign(n, 1);
  g.role.assign(n, p->all_input ? SNC_ROLE_INPUT : SNC_ROLE_INTERNAL);
  g.channel.resize(n);
  for (int i = 0; i < n; ++i) g.channel[i] = p->all_input ? i : -1;
  *out = b;
  return SNC_OK;
}

int snc_gen_random_recurrent(int32_t n_in, int32_t n_hidden, int32_t n_out, float density_in,
                             float density_rec, int32_t delay_min, int32_t delay_max,
                             float w_std, uint64_t seed, snc_graph_buf** out) {
  if (!out || n_in < 0 || n_hidden <= 0 || n_out < 0) return fail(SNC_E_INVALID, "bad sizes");
  const int n = n_in + n_hidden + n_out;
  const int h0 = n_in, o0 = n_in + n_hidden;
  delay_min = std::max(1, delay_min);
  delay_max = std::max(delay_min, delay_max);
  SplitMix64 rng(seed);
  std::vector<EdgeRec> edges;
  const uint32_t p_in = prob_to_u32(density_in), p_rec = prob_to_u32(density_rec);
  for (int a = 0; a < n_in; ++a)
    for (int j = 0; j < n_hidden; ++j)
      if ((uint32_t)(rng.next() >> 32) < p_in)
        edges.push_back({a, h0 + j, (float)(normal(rng) * w_std), 1});
  for (int a = 0; a < n_hidden; ++a)
    for (int j = 0; j < n_hidden; ++j) {
      if (a == j) continue;
      if ((uint32_t)(rng.next() >> 32) < p_rec) {
        const int d = delay_min + (int)rng.below((uint64_t)(delay_max - delay_min + 1));
        edges.push_back({h0 + a, h0 + j, (float)(normal(rng) * w_std), d});
      }
    }
  for (int a = 0; a < n_hidden; ++a)
    for (int k = 0; k < n_out; ++k)
      edges.push_back({h0 + a, o0 + k, (float)(normal(rng) * w_std), 1});
  auto* b = new snc_graph_buf();
  layered_roles({n_in, n_hidden, n_out}, b->g);
  pack_csr(n, edges, true, b->g);
  *out = b;
  return SNC_OK;
}

int snc_gen_static_snc(const int32_t* layers_p, int32_t n_layers, int32_t budget, float weight,
                       int32_t delay, float inhib, uint64_t seed, snc_graph_buf** out) {
  if (!out || !layers_p || n_layers < 1) return fail(SNC_E_INVALID, "bad layers");
  std::vector<int> layers(layers_p, layers_p + n_layers);
  auto* b = new snc_graph_buf();
  GraphArrays& g = b->g;
  layered_roles(layers, g);
  const int n = (int)g.role.size();
  long long fwd_pre = 0;
  for (size_t l = 0; l + 1 < layers.size(); ++l) fwd_pre += layers[l];
  const int fanout = fwd_pre > 0 ? std::max(1, (int)(budget / fwd_pre)) : 0;
  // The reference draws the inhibitory marks first, then one jitter per pre.
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  std::uniform_real_distribution<float> jit(-0.5f, 0.5f);
  for (int i = 0; i < n; ++i)
    if (g.role[i] != SNC_ROLE_INPUT && u01(rng) < inhib) g.sign[i] = -1;
  std::vector<EdgeRec> edges;
  int base = 0;
  for (size_t l = 0; l + 1 < layers.size(); ++l) {
    const int from_n = layers[l], to0 = base + from_n, to_n = layers[l + 1];
    const int k = std::min(fanout, to_n);
    for (int a = 0; a < from_n; ++a) {
      const double rel = from_n > 1 ? (double)a / (from_n - 1) : 0.5;
      const double pa = rel + jit(rng) / std::max(1, to_n);
      int centre = (int)std::lround(pa * (to_n - 1));
      centre = std::clamp(centre, 0, std::max(0, to_n - 1));
      const int lo = std::clamp(centre - k / 2, 0, std::max(0, to_n - k));
      for (int j = 0; j < k; ++j) edges.push_back({base + a, to0 + lo + j, weight, delay});
    }
    base += from_n;
  }
  pack_csr(n, edges, false, g);
  *out = b;
  return SNC_OK;
}

int snc_gen_random_sparse(const int32_t* layers_p, int32_t n_layers, int32_t budget,
                          float weight, int32_t delay, float inhib, uint64_t seed,
                          snc_graph_buf** out) {
  if (!out || !layers_p || n_layers < 1) return fail(SNC_E_INVALID, "bad layers");
  std::vector<int> layers(layers_p, layers_p + n_layers);
  auto* b = new snc_graph_buf();
  GraphArrays& g = b->g;
  layered_roles(layers, g);
  const int n = (int)g.role.size();
  long long possible = 0;
  for (size_t l = 0; l + 1 < layers.size(); ++l) possible += (long long)layers[l] * layers[l + 1];
  const double keep = possible > 0 ? std::min(1.0, (double)budget / possible) : 0.0;
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  for (int i = 0; i < n; ++i)
    if (g.role[i] != SNC_ROLE_INPUT && u01(rng) < inhib) g.sign[i] = -1;
  std::vector<EdgeRec> edges;
  int base = 0;
  for (size_t l = 0; l + 1 < layers.size(); ++l) {
    const int to0 = base + layers[l], to1 = to0 + layers[l + 1];
    for (int a = base; a < to0; ++a)
      for (int t = to0; t < to1; ++t)
        if (u01(rng) < keep) edges.push_back({a, t, weight, delay});
    base += layers[l];
  }
  pack_csr(n, edges, false, g);
  *out = b;
  return SNC_OK;
}

int snc_gen_dense(const int32_t* layers_p, int32_t n_layers, float weight, int32_t delay,
                  snc_graph_buf** out) {
  if (!out || !layers_p || n_layers < 1) return fail(SNC_E_INVALID, "bad layers");
  std::vector<int> layers(layers_p, layers_p + n_layers);
  auto* b = new snc_graph_buf();
  GraphArrays& g = b->g;
  layered_roles(layers, g);
  const int n = (int)g.role.size();
  std::vector<EdgeRec> edges;
  int base = 0;
  for (size_t l = 0; l + 1 < layers.size(); ++l) {
    const int to0 = base + layers[l], to1 = to0 + layers[l + 1];
    for (int a = base; a < to0; ++a)
      for (int t = to0; t < to1; ++t) edges.push_back({a, t, weight, delay});
    base += layers[l];
  }
  pack_csr(n, edges, false, g);
  *out = b;
  return SNC_OK;
}

int snc_gen_background(int32_t n, int32_t steps, double p, float amplitude, uint64_t seed,
                       snc_event_buf** out) {
  if (!out || n < 0 || steps < 0) return fail(SNC_E_INVALID, "bad sizes");
  auto* b = new snc_event_buf();
  const uint32_t thr = prob_to_u32(p);
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  b->ev.reserve((size_t)((double)n * steps * p * 1.05) + 16);
  for (int t = 0; t < steps; ++t)
    for (int i = 0; i < n; i += 4) {
      // One block covers neurons i..i+3 at step t.
      const Philox4 r = philox4x32_10((uint32_t)i, (uint32_t)t, 0x5EEDu, 0u, k0, k1);
      for (int j = 0; j < 4 && i + j < n; ++j)
        if (r.v[j] < thr) b->ev.push_back({0, i + j, t, amplitude});
    }
  *out = b;
  return SNC_OK;
}

void snc_gen_images(int32_t n, int32_t pixels, float zero_frac, uint64_t seed, uint8_t* out) {
  SplitMix64 rng(seed);
  for (int64_t k = 0; k < (int64_t)n * pixels; ++k) {
    const uint64_t r = rng.next();
    const double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
    out[k] = u < zero_frac ? 0 : (uint8_t)(1 + (rng.next() % 255));
  }
}

void snc_gen_prototype_images(int32_t n, int32_t pixels, int32_t classes, float zero_frac,
                              float noise, uint64_t seed, uint8_t* X, int32_t* y) {
  std::vector<uint8_t> proto((size_t)classes * pixels);
  snc_gen_images(classes, pixels, zero_frac, seed, proto.data());
  SplitMix64 rng(seed ^ 0x6A09E667F3BCC909ull);
  for (int i = 0; i < n; ++i) {
    const int c = i % classes;
    y[i] = c;
    for (int k = 0; k < pixels; ++k) {
      uint8_t v = proto[(size_t)c * pixels + k];
      if (rng.uniform() < noise) v = rng.uniform() < zero_frac ? 0 : (uint8_t)(1 + (rng.next() % 255));
      X[(size_t)i * pixels + k] = v;
    }
  }
}

int snc_gen_event_streams(int32_t n, int32_t channels, int32_t classes, int32_t n_centres,
                          float sd, int32_t events_per_sample, int32_t t_min, int32_t t_max,
                          uint64_t seed, int32_t* lengths, int32_t* labels,
                          snc_event_buf** out) {
  if (!out || n < 0 || channels <= 0 || classes <= 0 || n_centres <= 0 || t_min < 1 ||
      t_max < t_min)
    return fail(SNC_E_INVALID, "bad event-stream params");
  SplitMix64 rng(seed);
  // Mixture centres spread evenly over the channel axis, then one random
  // centre order (over time segments) per class.
  std::vector<double> centre(n_centres);
  for (int c = 0; c < n_centres; ++c) centre[c] = (c + 0.5) * channels / n_centres;
  std::vector<std::vector<int>> order(classes, std::vector<int>(n_centres));
  for (int k = 0; k < classes; ++k) {
    std::iota(order[k].begin(), order[k].end(), 0);
    for (int i = n_centres - 1; i > 0; --i) std::swap(order[k][i], order[k][rng.below(i + 1)]);
  }
  auto* b = new snc_event_buf();
  std::vector<snc_event> tmp;
  for (int i = 0; i < n; ++i) {
    const int lab = i % classes;
    const int T = t_min + (int)rng.below((uint64_t)(t_max - t_min + 1));
    lengths[i] = T;
    labels[i] = lab;
    tmp.clear();
    for (int e = 0; e < events_per_sample; ++e) {
      const int t = (int)rng.below((uint64_t)T);
      const int seg = std::min(n_centres - 1, (int)((int64_t)t * n_centres / T));
      const double ch = centre[order[lab][seg]] + sd * normal(rng);
      const int c = (int)std::lround(ch);
      if (c < 0 || c >= channels) continue;
      tmp.push_back({i, c, t, 1.0f});
    }
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](const snc_event& a, const snc_event& b) { return a.time < b.time; });
    b->ev.insert(b->ev.end(), tmp.begin(), tmp.end());
  }
  *out = b;
  return SNC_OK;
}

int snc_encode(int32_t enc, const float* x, int32_t n_ch, int32_t T, float gain,
               float max_rate, float thr, uint64_t seed, int32_t sample_id, snc_event_buf** out) {
  if (!out || (n_ch > 0 && !x) || T < 0) return fail(SNC_E_INVALID, "bad encode args");
  auto* b = new snc_event_buf();
  auto& ev = b->ev;
  switch (enc) {
    case 0:  // direct current on every step
      for (int c = 0; c < n_ch; ++c)
        if (x[c] > thr)
          for (int t = 0; t < T; ++t) ev.push_back({sample_id, c, t, x[c] * gain});
      break;
    case 1: {  // Bernoulli per (channel, step), one mt19937_64 stream
      std::mt19937_64 rng(seed);
      std::uniform_real_distribution<float> u01(0.0f, 1.0f);
      for (int c = 0; c < n_ch; ++c) {
        if (x[c] <= thr) continue;
        const float prob = std::min(1.0f, x[c] * max_rate);
        for (int t = 0; t < T; ++t)
          if (u01(rng) < prob) ev.push_back({sample_id, c, t, gain});
      }
      break;
    }
    case 2:  // time to first spike
      for (int c = 0; c < n_ch; ++c) {
        if (x[c] <= thr) continue;
        const float xc = std::min(1.0f, x[c]);
        const int t = std::clamp((int)std::lround((T - 1) * (1.0f - xc)), 0, std::max(0, T - 1));
        ev.push_back({sample_id, c, t, gain});
      }
      break;
    default:
      delete b;
      return fail(SNC_E_INVALID, "unknown encoder");
  }
  *out = b;
  return SNC_OK;
}

void snc_make_synthetic(int32_t n, int32_t dim, int32_t classes, float noise, uint64_t seed,
                        float* X, int32_t* y) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<float> u01(0.0f, 1.0f);
  std::normal_distribution<float> gauss(0.0f, noise);
  std::vector<float> proto((size_t)classes * dim);
  for (float& v : proto) v = u01(rng);
  for (int i = 0; i < n; ++i) {
    const int c = i % classes;
    for (int k = 0; k < dim; ++k)
      X[(size_t)i * dim + k] = std::clamp(proto[(size_t)c * dim + k] + gauss(rng), 0.0f, 1.0f);
    y[i] = c;
  }
}

const char* snc_hip_last_error(void) { return snc_amd::g_last_error.c_str(); }

}  // extern "C"
