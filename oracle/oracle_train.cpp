// CPU e-prop oracle -- TEST INFRASTRUCTURE ONLY.
//
// Restates the reference's per-sample e-prop Trainer (paths relative to the
// reference tree):
//   init          src/training/trainer.cpp:10-41   (weights uniform [0, w_init],
//                                                   then feedback B ~ N(0, fs),
//                                                   one mt19937_64(seed) stream)
//   run_sample    src/training/trainer.cpp:52-152  (inject, LIF + surrogate,
//                                                   trace, eligibility in double,
//                                                   delivery of signed weights)
//   apply_update  src/training/trainer.cpp:154-208 (softmax, delta -- supervised
//                                                   or reward advantage --,
//                                                   learning signal,
//                                                   w -= lr/T * L * E)
//   train_epoch   src/training/trainer.cpp:210-249 (shuffle, sample seeds)
// Compiled with -ffp-contract=off so each float expression keeps the
// reference's evaluation order.
//
// Extensions beyond the reference (used by the batched GPU trainer and pinned
// by their own tests): event-list inputs with per-sample lengths T_b,
// subtract reset (as in the forward oracle), a leaky non-spiking readout whose
// logits are the mean membrane over T_b, the minibatch-mean update of the
// reference GPU trainer (cuda_trainer.cu:163-172) with per-sample lr / T_b,
// and the Philox4x32-10 Poisson encoder the device uses.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "oracle.h"

namespace {

struct Trainer {
  // graph (owned copies)
  int n = 0, S = 0;
  std::vector<int32_t> row_ptr, post, delay, channel;
  std::vector<uint8_t> role;
  snc_train_config cfg{};
  int classes = 1, n_in_ch = 0, L = 2;
  std::vector<float> w, B;
  std::vector<int> internal_idx;
  std::vector<std::vector<int>> chan_to_neurons;
  std::vector<int64_t> syn_deliv, neur_fire;
  std::mt19937_64 action_rng;  // reward mode (trainer.cpp:39-40)
  float baseline = 0.0f;

  bool leaky_out(int j) const {
    return cfg.readout_mode == SNC_READOUT_LEAKY && role[j] == SNC_ROLE_OUTPUT;
  }

  struct SampleOut {
    std::vector<float> logits;
    long long spikes = 0, events = 0;
  };

  // One sample over T steps from per-step injections. elig (size S) is
  // accumulated when non-null; raster [T][ceil(n/32)] when non-null.
  SampleOut run(const std::vector<std::vector<std::pair<int, float>>>& by_step, int T,
                std::vector<double>* elig, uint32_t* raster, bool stats) {
    const float thr = cfg.lif.threshold, decay = cfg.lif.decay, reset = cfg.lif.reset;
    const int R = cfg.lif.refractory;
    const float gamma = cfg.surrogate_scale;
    const bool subtract = cfg.lif.reset_mode == SNC_RESET_SUBTRACT;
    const int words = (n + 31) / 32;
    std::vector<float> v(n, 0.0f), ext(n, 0.0f), tr(n, 0.0f), psi(n, 0.0f);
    std::vector<int> refr(n, 0);
    std::vector<uint8_t> fired(n, 0);
    std::vector<float> ring((size_t)L * n, 0.0f);
    SampleOut o;
    o.logits.assign(classes, 0.0f);
    std::vector<float> vsum(classes, 0.0f);
    for (int t = 0; t < T; ++t) {
      float* slot = &ring[(size_t)(t % L) * n];
      for (const auto& q : by_step[t]) ext[q.first] += q.second;
      for (int j = 0; j < n; ++j) {
        float vpre;
        if (leaky_out(j)) {
          // Leaky integrator: same membrane recursion, never fires.
          v[j] = decay * v[j] + ext[j] + slot[j];
          fired[j] = 0;
          psi[j] = cfg.readout_psi;
          if (channel[j] >= 0 && channel[j] < classes) vsum[channel[j]] += v[j];
          ext[j] = 0.0f;
          slot[j] = 0.0f;
          continue;
        }
        bool in_refr = false;
        if (refr[j] > 0) {
          --refr[j];
          in_refr = true;
          if (!subtract) v[j] = reset;
          vpre = subtract ? v[j] : reset;
          fired[j] = 0;
        } else {
          v[j] = decay * v[j] + ext[j] + slot[j];
          vpre = v[j];
          if (v[j] >= thr) {
            v[j] = subtract ? v[j] - thr : reset;
            refr[j] = R;
            fired[j] = 1;
          } else {
            fired[j] = 0;
          }
        }
        ext[j] = 0.0f;
        slot[j] = 0.0f;
        const float d = std::fabs(vpre - thr) / (thr > 0 ? thr : 1.0f);
        psi[j] = (subtract && in_refr) ? 0.0f : (d < 1.0f ? gamma * (1.0f - d) : 0.0f);
        if (fired[j]) {
          ++o.spikes;
          if (stats) ++neur_fire[j];
          if (raster) raster[(size_t)t * words + (j >> 5)] |= 1u << (j & 31);
          if (cfg.readout_mode == SNC_READOUT_SPIKING && role[j] == SNC_ROLE_OUTPUT &&
              channel[j] >= 0)
            o.logits[channel[j]] += 1.0f;
        }
      }
      for (int j = 0; j < n; ++j)
        tr[j] = decay * tr[j] + (1.0f - decay) * (fired[j] ? 1.0f : 0.0f);
      if (elig) {
        for (int i = 0; i < n; ++i) {
          const float tri = tr[i];
          if (tri == 0.0f) continue;
          for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
            (*elig)[e] += (double)psi[post[e]] * tri;
        }
      }
      for (int i = 0; i < n; ++i) {
        if (!fired[i]) continue;
        for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
          ring[(size_t)((t + delay[e]) % L) * n + post[e]] += w[e];
          ++o.events;
          if (stats) ++syn_deliv[e];
        }
      }
    }
    if (cfg.readout_mode == SNC_READOUT_LEAKY)
      for (int k = 0; k < classes; ++k) o.logits[k] = vsum[k] / (float)std::max(1, T);
    return o;
  }

  std::vector<std::vector<std::pair<int, float>>> schedule(const snc_event* ev, int64_t n_ev,
                                                           int T, int sample) const {
    std::vector<std::vector<std::pair<int, float>>> by_step(std::max(0, T));
    for (int64_t k = 0; k < n_ev; ++k) {
      const snc_event& e = ev[k];
      if (sample >= 0 && e.sample_id != sample) continue;
      if (e.time < 0 || e.time >= T) continue;
      if (e.channel_id < 0 || e.channel_id >= n_in_ch) continue;
      for (int nid : chan_to_neurons[e.channel_id]) by_step[e.time].push_back({nid, e.amplitude});
    }
    return by_step;
  }

  // delta = softmax(logits) - onehot(label), or in reward mode
  // (R - baseline) * (softmax - onehot(a)) with a ~ softmax drawn from the
  // action stream, R = 1[a == label], then the baseline's EMA step
  // (trainer.cpp:158-180).
  std::vector<float> delta_of(const std::vector<float>& logits, int label) {
    std::vector<float> p(classes);
    const float mx = *std::max_element(logits.begin(), logits.end());
    double sum = 0.0;
    for (int k = 0; k < classes; ++k) {
      p[k] = std::exp(logits[k] - mx);
      sum += p[k];
    }
    for (int k = 0; k < classes; ++k) p[k] /= (float)sum;
    std::vector<float> d(classes);
    if (cfg.reward_mode) {
      std::uniform_real_distribution<float> u(0.0f, 1.0f);
      const float r = u(action_rng);
      float cdf = 0.0f;
      int a = classes - 1;
      for (int k = 0; k < classes; ++k) {
        cdf += p[k];
        if (r <= cdf) {
          a = k;
          break;
        }
      }
      const float R = a == label ? 1.0f : 0.0f;
      const float adv = R - baseline;
      for (int k = 0; k < classes; ++k) d[k] = adv * (p[k] - (k == a ? 1.0f : 0.0f));
      baseline = cfg.reward_baseline_decay * baseline + (1.0f - cfg.reward_baseline_decay) * R;
      return d;
    }
    for (int k = 0; k < classes; ++k) d[k] = p[k] - (k == label ? 1.0f : 0.0f);
    return d;
  }

  // Per-neuron learning signal (trainer.cpp:182-195).
  std::vector<float> signal_of(const std::vector<float>& delta) const {
    std::vector<float> Ls(n, 0.0f);
    for (int j = 0; j < n; ++j) {
      if (role[j] == SNC_ROLE_OUTPUT && channel[j] >= 0 && channel[j] < classes) {
        Ls[j] = delta[channel[j]];
      } else if (cfg.train_hidden && internal_idx[j] >= 0) {
        const float* b = &B[(size_t)internal_idx[j] * classes];
        float s = 0.0f;
        for (int k = 0; k < classes; ++k) s += b[k] * delta[k];
        Ls[j] = s;
      }
    }
    return Ls;
  }

  // Per-sample SGD step (trainer.cpp:197-207).
  void apply_update(const std::vector<double>& elig, const std::vector<float>& logits, int label,
                    int T) {
    const std::vector<float> Ls = signal_of(delta_of(logits, label));
    const float scale = cfg.lr / std::max(1, T);
    for (int i = 0; i < n; ++i)
      for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
        const float lpost = Ls[post[e]];
        if (lpost == 0.0f) continue;
        w[e] -= scale * lpost * (float)elig[e];
        w[e] = std::clamp(w[e], -cfg.w_max, cfg.w_max);
      }
  }
};

}  // namespace

extern "C" {

void* oracle_trainer_create(const snc_graph_view* g, const snc_train_config* cfg) {
  auto* t = new Trainer();
  t->n = g->num_neurons;
  t->S = g->num_synapses;
  t->row_ptr.assign(g->row_ptr, g->row_ptr + t->n + 1);
  t->post.assign(g->post_ids, g->post_ids + t->S);
  t->delay.assign(g->delays, g->delays + t->S);
  t->role.assign(g->role, g->role + t->n);
  t->channel.assign(g->channel, g->channel + t->n);
  t->cfg = *cfg;
  int n_out = 0, md = 1;
  for (int i = 0; i < t->n; ++i) {
    if (t->role[i] == SNC_ROLE_OUTPUT) n_out = std::max(n_out, t->channel[i] + 1);
    if (t->role[i] == SNC_ROLE_INPUT) t->n_in_ch = std::max(t->n_in_ch, t->channel[i] + 1);
  }
  for (int e = 0; e < t->S; ++e) md = std::max(md, t->delay[e]);
  t->L = md + 1;
  t->classes = std::max(1, n_out);
  std::mt19937_64 rng(cfg->seed);
  std::uniform_real_distribution<float> wu(0.0f, cfg->w_init);
  t->w.resize(t->S);
  for (float& x : t->w) x = wu(rng);
  t->syn_deliv.assign(t->S, 0);
  t->neur_fire.assign(t->n, 0);
  t->chan_to_neurons.assign(t->n_in_ch, {});
  for (int i = 0; i < t->n; ++i)
    if (t->role[i] == SNC_ROLE_INPUT && t->channel[i] >= 0)
      t->chan_to_neurons[t->channel[i]].push_back(i);
  t->internal_idx.assign(t->n, -1);
  int n_int = 0;
  for (int i = 0; i < t->n; ++i)
    if (t->role[i] == SNC_ROLE_INTERNAL) t->internal_idx[i] = n_int++;
  std::normal_distribution<float> bn(0.0f, cfg->feedback_scale);
  t->B.resize((size_t)n_int * t->classes);
  for (float& b : t->B) b = bn(rng);
  t->action_rng.seed(cfg->seed ^ 0xA5A5C0FFEEULL);
  t->baseline = 1.0f / t->classes;
  return t;
}

void oracle_trainer_free(void* h) { delete static_cast<Trainer*>(h); }

void oracle_trainer_get_weights(void* h, float* w) {
  auto* t = static_cast<Trainer*>(h);
  std::copy(t->w.begin(), t->w.end(), w);
}
void oracle_trainer_set_weights(void* h, const float* w) {
  auto* t = static_cast<Trainer*>(h);
  std::copy(w, w + t->S, t->w.begin());
}
void oracle_trainer_get_feedback(void* h, float* B) {
  auto* t = static_cast<Trainer*>(h);
  std::copy(t->B.begin(), t->B.end(), B);
}
void oracle_trainer_reset_stats(void* h) {
  auto* t = static_cast<Trainer*>(h);
  std::fill(t->syn_deliv.begin(), t->syn_deliv.end(), 0);
  std::fill(t->neur_fire.begin(), t->neur_fire.end(), 0);
}
void oracle_trainer_stats(void* h, int64_t* syn, int64_t* neur) {
  auto* t = static_cast<Trainer*>(h);
  std::copy(t->syn_deliv.begin(), t->syn_deliv.end(), syn);
  std::copy(t->neur_fire.begin(), t->neur_fire.end(), neur);
}

// One sample from an event list (events with sample_id == sample, or all
// when sample < 0). Fills logits [classes]; elig [S] (double, overwritten) and
// raster when non-null. Returns spikes; *events gets synaptic events.
int64_t oracle_trainer_run_sample(void* h, const snc_event* ev, int64_t n_ev, int32_t sample,
                                  int32_t T, float* logits, double* elig, uint32_t* raster,
                                  int64_t* events, int32_t record_stats) {
  auto* t = static_cast<Trainer*>(h);
  std::vector<double> E;
  if (elig) E.assign(t->S, 0.0);
  auto o = t->run(t->schedule(ev, n_ev, T, sample), T, elig ? &E : nullptr, raster,
                  record_stats != 0);
  if (logits) std::copy(o.logits.begin(), o.logits.end(), logits);
  if (elig) std::copy(E.begin(), E.end(), elig);
  if (events) *events = o.events;
  return o.spikes;
}

// Reference per-sample SGD epoch on feature rows (trainer.cpp:210-249): the
// reference's shuffle and per-sample seeds, encode() with mt19937_64 draws.
double oracle_trainer_train_epoch_ref(void* h, const float* X, const int32_t* y, int32_t ns,
                                      int32_t dim, int32_t epoch, double* loss,
                                      int64_t* spikes, int64_t* events) {
  auto* t = static_cast<Trainer*>(h);
  std::vector<int> order(ns);
  std::iota(order.begin(), order.end(), 0);
  std::mt19937_64 rng(t->cfg.seed ^ (0x9e3779b97f4a7c15ULL * (uint64_t)(epoch + 1)));
  std::shuffle(order.begin(), order.end(), rng);
  const int T = t->cfg.num_steps;
  double loss_sum = 0.0;
  int correct = 0;
  long long sp = 0, evn = 0;
  std::vector<double> E(t->S);
  for (int idx : order) {
    const uint64_t sseed = t->cfg.seed ^ (0xD1B54A32D192ED03ULL * (uint64_t)(idx + 1)) ^
                           ((uint64_t)epoch << 32);
    int64_t nev = 0;
    oracle_encode(t->cfg.encoder, X + (size_t)idx * dim, t->n_in_ch, T, t->cfg.enc_gain,
                  t->cfg.enc_max_rate, t->cfg.enc_threshold, sseed, 0, nullptr, 0, &nev);
    std::vector<snc_event> ev((size_t)nev);
    oracle_encode(t->cfg.encoder, X + (size_t)idx * dim, t->n_in_ch, T, t->cfg.enc_gain,
                  t->cfg.enc_max_rate, t->cfg.enc_threshold, sseed, 0, ev.data(), nev, &nev);
    std::fill(E.begin(), E.end(), 0.0);
    auto o = t->run(t->schedule(ev.data(), nev, T, -1), T, &E, nullptr, true);
    sp += o.spikes;
    evn += o.events;
    const auto& c = o.logits;
    const float mx = *std::max_element(c.begin(), c.end());
    double s = 0.0;
    for (int k = 0; k < t->classes; ++k) s += std::exp(c[k] - mx);
    loss_sum += -(c[y[idx]] - mx) + std::log(s);
    int pred = 0;
    for (int k = 1; k < t->classes; ++k)
      if (c[k] > c[pred]) pred = k;
    if (pred == y[idx]) ++correct;
    t->apply_update(E, c, y[idx], T);
  }
  if (loss) *loss = ns ? loss_sum / ns : 0.0;
  if (spikes) *spikes = sp;
  if (events) *events = evn;
  return ns ? (double)correct / ns : 0.0;
}

// Minibatch-mean e-prop step from the current weights: every sample b runs
// its events (sample_id == b) for lengths[b] steps; then
//   c_{b,s} = (lr / T_b * L_b[post_s]) * float(E_{b,s})
//   w_s     = clamp(w_s - (sum_b c_{b,s}) / B, +-w_max)   (b ascending, float).
// In reward mode L_b comes from the reward delta, drawn and baselined in
// sample order; without apply no delta (and no action draw) is made.
// logits_out [B][classes]; raster_out [B][T_max][words] when non-null.
void oracle_trainer_minibatch(void* h, const snc_event* ev, int64_t n_ev, const int32_t* lengths,
                              const int32_t* labels, int32_t Bn, int32_t apply, float* logits_out,
                              uint32_t* raster_out, int32_t t_max, int64_t* spikes,
                              int64_t* events) {
  auto* t = static_cast<Trainer*>(h);
  const int words = (t->n + 31) / 32;
  std::vector<float> acc(t->S, 0.0f);
  std::vector<double> E(t->S);
  long long sp = 0, evn = 0;
  for (int b = 0; b < Bn; ++b) {
    std::fill(E.begin(), E.end(), 0.0);
    const int T = lengths[b];
    auto o = t->run(t->schedule(ev, n_ev, T, b), T, &E,
                    raster_out ? raster_out + (size_t)b * t_max * words : nullptr, true);
    sp += o.spikes;
    evn += o.events;
    if (logits_out) std::copy(o.logits.begin(), o.logits.end(), logits_out + (size_t)b * t->classes);
    if (!apply) continue;  // no update, no action draw
    const std::vector<float> Ls = t->signal_of(t->delta_of(o.logits, labels[b]));
    const float scale = t->cfg.lr / (float)std::max(1, T);
    for (int s = 0; s < t->S; ++s) acc[s] += (scale * Ls[t->post[s]]) * (float)E[s];
  }
  if (apply)
    for (int s = 0; s < t->S; ++s)
      t->w[s] = std::clamp(t->w[s] - acc[s] / (float)Bn, -t->cfg.w_max, t->cfg.w_max);
  if (spikes) *spikes = sp;
  if (events) *events = evn;
}

// Philox Poisson encoding of one sample, as the device does it: channel c
// fires at step t iff word (c % 4) of philox4x32_10(ctr = {c / 4, t, sample,
// epoch}, key = seed) is below p * 2^32 with p = min(1, x_c * max_rate).
int64_t oracle_encode_philox(const float* x, int32_t n_ch, int32_t T, float gain, float max_rate,
                             float thr, uint64_t seed, int32_t sample, int32_t epoch,
                             snc_event* out, int64_t cap) {
  int64_t k = 0;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int t = 0; t < T; ++t)
    for (int cb = 0; cb < (n_ch + 3) / 4; ++cb) {
      const uint32_t ctr[4] = {(uint32_t)cb, (uint32_t)t, (uint32_t)sample, (uint32_t)epoch};
      uint32_t r[4];
      oracle_philox4x32(ctr, key, r);
      for (int j = 0; j < 4; ++j) {
        const int c = cb * 4 + j;
        if (c >= n_ch || !(x[c] > thr)) continue;
        const float p = std::min(1.0f, x[c] * max_rate);
        const double p32 = (double)p * 4294967296.0;
        const uint32_t lim = p32 >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)p32;
        if (r[j] < lim) {
          if (k < cap && out) out[k] = {sample, c, t, gain};
          ++k;
        }
      }
    }
  return k;
}

}  // extern "C"
