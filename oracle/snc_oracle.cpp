// CPU oracle for SNC's event-driven spiking path -- TEST INFRASTRUCTURE ONLY.
//
// Restates the reference's CPU algorithms so the HIP backend can be checked
// against them on the same inputs, and timed beside them. Compiled with
// -ffp-contract=off: every float expression below is evaluated exactly in the
// order written, which is the ISO order of the reference's source lines.
//
// Functions and the reference lines they follow (paths relative to the
// reference tree):
//   build_injections  src/runtime/cpu_backend.cpp:47-65   (channel -> INPUT neurons, per step)
//   forward_cpu       src/runtime/cpu_backend.cpp:67-191  (LIF, ring, decode, delivery)
//   encode            src/training/encoders.cpp:25-71
//   make_synthetic    src/training/datasets.cpp:11-30
//   philox4x32        Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC'11
//   splitmix64        Steele, Lea, Flood, "Fast splittable PRNGs", OOPSLA'14
//
// Extension beyond the reference (pinned by closed-form tests, not goldens):
//   reset_mode == SNC_RESET_SUBTRACT: a spike subtracts the threshold instead
//   of setting v = reset, and a refractory step holds v instead of resetting.

#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Injection {
  int neuron;
  float amp;
};

// Per-step injection lists in the order the reference applies them: events in
// list order, and for each event the channel's INPUT neurons in ascending id
// order (cpu_backend.cpp:50-63).
std::vector<std::vector<Injection>> build_injections(const snc_graph_view& g,
                                                     const snc_event* ev, int64_t n_ev,
                                                     int steps) {
  int n_in_ch = 0;
  for (int i = 0; i < g.num_neurons; ++i)
    if (g.role[i] == SNC_ROLE_INPUT) n_in_ch = std::max(n_in_ch, g.channel[i] + 1);
  std::vector<std::vector<int>> by_channel(n_in_ch);
  for (int i = 0; i < g.num_neurons; ++i)
    if (g.role[i] == SNC_ROLE_INPUT && g.channel[i] >= 0) by_channel[g.channel[i]].push_back(i);
  std::vector<std::vector<Injection>> out(steps);
  for (int64_t k = 0; k < n_ev; ++k) {
    const snc_event& e = ev[k];
    if (e.time < 0 || e.time >= steps) continue;
    if (e.channel_id < 0 || e.channel_id >= n_in_ch) continue;
    for (int nid : by_channel[e.channel_id]) out[e.time].push_back({nid, e.amplitude});
  }
  return out;
}

int max_delay_of(const snc_graph_view& g) {
  int md = 1;
  for (int s = 0; s < g.num_synapses; ++s) md = std::max(md, g.delays[s]);
  return md;
}

}  // namespace

extern "C" int oracle_forward(const snc_graph_view* gp, const snc_event* events,
                              int64_t n_events, int32_t num_steps, const snc_lif* lifp,
                              int32_t use_omp, snc_forward_capture* cap,
                              snc_forward_stats* stats) {
  if (!gp || !lifp || !stats) return SNC_E_INVALID;
  const snc_graph_view& g = *gp;
  const snc_lif& lif = *lifp;
  const int n = g.num_neurons;
  std::memset(stats, 0, sizeof(*stats));
  stats->used_variant = use_omp ? 1 : 0;

  int n_out_ch = 0;
  for (int i = 0; i < n; ++i)
    if (g.role[i] == SNC_ROLE_OUTPUT) n_out_ch = std::max(n_out_ch, g.channel[i] + 1);
  std::vector<float> logits(n_out_ch, 0.0f);
  const int words = (n + 31) / 32;
  if (cap && cap->raster && num_steps > 0)
    std::memset(cap->raster, 0, sizeof(uint32_t) * (size_t)num_steps * words);
  if (cap && cap->neuron_spikes) std::memset(cap->neuron_spikes, 0, sizeof(int32_t) * n);
  if (n == 0 || num_steps <= 0) {
    if (cap && cap->logits) std::copy(logits.begin(), logits.end(), cap->logits);
    return SNC_OK;
  }

  // ring_len = max_delay + 1 so a deposit never lands on the slot being read
  // and cleared in the same step (cpu_backend.cpp:75-80).
  const int L = max_delay_of(g) + 1;
  std::vector<float> v(n, 0.0f), ext(n, 0.0f), ring((size_t)L * n, 0.0f);
  std::vector<int> refr(n, 0);
  std::vector<uint8_t> fired(n, 0);
  std::vector<int> fired_list;
  fired_list.reserve(n);
  const auto inj = build_injections(g, events, n_events, num_steps);
  const bool subtract = lif.reset_mode == SNC_RESET_SUBTRACT;

  long long spikes = 0, syn_events = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (int t = 0; t < num_steps; ++t) {
    float* slot = &ring[(size_t)(t % L) * n];
    for (const Injection& q : inj[t]) ext[q.neuron] += q.amp;

    // One neuron's integrate/fire; returns whether it fired.
    auto lif_one = [&](int i) -> bool {
      bool f = false;
      if (refr[i] > 0) {
        --refr[i];
        if (!subtract) v[i] = lif.reset;
      } else {
        v[i] = v[i] * lif.decay + ext[i] + slot[i];
        if (v[i] >= lif.threshold) {
          v[i] = subtract ? v[i] - lif.threshold : lif.reset;
          refr[i] = lif.refractory;
          f = true;
        }
      }
      ext[i] = 0.0f;
      slot[i] = 0.0f;
      return f;
    };

    fired_list.clear();
#ifdef _OPENMP
    if (use_omp) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) fired[i] = lif_one(i) ? 1 : 0;
      for (int i = 0; i < n; ++i)
        if (fired[i]) fired_list.push_back(i);
    } else
#endif
    {
      for (int i = 0; i < n; ++i)
        if (lif_one(i)) fired_list.push_back(i);
    }
    spikes += (long long)fired_list.size();
    if (cap && cap->step_spikes) cap->step_spikes[t] = (int64_t)fired_list.size();
    for (int i : fired_list) {
      if (g.role[i] == SNC_ROLE_OUTPUT && g.channel[i] >= 0) logits[g.channel[i]] += 1.0f;
      if (cap && cap->neuron_spikes) cap->neuron_spikes[i] += 1;
      if (cap && cap->raster) cap->raster[(size_t)t * words + (i >> 5)] |= 1u << (i & 31);
    }

    // Expand fired rows into the delay ring: fired ids ascending, CSR order
    // within a row (cpu_backend.cpp:173-181). That order fixes the float sum.
#ifdef _OPENMP
    if (use_omp) {
      long long local = 0;
      const int nf = (int)fired_list.size();
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : local)
      for (int fi = 0; fi < nf; ++fi) {
        const int i = fired_list[fi];
        const float s = (float)g.sign[i];
        for (int e = g.row_ptr[i]; e < g.row_ptr[i + 1]; ++e) {
          float* dst = &ring[(size_t)((t + g.delays[e]) % L) * n + g.post_ids[e]];
#pragma omp atomic
          *dst += s * g.weights[e];
          ++local;
        }
      }
      syn_events += local;
    } else
#endif
    {
      for (int i : fired_list) {
        const float s = (float)g.sign[i];
        for (int e = g.row_ptr[i]; e < g.row_ptr[i + 1]; ++e) {
          ring[(size_t)((t + g.delays[e]) % L) * n + g.post_ids[e]] += s * g.weights[e];
          ++syn_events;
        }
      }
    }
  }
  const auto t1 = std::chrono::steady_clock::now();

  stats->spikes = spikes;
  stats->synaptic_events = syn_events;
  stats->ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (cap && cap->logits) std::copy(logits.begin(), logits.end(), cap->logits);
  if (cap && cap->final_v) std::copy(v.begin(), v.end(), cap->final_v);
  return SNC_OK;
}

extern "C" int oracle_encode(int32_t enc, const float* x, int32_t n_ch, int32_t T, float gain,
                             float max_rate, float thr, uint64_t seed, int32_t sample_id,
                             snc_event* out, int64_t cap, int64_t* n_out) {
  std::vector<snc_event> ev;
  if (enc == 0) {  // direct: constant current every step
    for (int c = 0; c < n_ch; ++c) {
      if (x[c] <= thr) continue;
      for (int t = 0; t < T; ++t) ev.push_back({sample_id, c, t, x[c] * gain});
    }
  } else if (enc == 1) {  // poisson: one shared mt19937_64 stream, channel-major
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<float> u01(0.0f, 1.0f);
    for (int c = 0; c < n_ch; ++c) {
      if (x[c] <= thr) continue;
      const float p = std::min(1.0f, x[c] * max_rate);
      for (int t = 0; t < T; ++t)
        if (u01(rng) < p) ev.push_back({sample_id, c, t, gain});
    }
  } else if (enc == 2) {  // latency: one spike, brighter is earlier
    for (int c = 0; c < n_ch; ++c) {
      if (x[c] <= thr) continue;
      const float xc = std::min(1.0f, x[c]);
      int t = (int)std::lround((T - 1) * (1.0f - xc));
      t = std::clamp(t, 0, T - 1);
      ev.push_back({sample_id, c, t, gain});
    }
  } else {
    return SNC_E_INVALID;
  }
  if (n_out) *n_out = (int64_t)ev.size();
  const int64_t k = std::min<int64_t>(cap, (int64_t)ev.size());
  if (out && k > 0) std::memcpy(out, ev.data(), sizeof(snc_event) * (size_t)k);
  return SNC_OK;
}

extern "C" void oracle_make_synthetic(int32_t n, int32_t dim, int32_t classes, float noise,
                                      uint64_t seed, float* X, int32_t* y) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<float> u01(0.0f, 1.0f);
  std::normal_distribution<float> gauss(0.0f, noise);
  std::vector<float> proto((size_t)classes * dim);
  for (float& p : proto) p = u01(rng);
  for (int i = 0; i < n; ++i) {
    const int c = i % classes;
    for (int k = 0; k < dim; ++k)
      X[(size_t)i * dim + k] = std::clamp(proto[(size_t)c * dim + k] + gauss(rng), 0.0f, 1.0f);
    y[i] = c;
  }
}

extern "C" void oracle_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  // Ten rounds of the Philox-4x32 S-box with the published multipliers and
  // Weyl key increments (Salmon et al. 2011, Table 2 / Section 4.1).
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c0;
    const uint64_t p1 = (uint64_t)M1 * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += W0;
    k1 += W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

extern "C" uint64_t oracle_splitmix64(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
