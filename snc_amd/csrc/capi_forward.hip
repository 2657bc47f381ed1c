// C ABI of the forward path: device graph upload, forward plans, one-shot
// forward. Host code; the kernels live in forward.hip / sort.hip / persist.hip.
//
// Reference contract kept (src/runtime/cuda_backend.cu:186-361): synchronous,
// validated input, non-zero status on any allocation or launch failure so the
// caller can fall back, stats.ms covering only the step loop. Differences: the
// graph is uploaded once and stays resident (snc_hip_graph_upload), and a plan
// keeps one sample's inputs resident across runs.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "autotune.hpp"
#include "common.hpp"
#include "hip_err.hpp"
#include "pull.hpp"
#include "device_graph.hpp"
#include "forward_plan.hpp"
#include "persist.hpp"
#include "snc_hip.h"

namespace snc_amd {

namespace {

template <typename T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    hip_clear();
    return fail(SNC_E_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
  }
  account_alloc(count * sizeof(T));
  return SNC_OK;
}

template <typename T>
int dupload(T** p, const T* h, size_t count) {
  int rc = dalloc(p, count);
  if (rc) return rc;
  if (count && hipMemcpy(*p, h, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return fail(SNC_E_HIP, "hipMemcpy H2D failed");
  return SNC_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

bool device_ok() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return false;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return false;
  return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

}  // namespace

DevGraph::~DevGraph() {
  dfree(row_ptr);
  dfree(post);
  dfree(wsig);
  dfree(delay);
  dfree(cpos);
  dfree(ccol);
  dfree(cval);
  dfree(cdel);
  dfree(heavy_ids);
  dfree(is_heavy);
}

}  // namespace snc_amd

using namespace snc_amd;

struct snc_hip_plan;
static int sort_alloc(snc_hip_plan* P, const std::vector<int32_t>& inj_ptr,
                      const std::vector<int32_t>& inj_nid);
static int autotune_search(snc_hip_graph* G, const snc_event* events, int64_t n_events,
                           int num_samples, int num_steps, const snc_lif* lif, bool exact_only,
                           snc_plan_config* out);

struct snc_hip_plan {
  snc_hip_graph* G = nullptr;
  FwdPlanCore core;
  StepArgs a{};
  int T = 0;
  int B = 1;  // samples (grid.y of the step kernels)
  snc_lif lif{};
  int variant = SNC_VARIANT_ATOMIC;
  snc_plan_config cfg{};
  bool persistent = false;
  PersistConfig pcfg{};
  // hub set of this plan's threshold (hip-bucket and the persistent kernel)
  std::vector<int32_t> heavy_host;
  int32_t* heavy_ids = nullptr;
  uint8_t* is_heavy = nullptr;
  std::vector<int32_t> h_inj_ptr;
  // owned device buffers
  int32_t *inj_ptr = nullptr, *inj_nid = nullptr;
  float* inj_amp = nullptr;
  float *v = nullptr, *ext = nullptr, *ring = nullptr;
  int32_t *refr = nullptr, *fired = nullptr, *fcnt = nullptr, *nspk = nullptr;
  uint8_t* dirty = nullptr;
  uint32_t* bits = nullptr;
  size_t n_bits_words = 0;
  unsigned long long *blk_spikes = nullptr, *blk_events = nullptr;
  uint32_t* raster = nullptr;
  uint32_t* fbits = nullptr;
  int32_t* heavy_blk_ptr = nullptr;
  int32_t *inj_blk_ptr = nullptr, *blk_nid = nullptr;
  float *blk_amp = nullptr, *inj_heavy = nullptr;
  int2* inj_tab = nullptr;
  uint32_t* F = nullptr;      // hip-pull: fired words [L][n], 4 or 8 bytes each
  float* hub_list = nullptr;  // hip-pull: hub arrivals per chunk and sample
  uint16_t* hub_cnt = nullptr;  // hip-pull: their counts
  size_t F_words = 0;         // hip-pull: 32-bit words of F
  uint16_t* fsum = nullptr;   // hip-pull: fired summary [L][sum_words]
  size_t fsum_total = 0;
  PullArgs pa{};
  size_t n_blk_slots = 0;
  int halves = 1;
  size_t dirty_stride = 0;
  size_t bits_total = 0, dirty_total = 0, ring_total = 0;  // elements allocated
  std::vector<int64_t> last_spikes, last_events;           // per sample, last run
  int4* sseg = nullptr;  // hip-sort segments, each synapse's segment, lists, counters
  int32_t *sseg_e = nullptr, *scnt = nullptr, *sinj_ptr = nullptr;
  uint2* slist = nullptr;
  int n_seg = 0;
  std::vector<int32_t> scnt_init;  // scnt at reset: each counter at its segment's first column position
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};  // [raster off, on]
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int64_t dev_bytes = 0;  // device memory this plan allocated
  double kernel_avg_ms = 0.0;
  int64_t kernel_launches = 0;

  ~snc_hip_plan() {
    for (auto& g : graph_exec)
      if (g) (void)hipGraphExecDestroy(g);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (core.stream) (void)hipStreamDestroy(core.stream);
    dfree(inj_ptr);
    dfree(inj_nid);
    dfree(inj_amp);
    dfree(v);
    dfree(ext);
    dfree(ring);
    dfree(refr);
    dfree(fired);
    dfree(fcnt);
    dfree(nspk);
    dfree(dirty);
    dfree(bits);
    dfree(blk_spikes);
    dfree(blk_events);
    dfree(raster);
    dfree(fbits);
    dfree(heavy_blk_ptr);
    dfree(inj_blk_ptr);
    dfree(blk_nid);
    dfree(blk_amp);
    dfree(inj_tab);
    dfree(F);
    dfree(hub_list);
    dfree(hub_cnt);
    dfree(fsum);
    dfree(inj_heavy);
    dfree(heavy_ids);
    dfree(is_heavy);
    dfree(sseg);
    dfree(sseg_e);
    dfree(slist);
    dfree(scnt);
    dfree(sinj_ptr);
  }
};

// State of the hip-sort strategy: the target segments, their arrival queues
// and counters per ring slot, and each step's injections per segment. A
// queue of slot s holds at most one record per column position of its
// segment (a synapse deposits into slot s once, and the queue is emptied at
// the step that reads slot s before the synapse can deposit into it again),
// so the queues of one slot share one P-sized array, segment s at its first
// column position.
static int sort_alloc(snc_hip_plan* P, const std::vector<int32_t>& inj_ptr,
                      const std::vector<int32_t>& inj_nid) {
  const DevGraph& g = P->G->g;
  std::vector<int32_t> ccol((size_t)g.n + 1, 0);
  for (int j = 0; j < g.n; ++j) ccol[j + 1] = ccol[j] + (g.in_degree[j] + 31) / 32 * 32;
  std::vector<int4> seg;
  std::vector<int32_t> sow;
  int rc = sort_prepare();
  if (!rc) rc = sort_build_segments(ccol, g.n, seg, sow);
  if (rc) return rc;
  // each synapse's segment, from its column position (host copy of cpos)
  std::vector<int32_t> cpos((size_t)std::max<int64_t>(1, g.S), 0), sse((size_t)std::max<int64_t>(1, g.S), 0);
  if (g.S > 0 && hipMemcpy(cpos.data(), g.cpos, sizeof(int32_t) * (size_t)g.S,
                           hipMemcpyDeviceToHost) != hipSuccess)
    return fail(SNC_E_HIP, "column positions readback failed");
  for (int64_t e = 0; e < g.S; ++e) sse[(size_t)e] = sow[(size_t)(cpos[(size_t)e] >> 5)];
  P->n_seg = (int)seg.size();
  if (!rc) rc = dupload(&P->sseg, seg.data(), seg.size());
  if (!rc) rc = dupload(&P->sseg_e, sse.data(), sse.size());
  const size_t L = (size_t)g.L;
  if (L * (size_t)std::max<int64_t>(1, g.P) * sizeof(uint2) > (size_t)kSortQueueBytesMax)
    return fail(SNC_E_UNSUPPORTED, "hip-sort arrival queues (ring slots x column positions x 8 B) "
                                   "above 64 GiB");
  if (!rc) rc = dalloc(&P->slist, L * (size_t)std::max<int64_t>(1, g.P));
  if (!rc) rc = dalloc(&P->scnt, L * (size_t)std::max(1, P->n_seg) * kSortCntStride);
  P->scnt_init.assign(L * (size_t)P->n_seg * kSortCntStride, 0);
  for (size_t sl = 0; sl < L; ++sl)
    for (int q = 0; q < P->n_seg; ++q)
      P->scnt_init[(sl * P->n_seg + q) * kSortCntStride] = seg[q].z;
  const std::vector<int32_t> sinj = sort_injection_ranges(seg, inj_ptr, inj_nid, P->T);
  if (!rc) rc = dupload(&P->sinj_ptr, sinj.data(), sinj.size());
  return rc;
}

extern "C" {

int snc_hip_available(void) { return device_ok() ? 1 : 0; }

int snc_hip_graph_upload(const snc_graph_view* gv, snc_hip_graph** out) {
  if (!gv || !out) return fail(SNC_E_INVALID, "null argument");
  *out = nullptr;
  if (!device_ok()) return fail(SNC_E_NODEVICE, "no gfx950 device");
  const std::string err = validate_view(*gv);
  if (!err.empty()) return fail(SNC_E_INVALID, "graph validation failed: " + err);
  const int n = gv->num_neurons;
  const int64_t S = gv->num_synapses;
  std::unique_ptr<snc_hip_graph> H(new snc_hip_graph());
  DevGraph& g = H->g;
  AllocAccount acct(&g.dev_bytes);
  g.n = n;
  g.S = S;
  int md = 1;
  for (int64_t e = 0; e < S; ++e) md = std::max(md, (int)gv->delays[e]);
  if (md > 255) return fail(SNC_E_UNSUPPORTED, "delays above 255 steps are not supported");
  g.max_delay = md;
  g.L = md + 1;

  std::vector<float> wsig(S);
  std::vector<uint8_t> del8(S);
  g.out_degree.assign(n, 0);
  g.in_degree.assign(n, 0);
  for (int i = 0; i < n; ++i) {
    const float s = (float)gv->sign[i];
    g.out_degree[i] = gv->row_ptr[i + 1] - gv->row_ptr[i];
    for (int e = gv->row_ptr[i]; e < gv->row_ptr[i + 1]; ++e) {
      wsig[e] = s * gv->weights[e];
      del8[e] = (uint8_t)gv->delays[e];
      g.in_degree[gv->post_ids[e]]++;
    }
  }
  // Padded incoming columns, ordered (delay descending, pre ascending, CSR
  // position ascending). For one arrival step the CPU adds a target's
  // deposits in order of their origin step (earlier first, i.e. longer delay
  // first) and, within an origin step, in CSR order; so walking the column
  // in this order over the deposits that arrive together reproduces the
  // CPU's sum. Per origin step (one delay) it is the plain CSR order, which
  // the ring-based folds rely on.
  std::vector<int32_t> ccol((size_t)n + 1, 0);
  int64_t P = 0;
  for (int j = 0; j < n; ++j) {
    ccol[j] = (int32_t)P;
    P += ((int64_t)g.in_degree[j] + 31) / 32 * 32;
    if (P > 0x7fffff00LL) return fail(SNC_E_UNSUPPORTED, "graph too large for 32-bit columns");
  }
  ccol[n] = (int32_t)P;
  g.P = P;
  std::vector<int32_t> cpos(S), fillc(n, 0), csr_order(S);
  std::vector<int64_t> cstart((size_t)n + 1, 0);
  for (int j = 0; j < n; ++j) cstart[j + 1] = cstart[j] + g.in_degree[j];
  for (int64_t e = 0; e < S; ++e) {  // each column's synapses in CSR order
    const int j = gv->post_ids[e];
    csr_order[cstart[j] + fillc[j]++] = (int32_t)e;
  }
  std::vector<float> cval((size_t)P, 0.0f);
  std::vector<uint8_t> cdel((size_t)P, 0);
  {
    int32_t cnt[257];
    for (int j = 0; j < n; ++j) {  // stable counting sort by delay, descending
      std::fill(cnt, cnt + 257, 0);
      for (int64_t k = cstart[j]; k < cstart[j + 1]; ++k) cnt[255 - del8[csr_order[k]]]++;
      int32_t run = ccol[j];
      for (int b = 0; b < 256; ++b) {
        const int32_t c = cnt[b];
        cnt[b] = run;
        run += c;
      }
      for (int64_t k = cstart[j]; k < cstart[j + 1]; ++k) {
        const int32_t e = csr_order[k];
        const int32_t p = cnt[255 - del8[e]]++;
        cpos[e] = p;
        cval[p] = wsig[e];
        cdel[p] = del8[e];
      }
    }
  }
  // Hubs at the default threshold, for the profile (plans build their own).
  std::vector<int32_t> heavy;
  int64_t heavy_syn = 0;
  for (int j = 0; j < n; ++j)
    if (g.in_degree[j] > g.heavy_threshold) {
      heavy.push_back(j);
      heavy_syn += g.in_degree[j];
    }
  g.n_heavy = (int)heavy.size();
  g.heavy_host = heavy;

  // Device row order: each row's synapses stably sorted by post. The lanes
  // of a wave expanding a row then deposit into neighbouring words, so one
  // atomic instruction touches few cache lines; scattered device atomics
  // retire per line, not per lane. Results do not depend on this order: the
  // ordered folds follow column positions (cpos), and a row's synapses to one
  // post keep their CSR order.
  std::vector<int32_t> dpost(S), dcpos(S);
  std::vector<float> dwsig(S);
  std::vector<uint8_t> ddel(S);
  {
    std::vector<int32_t> ord;
    for (int i = 0; i < n; ++i) {
      const int rb = gv->row_ptr[i], re = gv->row_ptr[i + 1];
      ord.resize(re - rb);
      for (int e = rb; e < re; ++e) ord[e - rb] = e;
      std::stable_sort(ord.begin(), ord.end(),
                       [&](int32_t x, int32_t y) { return gv->post_ids[x] < gv->post_ids[y]; });
      for (int k = 0; k < re - rb; ++k) {
        const int32_t e = ord[k];
        dpost[rb + k] = gv->post_ids[e];
        dwsig[rb + k] = wsig[e];
        ddel[rb + k] = del8[e];
        dcpos[rb + k] = cpos[e];
      }
    }
  }

  int rc = SNC_OK;
  if (!rc) rc = dupload(&g.row_ptr, gv->row_ptr, (size_t)n + 1);
  if (!rc) rc = dupload(&g.post, dpost.data(), (size_t)S);
  if (!rc) rc = dupload(&g.wsig, dwsig.data(), (size_t)S);
  if (!rc) rc = dupload(&g.delay, ddel.data(), (size_t)S);
  if (!rc) rc = dupload(&g.cpos, dcpos.data(), (size_t)S);
  if (!rc) rc = dupload(&g.ccol, ccol.data(), (size_t)n + 1);
  if (!rc) rc = dupload(&g.cval, cval.data(), (size_t)P);
  if (!rc) rc = dupload(&g.cdel, cdel.data(), (size_t)P);
  if (rc) return rc;

  g.role.assign(gv->role, gv->role + n);
  g.channel.assign(gv->channel, gv->channel + n);
  int n_in_ch = 0;
  for (int i = 0; i < n; ++i) {
    if (g.role[i] == SNC_ROLE_OUTPUT) g.n_out_channels = std::max(g.n_out_channels, g.channel[i] + 1);
    if (g.role[i] == SNC_ROLE_INPUT) n_in_ch = std::max(n_in_ch, g.channel[i] + 1);
  }
  g.chan_to_inputs.assign(n_in_ch, {});
  for (int i = 0; i < n; ++i)
    if (g.role[i] == SNC_ROLE_INPUT && g.channel[i] >= 0) g.chan_to_inputs[g.channel[i]].push_back(i);

  snc_graph_profile& pf = g.profile;
  pf.num_neurons = n;
  pf.num_synapses = S;
  pf.max_delay = md;
  pf.ring_len = g.L;
  pf.max_in_degree = n ? *std::max_element(g.in_degree.begin(), g.in_degree.end()) : 0;
  pf.max_out_degree = n ? *std::max_element(g.out_degree.begin(), g.out_degree.end()) : 0;
  pf.mean_in_degree = n ? (double)S / n : 0.0;
  pf.heavy_fraction = S ? (double)heavy_syn / (double)S : 0.0;
  pf.heavy_threshold = g.heavy_threshold;
  pf.num_heavy = g.n_heavy;
  *out = H.release();
  return SNC_OK;
}

void snc_hip_graph_free(snc_hip_graph* g) { delete g; }

int snc_hip_graph_profile_get(const snc_hip_graph* g, snc_graph_profile* out) {
  if (!g || !out) return fail(SNC_E_INVALID, "null argument");
  *out = g->g.profile;
  return SNC_OK;
}

// Builds a plan for B samples. With B = 1 every event belongs to the one
// sample whatever its sample_id (the reference's forward reads one sample's
// events); with B > 1 an event belongs to sample sample_id and events of
// other ids are ignored. Batched plans run the fused step kernels only.
static int plan_build(snc_hip_graph* G, const snc_event* events, int64_t n_events, int B,
                      int32_t num_steps, const snc_lif* lif, const snc_plan_config* cfg,
                      snc_hip_plan** out) {
  hip_clear();
  if (!G || !lif || !out || !cfg || (n_events > 0 && !events) || num_steps < 0 || B < 1)
    return fail(SNC_E_INVALID, "bad plan arguments");
  *out = nullptr;
  if (lif->reset_mode != SNC_RESET_TO_VALUE && lif->reset_mode != SNC_RESET_SUBTRACT)
    return fail(SNC_E_INVALID, "unknown reset mode");
  const int variant = cfg->variant;
  if (variant != SNC_VARIANT_ATOMIC && variant != SNC_VARIANT_BUCKET &&
      variant != SNC_VARIANT_SORT && variant != SNC_VARIANT_PULL)
    return fail(SNC_E_INVALID, "unknown variant");
  const bool pull = variant == SNC_VARIANT_PULL;
  if (pull) {
    if (!pull_block_ok(cfg->block_neurons))
      return fail(SNC_E_INVALID, "hip-pull blocks hold 64 to 256 neurons, a multiple of 16");
    if (cfg->heavy_threshold < 1 || cfg->heavy_threshold > kPullMaxCap)
      return fail(SNC_E_INVALID, "hip-pull heavy_threshold (light in-degree cap) must be in [1, 1024]");
    if (cfg->persistent) return fail(SNC_E_UNSUPPORTED, "hip-pull has no persistent kernel");
    if (lif->refractory > 255)
      return fail(SNC_E_UNSUPPORTED, "hip-pull keeps refractory counters in one byte (at most 255 steps)");
    if (B > kPullMaxSamples) return fail(SNC_E_UNSUPPORTED, "hip-pull plans hold at most 64 samples");
  } else {
    if (!fused_block_ok(cfg->block_neurons))
      return fail(SNC_E_INVALID, "block_neurons must be 128, 256 or 512");
    if (cfg->heavy_threshold < 1 || cfg->heavy_threshold > fused_light_max_in_degree())
      return fail(SNC_E_INVALID, "heavy_threshold must be in [1, 512]");
  }
  if (B > 1 && (variant == SNC_VARIANT_SORT || cfg->persistent))
    return fail(SNC_E_UNSUPPORTED, "batched plans run hip-atomic, hip-bucket or hip-pull");
  const DevGraph& g = G->g;
  std::unique_ptr<snc_hip_plan> P(new snc_hip_plan());
  AllocAccount acct(&P->dev_bytes);
  P->G = G;
  P->T = num_steps;
  P->B = B;
  P->lif = *lif;
  P->cfg = *cfg;
  P->variant = variant;
  P->core.variant = variant;
  const int n = g.n;

  // Hub set of this plan's threshold.
  std::vector<uint8_t> is_heavy_h((size_t)n, 0);
  for (int j = 0; j < n; ++j)
    if (g.in_degree[j] > cfg->heavy_threshold) {
      P->heavy_host.push_back(j);
      is_heavy_h[j] = 1;
    }
  const int n_heavy = (int)P->heavy_host.size();

  // Events of each sample, in input order.
  std::vector<int64_t> ev_ptr((size_t)B + 1, 0), ev_idx;
  if (B == 1) {
    ev_ptr[1] = n_events;
  } else {
    for (int64_t k = 0; k < n_events; ++k)
      if (events[k].sample_id >= 0 && events[k].sample_id < B) ev_ptr[events[k].sample_id + 1]++;
    for (int b = 0; b < B; ++b) ev_ptr[b + 1] += ev_ptr[b];
    ev_idx.resize((size_t)ev_ptr[B]);
    std::vector<int64_t> fill(ev_ptr.begin(), ev_ptr.end() - 1);
    for (int64_t k = 0; k < n_events; ++k)
      if (events[k].sample_id >= 0 && events[k].sample_id < B) ev_idx[fill[events[k].sample_id]++] = k;
  }

  // Per sample: injections resolved channel -> INPUT neurons in the
  // reference's order (cpu_backend.cpp:50-63) and folded per (step, neuron)
  // in that order; then regrouped per (step, block of block_neurons neurons)
  // for the fused kernels, hub targets (hip-bucket) into a dense
  // [step][hub] table.
  const int n_in_ch = (int)g.chan_to_inputs.size();
  const int nb = cfg->block_neurons;
  const int R = std::max(1, (n + nb - 1) / nb);
  const bool hubs = (variant == SNC_VARIANT_BUCKET || pull) && n_heavy > 0;
  std::vector<int32_t> hub_index(hubs ? n : 0, -1);
  for (int k = 0; hubs && k < n_heavy; ++k) hub_index[P->heavy_host[k]] = k;
  std::vector<int32_t> blk_ptr((size_t)B * num_steps * (R + 1), 0), bnid;
  std::vector<float> bamp, hinj(hubs ? (size_t)B * num_steps * n_heavy : 1, 0.0f);
  std::vector<int32_t> ip((size_t)num_steps + 1, 0), inid;  // sample 0 (hip-sort, persistent)
  std::vector<float> iamp;
  for (int smp = 0; smp < B; ++smp) {
    std::vector<std::vector<std::pair<int32_t, float>>> by_step(num_steps);
    for (int64_t q = ev_ptr[smp]; q < ev_ptr[smp + 1]; ++q) {
      const snc_event& e = events[B == 1 ? q : ev_idx[q]];
      if (e.time < 0 || e.time >= num_steps) continue;
      if (e.channel_id < 0 || e.channel_id >= n_in_ch) continue;
      for (int nid : g.chan_to_inputs[e.channel_id]) by_step[e.time].push_back({nid, e.amplitude});
    }
    for (int t = 0; t < num_steps; ++t) {
      auto& lst = by_step[t];
      std::stable_sort(lst.begin(), lst.end(),
                       [](const auto& x, const auto& y) { return x.first < y.first; });
      int32_t* bp = &blk_ptr[((size_t)smp * num_steps + t) * (R + 1)];
      int blk = 0;
      bp[0] = (int32_t)bnid.size();
      int64_t cnt = 0;
      for (size_t k = 0; k < lst.size();) {
        float acc = 0.0f;
        size_t q = k;
        for (; q < lst.size() && lst[q].first == lst[k].first; ++q) acc += lst[q].second;
        const int j = lst[k].first;
        k = q;
        ++cnt;
        if (smp == 0) {
          inid.push_back(j);
          iamp.push_back(acc);
        }
        if (hubs && hub_index[j] >= 0) {
          hinj[((size_t)smp * num_steps + t) * n_heavy + hub_index[j]] = acc;
          continue;
        }
        while (blk < j / nb) bp[++blk] = (int32_t)bnid.size();
        bnid.push_back(j);
        bamp.push_back(acc);
      }
      while (blk < R) bp[++blk] = (int32_t)bnid.size();
      if (smp == 0) ip[t + 1] = (int32_t)inid.size();
      P->core.max_inj_per_step = std::max<int64_t>(P->core.max_inj_per_step, cnt);
    }
    if (bnid.size() > (size_t)0x7fffffff) return fail(SNC_E_UNSUPPORTED, "too many injections");
  }
  P->h_inj_ptr = ip;

  // State layout. hip-atomic and hip-sort: ring [B][L][n]. hip-bucket with
  // the fused kernels: one arrival bitmap per ring slot, [B][L][P/32], and no
  // ring. Persistent: ring plus bitmap and dirty flags double-buffered by
  // step parity.
  P->persistent = B == 1 && !pull && persist_choose(g, variant, cfg->persistent != 0, P->pcfg);
  P->cfg.persistent = P->persistent ? 1 : 0;
  const bool bucket = variant == SNC_VARIANT_BUCKET;
  const bool slots = bucket && !P->persistent;
  const int halves = P->persistent ? 2 : 1;
  const size_t dirty_stride = ((size_t)n + 3) / 4 * 4;
  P->n_bits_words = (size_t)(g.P / 32);
  P->dirty_stride = dirty_stride;
  const size_t planes = slots ? (size_t)g.L : (bucket ? 2 : 0);
  const size_t dirty_planes = bucket && !slots ? 2 : 0;
  P->bits_total = (size_t)B * planes * P->n_bits_words;
  P->dirty_total = (size_t)B * dirty_planes * dirty_stride;
  P->ring_total = slots || pull ? 0 : (size_t)B * g.L * n;
  int rc = SNC_OK;
  if (!rc) rc = dupload(&P->heavy_ids, P->heavy_host.data(), P->heavy_host.size());
  if (!rc) rc = dupload(&P->is_heavy, is_heavy_h.data(), is_heavy_h.size());
  if (!rc) rc = dupload(&P->inj_blk_ptr, blk_ptr.data(), blk_ptr.size());
  if (!rc) rc = dupload(&P->blk_nid, bnid.data(), bnid.size());
  if (!rc) rc = dupload(&P->blk_amp, bamp.data(), bamp.size());
  if (!rc) rc = dupload(&P->inj_heavy, hinj.data(), hinj.size());
  if (!rc) rc = dupload(&P->inj_ptr, ip.data(), ip.size());
  if (!rc) rc = dupload(&P->inj_nid, inid.data(), inid.size());
  if (!rc) rc = dupload(&P->inj_amp, iamp.data(), iamp.size());
  if (!rc) rc = dalloc(&P->v, (size_t)B * n);
  if (!rc) rc = dalloc(&P->ext, (size_t)n * halves);
  if (!rc && P->ring_total) rc = dalloc(&P->ring, P->ring_total);
  if (!rc) rc = dalloc(&P->refr, (size_t)B * n);
  if (!rc) rc = dalloc(&P->fired, (size_t)n);
  if (!rc) rc = dalloc(&P->fcnt, 2);
  if (!rc) rc = dalloc(&P->nspk, (size_t)B * n);
  if (!rc && P->dirty_total) rc = dalloc(&P->dirty, P->dirty_total);  // cleared in 4-byte words
  if (!rc && P->bits_total) rc = dalloc(&P->bits, P->bits_total);
  P->core.lif_grid = std::max(1, (n + 255) / 256);
  P->core.expand_grid = std::max(1, std::min(2048, (n + 255) / 256));
  P->core.n_heavy = bucket ? n_heavy : 0;
  P->n_blk_slots = (size_t)std::max(4096, R + n_heavy + 1);
  if (!rc) rc = dalloc(&P->blk_spikes, (size_t)B * P->n_blk_slots);
  if (!rc) rc = dalloc(&P->blk_events, (size_t)B * P->n_blk_slots);
  const bool fused = !P->persistent && variant != SNC_VARIANT_SORT && !pull;
  if (!rc && fused && num_steps > 0)
    rc = dalloc(&P->inj_tab, (size_t)B * num_steps * R * kInjInline);
  if (!rc && variant == SNC_VARIANT_SORT) rc = sort_alloc(P.get(), ip, inid);
  if (!rc && variant == SNC_VARIANT_SORT && (size_t)P->n_seg > P->n_blk_slots) {
    // stats slots: one per segment workgroup
    dfree(P->blk_spikes);
    dfree(P->blk_events);
    P->n_blk_slots = (size_t)P->n_seg;
    rc = dalloc(&P->blk_spikes, (size_t)B * P->n_blk_slots);
    if (!rc) rc = dalloc(&P->blk_events, (size_t)B * P->n_blk_slots);
  }
  if (!rc && P->persistent) {
    const std::vector<int32_t> hp = persist_heavy_ptr(g, P->heavy_host, P->pcfg);
    rc = dupload(&P->heavy_blk_ptr, hp.data(), hp.size());
    P->pcfg.heavy_blk_ptr = P->heavy_blk_ptr;
    P->pcfg.bits_words = (int64_t)P->n_bits_words;
    P->pcfg.dirty_stride = (int64_t)dirty_stride;
  }
  P->halves = halves;
  if (rc) return rc;
  if (hipStreamCreateWithFlags(&P->core.stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&P->ev0) != hipSuccess || hipEventCreate(&P->ev1) != hipSuccess)
    return fail(SNC_E_HIP, "stream/event creation failed");

  StepArgs& a = P->a;
  a.n = n;
  a.L = g.L;
  a.T = num_steps;
  a.words = (n + 31) / 32;
  a.lif = {lif->threshold, lif->decay, lif->reset, lif->refractory,
           lif->reset_mode == SNC_RESET_SUBTRACT ? 1 : 0};
  a.row_ptr = g.row_ptr;
  a.post = g.post;
  a.wsig = g.wsig;
  a.delay = g.delay;
  a.cpos = g.cpos;
  a.ccol = g.ccol;
  a.cval = g.cval;
  a.cdel = g.cdel;
  a.heavy_ids = P->heavy_ids;
  a.inj_ptr = P->inj_ptr;
  a.inj_nid = P->inj_nid;
  a.inj_amp = P->inj_amp;
  a.v = P->v;
  a.refr = P->refr;
  a.ext = P->ext;
  a.ring = P->ring;
  a.dirty = P->dirty;
  a.bits = P->bits;
  a.fired = P->fired;
  a.fcnt = P->fcnt;
  a.nspk = P->nspk;
  a.blk_spikes = P->blk_spikes;
  a.blk_events = P->blk_events;
  a.raster = nullptr;
  a.sseg = P->sseg;
  a.sseg_e = P->sseg_e;
  a.slist = P->slist;
  a.scnt = P->scnt;
  a.slist_stride = g.P;
  a.sinj_ptr = P->sinj_ptr;
  a.n_seg = P->n_seg;
  a.sort_slot0 = 0;
  a.fbits = P->fbits;
  a.inj_blk_ptr = P->inj_blk_ptr;
  a.blk_nid = P->blk_nid;
  a.blk_amp = P->blk_amp;
  a.inj_heavy = P->inj_heavy;
  a.inj_tab = P->inj_tab;
  a.is_heavy = P->is_heavy;
  a.n_reg_blocks = R;
  a.blk = nb;
  a.n_heavy = hubs ? n_heavy : 0;
  a.bits_half = (int64_t)P->n_bits_words;
  a.dirty_half = (int64_t)dirty_stride;
  a.bits_slot = (int64_t)P->n_bits_words;
  a.B = B;
  a.ring_stride = (int64_t)g.L * n;
  a.bits_stride = (int64_t)(planes * P->n_bits_words);
  a.dirty_stride = (int64_t)(dirty_planes * dirty_stride);
  a.blk_stride = (int64_t)P->n_blk_slots;
  a.inj_stride = (int64_t)num_steps * (R + 1);
  a.hinj_stride = (int64_t)num_steps * n_heavy;
  a.tab_stride = (int64_t)num_steps * R * kInjInline;
  a.raster_stride = (int64_t)num_steps * a.words;
  if (pull) {
    // The light block's LDS injection pool holds the largest (step, block).
    int64_t pool = 64;
    for (int t = 0; t < num_steps; ++t)
      for (int q = 0; q < R; ++q) {
        int64_t tot = 0;
        for (int smp = 0; smp < B; ++smp) {
          const int32_t* bp = &blk_ptr[((size_t)smp * num_steps + t) * (R + 1) + q];
          tot += bp[1] - bp[0];
        }
        pool = std::max(pool, tot);
      }
    pool = (pool + 63) / 64 * 64;
    if (pool > (1 << 20)) return fail(SNC_E_UNSUPPORTED, "hip-pull: injection pool exceeds LDS");
    // Halo-indexed columns per light cap, built once per graph and kept
    // while it lives (plans of different caps may be alive together).
    PullGraph* pgp = nullptr;
    for (PullGraph* q : G->pulls)
      if (q->cap == cfg->heavy_threshold && q->blk == nb) pgp = q;
    if (!pgp) {
      std::unique_ptr<PullGraph> fresh(new PullGraph());
      AllocAccount gacct(&fresh->dev_bytes);
      const int brc = pull_prepare(g, cfg->heavy_threshold, nb, *fresh);
      if (brc) return brc;
      pgp = fresh.release();
      G->pulls.push_back(pgp);
    }
    PullGraph& pg = *pgp;
    if (pg.hub_host != P->heavy_host) return fail(SNC_E_UNSUPPORTED, "hip-pull hub sets differ");
    if (pg.n_blocks != R) return fail(SNC_E_INVALID, "hip-pull block count differs from the plan's");
    int lrc = pull_set_lds(B, pg.halo_words, (int)pool, nb);
    if (lrc) return lrc;
    const size_t word_u32 = (size_t)pull_word_bytes(B) / 4;
    // the step kernels index every array with 32-bit byte offsets
    if ((uint64_t)g.P * 4 >= (1ull << 32) || (uint64_t)n * B * 4 >= (1ull << 32) ||
        (uint64_t)g.L * n * word_u32 * 4 >= (1ull << 32))
      return fail(SNC_E_UNSUPPORTED, "hip-pull: graph or batch too large for 32-bit offsets");
    int prc = dalloc(&P->F, (size_t)g.L * n * word_u32);
    // hub lists: a row of lstride values per hub entry, plus the rows (and
    // one row's lanes) that loads past a chunk's list may touch
    const int lstride = (B + 7) / 8 * 8;
    if ((uint64_t)(pg.hub_entries + 16) * lstride * 4 >= (1ull << 32))
      return fail(SNC_E_UNSUPPORTED, "hip-pull: hub lists too large for 32-bit offsets");
    if (!prc) prc = dalloc(&P->hub_list, (size_t)(pg.hub_entries + 16) * lstride + kPullMaxSamples);
    if (!prc)
      prc = dalloc(&P->hub_cnt, (size_t)std::max(1, pg.n_chunks) * lstride + kPullMaxSamples);
    const int sum_words = (n + 15) / 16;
    // an even count: the fill kernel clears whole 32-bit words
    P->fsum_total = ((size_t)g.L * sum_words + 1) / 2 * 2;
    if (!prc) prc = dalloc(&P->fsum, P->fsum_total);
    if (prc) return prc;
    P->F_words = (size_t)g.L * n * word_u32;
    PullArgs& pa = P->pa;
    pa.n = n;
    pa.L = g.L;
    pa.B = B;
    pa.T = num_steps;
    pa.words = a.words;
    pa.n_blocks = R;
    pa.blk = nb;
    pa.pool = (int)pool;
    pa.halo = pg.halo_words;
    pa.direct = pg.direct ? 1 : 0;
    pa.lif = a.lif;
    pa.lw = pg.lw;
    pa.lh = pg.lh;
    pa.ldir = pg.ldir;
    pa.loff = pg.loff;
    pa.lcnt = pg.lcnt;
    pa.hub_of = pg.hub_of;
    pa.fsum = P->fsum;
    pa.sum_words = sum_words;
    pa.halo_ptr = pg.halo_ptr;
    pa.halo_pk = pg.halo;
    pa.hub_ids = pg.hub_ids;
    pa.hub_pk = pg.hub_pk;
    pa.hub_w = pg.hub_w;
    pa.chunk_hub = pg.chunk_hub;
    pa.chunk_pos = pg.chunk_pos;
    pa.chunk_len = pg.chunk_len;
    pa.hub_chunk_ptr = pg.hub_chunk_ptr;
    pa.blk_order = pg.blk_order;
    pa.n_hub = pg.n_hub;
    pa.n_chunks = pg.n_chunks;
    pa.inj_blk_ptr = P->inj_blk_ptr;
    pa.blk_nid = P->blk_nid;
    pa.blk_amp = P->blk_amp;
    pa.inj_hub = P->inj_heavy;
    pa.inj_stride = a.inj_stride;
    pa.hinj_stride = a.hinj_stride;
    pa.F = P->F;
    pa.hub_list = P->hub_list;
    pa.hub_cnt = P->hub_cnt;
    pa.lstride = lstride;
    pa.v = P->v;  // hip-pull state is sample-minor: [n][B]
    pa.refr = reinterpret_cast<uint8_t*>(P->refr);  // one byte per neuron and sample
    pa.nspk = P->nspk;
    pa.raster = nullptr;
    pa.raster_stride = a.raster_stride;
  }
  if (P->inj_tab) {
    rc = fwd_launch_inj_table(P->core.stream, a);
    if (!rc) {
      const hipError_t es = hipStreamSynchronize(P->core.stream);
      if (es != hipSuccess) rc = hip_fail(es, "injection table build");
    }
    if (rc) return rc;
  }

  *out = P.release();
  return SNC_OK;
}

int snc_hip_plan_create_config(snc_hip_graph* G, const snc_event* events, int64_t n_events,
                               int32_t num_steps, const snc_lif* lif,
                               const snc_plan_config* cfg, snc_hip_plan** out) {
  return plan_build(G, events, n_events, 1, num_steps, lif, cfg, out);
}

int snc_hip_plan_get_config(const snc_hip_plan* p, snc_plan_config* out) {
  if (!p || !out) return fail(SNC_E_INVALID, "null argument");
  *out = p->cfg;
  return SNC_OK;
}

int snc_hip_graph_tune_report(const snc_hip_graph* g, int32_t exact_only, int32_t batched,
                              snc_tune_report* out) {
  if (!g || !out) return fail(SNC_E_INVALID, "null argument");
  const int k = (exact_only ? 1 : 0) + (batched ? 2 : 0);
  if (!g->tuned[k]) return fail(SNC_E_INVALID, "no autotuner search has run for this graph");
  *out = g->tune[k];
  return SNC_OK;
}

int snc_hip_plan_create_batch(snc_hip_graph* G, const snc_event* events, int64_t n_events,
                              int32_t num_samples, int32_t num_steps, const snc_lif* lif,
                              int32_t variant, const snc_plan_config* cfg, snc_hip_plan** out) {
  if (!G || !lif || !out || (n_events > 0 && !events) || num_steps < 0 || num_samples < 1)
    return fail(SNC_E_INVALID, "bad plan arguments");
  if (cfg) return plan_build(G, events, n_events, num_samples, num_steps, lif, cfg, out);
  snc_plan_config c;
  if (variant == SNC_VARIANT_AUTO || variant == SNC_VARIANT_AUTO_EXACT) {
    const int rc = autotune_search(G, events, n_events, num_samples, num_steps, lif,
                                   variant == SNC_VARIANT_AUTO_EXACT, &c);
    if (rc) return rc;
  } else if (variant == SNC_VARIANT_ATOMIC || variant == SNC_VARIANT_BUCKET ||
             variant == SNC_VARIANT_PULL || (variant == SNC_VARIANT_SORT && num_samples == 1)) {
    c = default_plan_config(G->g, variant);
  } else {
    return fail(SNC_E_INVALID,
                "unknown variant (batched plans: hip-atomic, hip-bucket, hip-pull, auto)");
  }
  return plan_build(G, events, n_events, num_samples, num_steps, lif, &c, out);
}

int snc_hip_plan_create(snc_hip_graph* G, const snc_event* events, int64_t n_events,
                        int32_t num_steps, const snc_lif* lif, int32_t variant,
                        snc_hip_plan** out) {
  return snc_hip_plan_create_batch(G, events, n_events, 1, num_steps, lif, variant, nullptr, out);
}

int32_t snc_hip_plan_variant(const snc_hip_plan* p) { return p ? p->variant : -1; }

void snc_hip_plan_free(snc_hip_plan* p) { delete p; }

}  // extern "C"

namespace {

// Clear all per-run state with fill kernels (no memset nodes), then stage the
// first step's injections.
int reset_state(snc_hip_plan* P, bool raster) {
  const DevGraph& g = P->G->g;
  hipStream_t st = P->core.stream;
  const size_t n = (size_t)g.n, B = (size_t)P->B;
  int rc = SNC_OK;
  auto fill = [&](void* ptr, size_t bytes) {
    if (!rc && ptr && bytes) rc = fwd_launch_fill(st, ptr, 0u, (bytes + 3) / 4);
  };
  fill(P->v, sizeof(float) * B * n);
  fill(P->ext, sizeof(float) * n * P->halves);
  fill(P->ring, sizeof(float) * P->ring_total);
  fill(P->refr, sizeof(int32_t) * B * n);
  fill(P->fcnt, sizeof(int32_t) * 2);
  if (P->scnt) {  // hip-sort: each counter starts at its segment's list start
    const hipError_t ec = hipMemcpyAsync(P->scnt, P->scnt_init.data(),
                                         sizeof(int32_t) * P->scnt_init.size(),
                                         hipMemcpyHostToDevice, st);
    if (!rc && ec != hipSuccess) rc = hip_fail(ec, "segment counter reset");
  }
  fill(P->nspk, sizeof(int32_t) * B * n);
  fill(P->dirty, P->dirty_total);
  fill(P->bits, sizeof(uint32_t) * P->bits_total);
  if (P->F) fill(P->F, sizeof(uint32_t) * P->F_words);
  if (P->fsum) fill(P->fsum, sizeof(uint16_t) * P->fsum_total);
  fill(P->blk_spikes, sizeof(unsigned long long) * B * P->n_blk_slots);
  fill(P->blk_events, sizeof(unsigned long long) * B * P->n_blk_slots);
  if (raster) fill(P->raster, sizeof(uint32_t) * B * (size_t)P->T * P->a.words);
  if (!rc && P->persistent) rc = fwd_launch_first_injection(P->core, P->a);
  return rc;
}

int ensure_raster(snc_hip_plan* P) {
  if (P->raster) return SNC_OK;
  AllocAccount acct(&P->dev_bytes);
  return dalloc(&P->raster, (size_t)P->B * std::max(1, P->T) * P->a.words);
}

// The step chain: one fused launch per step (hip-atomic, hip-bucket) or the
// hip-sort kernel chain of each step, all on the plan's stream.
int launch_chain(snc_hip_plan* P, const StepArgs& a) {
  for (int t = 0; t < P->T; ++t) {
    int rc;
    if (P->variant == SNC_VARIANT_SORT) {
      rc = sort_launch_step(P->core, a, t);
    } else if (P->variant == SNC_VARIANT_PULL) {
      PullArgs pa = P->pa;
      pa.raster = a.raster;
      rc = pull_launch_step(pa, t, P->core.stream);
    } else {
      rc = fused_launch_step(P->core, a, t);
    }
    if (rc) return rc;
  }
  return SNC_OK;
}

bool use_graphs() {
  const char* e = std::getenv("SNC_HIP_GRAPH");
  return !(e && e[0] == '0');
}

int launch_loop(snc_hip_plan* P, bool raster) {
  StepArgs a = P->a;
  a.raster = raster ? P->raster : nullptr;
  if (P->persistent) return persist_launch(P->G->g, P->core, a, P->pcfg);
  if (!use_graphs() || P->T == 0) return launch_chain(P, a);
  hipGraphExec_t& ge = P->graph_exec[raster ? 1 : 0];
  if (!ge) {
    hipGraph_t graph = nullptr;
    const hipError_t eb = hipStreamBeginCapture(P->core.stream, hipStreamCaptureModeThreadLocal);
    if (eb != hipSuccess) return hip_fail(eb, "hipStreamBeginCapture");
    int rc = launch_chain(P, a);
    hipError_t e = hipStreamEndCapture(P->core.stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess || !graph) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      ge = nullptr;
      return hip_fail(e, "hipGraphInstantiate");
    }
  }
  const hipError_t el = hipGraphLaunch(ge, P->core.stream);
  return el == hipSuccess ? SNC_OK : hip_fail(el, "hipGraphLaunch");
}

}  // namespace

extern "C" {

int snc_hip_plan_launch(snc_hip_plan* P) {
  if (!P) return fail(SNC_E_INVALID, "null plan");
  hip_clear();
  int rc = reset_state(P, false);
  if (!rc) rc = launch_loop(P, false);
  return rc;
}

int snc_hip_plan_run(snc_hip_plan* P, snc_forward_capture* cap, snc_forward_stats* stats) {
  if (!P || !stats) return fail(SNC_E_INVALID, "null argument");
  hip_clear();
  const DevGraph& g = P->G->g;
  const bool want_raster = cap && (cap->raster || cap->step_spikes);
  int rc = SNC_OK;
  if (want_raster) rc = ensure_raster(P);
  if (!rc) rc = reset_state(P, want_raster);
  if (rc) return rc;
  hipStream_t st = P->core.stream;
  if (hipEventRecord(P->ev0, st) != hipSuccess) return fail(SNC_E_HIP, "hipEventRecord failed");
  rc = launch_loop(P, want_raster);
  if (rc) return rc;
  if (hipEventRecord(P->ev1, st) != hipSuccess) return fail(SNC_E_HIP, "hipEventRecord failed");
  const hipError_t es = hipStreamSynchronize(st);
  if (es != hipSuccess) return hip_fail(es, "step loop");
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, P->ev0, P->ev1);
  P->last_ms = ms;

  const size_t n = (size_t)g.n, B = (size_t)P->B, slots = P->n_blk_slots;
  std::vector<unsigned long long> bs(B * slots), be(B * slots);
  std::vector<int32_t> nspk(B * n);
  const bool pull = P->variant == SNC_VARIANT_PULL;
  if (hipMemcpy(nspk.data(), P->nspk, sizeof(int32_t) * B * n, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(SNC_E_HIP, "result readback failed");
  std::memset(stats, 0, sizeof(*stats));
  P->last_spikes.assign(B, 0);
  P->last_events.assign(B, 0);
  if (pull) {
    // hip-pull keeps no per-step counters and its state is sample-minor:
    // transpose, then count spikes and events per sample from the per-neuron
    // counts (each spike of neuron i carries its out-degree's events).
    std::vector<int32_t> t(B * n);
    for (size_t i = 0; i < n; ++i)
      for (size_t b = 0; b < B; ++b) t[b * n + i] = nspk[i * B + b];
    nspk.swap(t);
    for (size_t b = 0; b < B; ++b)
      for (size_t i = 0; i < n; ++i) {
        P->last_spikes[b] += nspk[b * n + i];
        P->last_events[b] += (int64_t)nspk[b * n + i] * g.out_degree[i];
      }
  } else {
    if (hipMemcpy(bs.data(), P->blk_spikes, bs.size() * sizeof(bs[0]), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(be.data(), P->blk_events, be.size() * sizeof(be[0]), hipMemcpyDeviceToHost) != hipSuccess)
      return fail(SNC_E_HIP, "result readback failed");
    for (size_t b = 0; b < B; ++b)
      for (size_t k = 0; k < slots; ++k) {
        P->last_spikes[b] += (int64_t)bs[b * slots + k];
        P->last_events[b] += (int64_t)be[b * slots + k];
      }
  }
  for (size_t b = 0; b < B; ++b) {
    stats->spikes += P->last_spikes[b];
    stats->synaptic_events += P->last_events[b];
  }
  stats->used_variant = P->variant;
  stats->ms = ms;
  if (cap) {
    if (cap->neuron_spikes) std::copy(nspk.begin(), nspk.end(), cap->neuron_spikes);
    if (cap->logits) {
      // Output spikes summed per channel in float, as the CPU decode does
      // (cpu_backend.cpp:149-151); exact while counts stay below 2^24.
      const size_t C = (size_t)g.n_out_channels;
      std::fill(cap->logits, cap->logits + B * C, 0.0f);
      for (size_t b = 0; b < B; ++b)
        for (size_t i = 0; i < n; ++i)
          if (g.role[i] == SNC_ROLE_OUTPUT && g.channel[i] >= 0)
            cap->logits[b * C + g.channel[i]] += (float)nspk[b * n + i];
    }
    if (cap->final_v) {
      if (pull) {
        std::vector<float> vv(B * n);
        if (hipMemcpy(vv.data(), P->v, sizeof(float) * B * n, hipMemcpyDeviceToHost) != hipSuccess)
          return fail(SNC_E_HIP, "voltage readback failed");
        for (size_t i = 0; i < n; ++i)
          for (size_t b = 0; b < B; ++b) cap->final_v[b * n + i] = vv[i * B + b];
      } else if (hipMemcpy(cap->final_v, P->v, sizeof(float) * B * n, hipMemcpyDeviceToHost) != hipSuccess) {
        return fail(SNC_E_HIP, "voltage readback failed");
      }
    }
    if (want_raster) {
      const size_t W = (size_t)P->a.words, T = (size_t)P->T;
      std::vector<uint32_t> r(B * T * W);
      if (!r.empty() && hipMemcpy(r.data(), P->raster, r.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(SNC_E_HIP, "raster readback failed");
      if (cap->raster) std::copy(r.begin(), r.end(), cap->raster);
      if (cap->step_spikes)
        for (size_t bt = 0; bt < B * T; ++bt) {
          int64_t c = 0;
          for (size_t w = 0; w < W; ++w) c += __builtin_popcount(r[bt * W + w]);
          cap->step_spikes[bt] = c;
        }
    }
  }
  return SNC_OK;
}

int32_t snc_hip_plan_num_samples(const snc_hip_plan* p) { return p ? p->B : -1; }

int snc_hip_graph_memory(const snc_hip_graph* g, int64_t* bytes) {
  if (!g || !bytes) return fail(SNC_E_INVALID, "null argument");
  int64_t b = g->g.dev_bytes;
  for (const PullGraph* q : g->pulls) b += q->dev_bytes;
  *bytes = b;
  return SNC_OK;
}

int snc_hip_plan_memory(const snc_hip_plan* p, int64_t* bytes) {
  if (!p || !bytes) return fail(SNC_E_INVALID, "null argument");
  *bytes = p->dev_bytes;
  return SNC_OK;
}

int snc_hip_plan_sample_stats(const snc_hip_plan* p, int64_t* spikes, int64_t* events) {
  if (!p || !spikes || !events) return fail(SNC_E_INVALID, "null argument");
  if (p->last_spikes.size() != (size_t)p->B) return fail(SNC_E_INVALID, "the plan has not run");
  std::copy(p->last_spikes.begin(), p->last_spikes.end(), spikes);
  std::copy(p->last_events.begin(), p->last_events.end(), events);
  return SNC_OK;
}

int snc_hip_plan_kernel_time(const snc_hip_plan* p, double* avg_ms, int64_t* launches) {
  if (!p || !avg_ms || !launches) return fail(SNC_E_INVALID, "null argument");
  if (p->persistent) {
    *avg_ms = p->last_ms;
    *launches = 1;
  } else {
    *avg_ms = p->T > 0 ? p->last_ms / p->T : 0.0;
    *launches = p->T;
  }
  return SNC_OK;
}

int snc_hip_synchronize(void) {
  return hipDeviceSynchronize() == hipSuccess ? SNC_OK : fail(SNC_E_HIP, "hipDeviceSynchronize failed");
}

int snc_hip_plan_profile(snc_hip_plan* P, double* kernel_ms, int64_t* launches) {
  if (!P || !kernel_ms || !launches) return fail(SNC_E_INVALID, "null argument");
  for (int k = 0; k < SNC_KERNEL_CLASSES; ++k) {
    kernel_ms[k] = 0.0;
    launches[k] = 0;
  }
  int rc = reset_state(P, false);
  if (rc) return rc;
  hipStream_t st = P->core.stream;
  hipEvent_t ev[4];
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(SNC_E_HIP, "hipEventCreate failed");
  StepArgs a = P->a;
  a.raster = nullptr;
  auto span = [&](int k, hipEvent_t b, hipEvent_t e) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, b, e) == hipSuccess) {
      kernel_ms[k] += ms;
      launches[k] += 1;
    }
  };
  if (P->persistent) {
    (void)hipEventRecord(ev[0], st);
    rc = persist_launch(P->G->g, P->core, a, P->pcfg);
    (void)hipEventRecord(ev[1], st);
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(SNC_E_HIP, "run failed");
    if (!rc) span(SNC_KERNEL_PERSISTENT, ev[0], ev[1]);
  } else {
    for (int t = 0; t < P->T && !rc; ++t) {
      (void)hipEventRecord(ev[0], st);
      if (P->variant == SNC_VARIANT_SORT) {
        rc = sort_launch_step(P->core, a, t);
        (void)hipEventRecord(ev[3], st);
        if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(SNC_E_HIP, "run failed");
        if (!rc) span(SNC_KERNEL_SORT_STEP, ev[0], ev[3]);
        continue;
      }
      if (P->variant == SNC_VARIANT_PULL) {
        PullArgs pa = P->pa;
        pa.raster = a.raster;
        rc = pull_launch_step(pa, t, st);
      } else {
        rc = fused_launch_step(P->core, a, t);
      }
      (void)hipEventRecord(ev[1], st);
      if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(SNC_E_HIP, "run failed");
      if (rc) break;
      span(SNC_KERNEL_STEP, ev[0], ev[1]);
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
}

int snc_hip_forward(const snc_graph_view* gv, const snc_event* events, int64_t n_events,
                    int32_t num_steps, const snc_lif* lif, int32_t variant,
                    snc_forward_capture* cap, snc_forward_stats* stats) {
  snc_hip_graph* G = nullptr;
  int rc = snc_hip_graph_upload(gv, &G);
  if (rc) return rc;
  snc_hip_plan* P = nullptr;
  rc = snc_hip_plan_create(G, events, n_events, num_steps, lif, variant, &P);
  if (!rc) rc = snc_hip_plan_run(P, cap, stats);
  snc_hip_plan_free(P);
  snc_hip_graph_free(G);
  return rc;
}

}  // extern "C"

// The autotuner's search (cached per graph, separately for single-sample
// and batched plans; a batched search is redone for another sample count,
// since the fastest strategy changes with it): every candidate runs the
// first trial_steps steps of this input twice (the first run captures and
// instantiates its graph); the device time of the second run decides.
static int autotune_search(snc_hip_graph* G, const snc_event* events, int64_t n_events,
                           int num_samples, int num_steps, const snc_lif* lif, bool exact_only,
                           snc_plan_config* out) {
  const int k = (exact_only ? 1 : 0) + (num_samples > 1 ? 2 : 0);
  if (G->tuned[k] && G->tuned_samples[k] == num_samples) {
    *out = G->tune[k].cand[G->tune[k].chosen];
    return SNC_OK;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const std::vector<snc_plan_config> cands =
      autotune_candidates(G->g, exact_only, num_samples == 1, num_samples);
  snc_tune_report rep{};
  rep.trial_steps = autotune_trial_steps(num_steps);
  int best = -1;
  double best_ms = 0.0;
  for (const snc_plan_config& c : cands) {
    snc_hip_plan* P = nullptr;
    int rc = plan_build(G, events, n_events, num_samples, rep.trial_steps, lif, &c, &P);
    snc_forward_stats st{};
    if (!rc) rc = snc_hip_plan_run(P, nullptr, &st);
    if (!rc) rc = snc_hip_plan_run(P, nullptr, &st);
    snc_hip_plan_free(P);
    if (rc) {
      (void)hipGetLastError();  // a candidate that cannot run leaves no error behind
      fprintf(stderr, "snc_hip autotuner: candidate variant %d block %d hub %d persistent %d "
                      "skipped: %s\n",
              c.variant, c.block_neurons, c.heavy_threshold, c.persistent, snc_hip_last_error());
    }
    const double ms = rc ? -1.0 : st.ms / rep.trial_steps;  // -1: could not run here
    if (rep.n_candidates >= SNC_TUNE_MAX_CANDIDATES) break;
    rep.cand[rep.n_candidates] = c;
    rep.ms_per_step[rep.n_candidates] = ms;
    rep.status[rep.n_candidates] = rc;
    if (rc)
      std::snprintf(rep.error[rep.n_candidates], SNC_TUNE_ERROR_LEN, "%s", snc_hip_last_error());
    if (ms >= 0.0 && (best < 0 || ms < best_ms)) {
      best = rep.n_candidates;
      best_ms = ms;
    }
    ++rep.n_candidates;
  }
  if (best < 0) return fail(SNC_E_HIP, "autotuner: no candidate configuration ran");
  rep.chosen = best;
  rep.search_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  G->tune[k] = rep;
  G->tuned[k] = true;
  G->tuned_samples[k] = num_samples;
  *out = rep.cand[best];
  return SNC_OK;
}
