// C ABI of the e-prop trainer session (cudatrain::*, cuda_trainer.hpp:24-57).
// Host code; kernels are in train.hip.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "common.hpp"
#include "hip_err.hpp"
#include "forward_plan.hpp"
#include "snc_hip.h"
#include "train.hpp"

namespace snc_amd {
namespace {

template <typename T>
int talloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return fail(SNC_E_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
  }
  return SNC_OK;
}

template <typename T>
int tupload(T** p, const T* h, size_t count) {
  int rc = talloc(p, count);
  if (!rc && count && hipMemcpy(*p, h, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(SNC_E_HIP, "hipMemcpy H2D failed");
  return rc;
}

template <typename T>
void tfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

}  // namespace
}  // namespace snc_amd

using namespace snc_amd;

struct snc_hip_train {
  snc_train_config cfg{};
  int n = 0, S = 0, L = 2, classes = 1, n_internal = 0, batch = 1, max_steps = 1;
  int64_t P = 0;
  bool f64 = true, ring_lds = true;
  size_t lds_max = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  TrainArgs a{};
  std::vector<int32_t> h_src, h_inputs;
  // device graph
  int32_t *row_ptr = nullptr, *post = nullptr, *src = nullptr, *cpos = nullptr, *ccol = nullptr,
          *csc_e = nullptr, *channel = nullptr, *internal_row = nullptr, *inputs = nullptr;
  uint8_t *cdel = nullptr, *sdel = nullptr, *role = nullptr;
  float *w = nullptr, *wc = nullptr, *Bfb = nullptr;
  int2* eunit = nullptr;  // eligibility work units
  int32_t *ecol_pre = nullptr, *ecol_syn = nullptr;
  int* wq = nullptr;       // eligibility work queue [2 * batch]
  float* Ls_g = nullptr;   // learning signals [batch][n]
  int n_cu = 1;
  // dataset
  int n_data = 0, dim = 0;
  std::vector<int32_t> h_labels, h_lengths;
  float* X = nullptr;
  int32_t *labels = nullptr, *lengths = nullptr, *ev_base = nullptr, *ev_step_ptr = nullptr,
          *ev_nid = nullptr;
  float* ev_amp = nullptr;
  int64_t max_inj = 0;
  // batch buffers
  int32_t* idx = nullptr;
  float* hist = nullptr;  // [batch][max_steps][2][hn] psi / trace histories
  float *C = nullptr, *ring_g = nullptr, *logits = nullptr;
  uint2* arq = nullptr;  // arrival queues [batch][L][S] (ring not in LDS)
  double* loss = nullptr;
  int32_t* correct = nullptr;
  unsigned long long *blk_spikes = nullptr, *blk_events = nullptr, *neur_fire = nullptr;
  uint32_t* raster = nullptr;
  size_t raster_words = 0;
  // plain state
  PlainState ps{};
  float *p_v = nullptr, *p_tr = nullptr, *p_psi = nullptr, *p_ext = nullptr, *p_ring = nullptr,
        *p_logit = nullptr, *p_L = nullptr;
  int32_t *p_refr = nullptr, *p_fired = nullptr, *p_nf = nullptr;
  uint32_t *p_fw = nullptr, *p_bits = nullptr;
  void* p_E = nullptr;
  unsigned long long *p_spikes = nullptr, *p_events = nullptr;
  int32_t* p_cur = nullptr;
  // reward mode: the action stream (trainer.cpp:39) and device state
  std::mt19937_64 action_rng;
  float *r_u = nullptr, *r_base = nullptr, *r_delta = nullptr, *r_L = nullptr, *r_adv = nullptr;
  int32_t* r_act = nullptr;
  // captured plain-path chains, keyed by (chain length, update on/off)
  std::map<std::pair<int, int>, hipGraphExec_t> plain_graphs;

  void drop_graphs() {
    for (auto& kv : plain_graphs) (void)hipGraphExecDestroy(kv.second);
    plain_graphs.clear();
  }

  ~snc_hip_train() {
    drop_graphs();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
    for (void* p : {(void*)row_ptr, (void*)post, (void*)src, (void*)cpos, (void*)ccol,
                    (void*)csc_e, (void*)channel, (void*)internal_row, (void*)inputs, (void*)cdel,
                    (void*)sdel,
                    (void*)role, (void*)w, (void*)wc, (void*)Bfb, (void*)eunit,
                    (void*)ecol_pre, (void*)ecol_syn, (void*)wq, (void*)Ls_g, (void*)X,
                    (void*)labels,
                    (void*)lengths, (void*)ev_base, (void*)ev_step_ptr, (void*)ev_nid,
                    (void*)ev_amp, (void*)idx, (void*)hist, (void*)C, (void*)ring_g, (void*)arq,
                    (void*)logits,
                    (void*)loss, (void*)correct, (void*)blk_spikes, (void*)blk_events,
                    (void*)neur_fire, (void*)raster, (void*)p_v, (void*)p_tr, (void*)p_psi,
                    (void*)p_ext, (void*)p_ring, (void*)p_logit, (void*)p_L, (void*)p_refr,
                    (void*)p_fired, (void*)p_nf, (void*)p_fw, (void*)p_bits, p_E,
                    (void*)p_spikes, (void*)p_events, (void*)p_cur, (void*)r_u,
                    (void*)r_base, (void*)r_delta, (void*)r_L, (void*)r_adv, (void*)r_act})
      if (p) (void)hipFree(p);
  }
};

namespace {

// Step granule of the plain path's captured chains: 8 steps, or more when
// max_steps exceeds 256, so that at most 32 chain lengths exist.
int plain_granule(int max_steps) { return std::max(8, (max_steps + 31) / 32); }

int fill0(hipStream_t st, void* p, size_t bytes) {
  return p && bytes ? fwd_launch_fill(st, p, 0u, (bytes + 3) / 4) : SNC_OK;
}

int check_labels(const snc_hip_train* s, const int32_t* labels, int32_t n) {
  for (int i = 0; i < n; ++i)
    if (labels[i] < 0 || labels[i] >= s->classes)
      return fail(SNC_E_INVALID, "label outside [0, num_output_channels)");
  return SNC_OK;
}

int check_lengths(snc_hip_train* s, const int32_t* lengths, int32_t n, std::vector<int32_t>& out) {
  out.assign(n, s->cfg.num_steps);
  if (lengths)
    for (int i = 0; i < n; ++i) out[i] = lengths[i];
  for (int i = 0; i < n; ++i)
    if (out[i] < 0 || out[i] > s->max_steps)
      return fail(SNC_E_INVALID, "sample length outside [0, max_steps]");
  return SNC_OK;
}

}  // namespace

extern "C" {

int snc_hip_train_create(const snc_graph_view* g, const snc_train_config* cfg, int32_t batch,
                         int32_t max_steps, int32_t elig, snc_hip_train** out) {
  if (!g || !cfg || !out || batch < 1 || max_steps < 1) return fail(SNC_E_INVALID, "bad arguments");
  *out = nullptr;
  if (cfg->encoder != SNC_ENC_DIRECT && cfg->encoder != SNC_ENC_POISSON &&
      cfg->encoder != SNC_ENC_LATENCY)
    return fail(SNC_E_INVALID, "unknown encoder");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(SNC_E_NODEVICE, "no device");
  const std::string err = validate_view(*g);
  if (!err.empty()) return fail(SNC_E_INVALID, "graph validation failed: " + err);
  std::unique_ptr<snc_hip_train> s(new snc_hip_train());
  s->cfg = *cfg;
  s->n = g->num_neurons;
  s->S = g->num_synapses;
  s->batch = batch;
  s->max_steps = max_steps;
  s->f64 = elig == SNC_ELIG_F64;
  const int n = s->n, S = s->S;
  int n_out = 0, md = 1;
  for (int i = 0; i < n; ++i)
    if (g->role[i] == SNC_ROLE_OUTPUT) n_out = std::max(n_out, g->channel[i] + 1);
  for (int e = 0; e < S; ++e) md = std::max(md, (int)g->delays[e]);
  if (md > 255) return fail(SNC_E_UNSUPPORTED, "delays above 255 steps are not supported");
  s->L = md + 1;
  s->classes = std::max(1, n_out);
  if (s->classes > 64) return fail(SNC_E_UNSUPPORTED, "more than 64 output channels");

  // Weights then feedback from one mt19937_64(seed) stream (trainer.cpp:15-37).
  std::mt19937_64 rng(cfg->seed);
  std::uniform_real_distribution<float> wu(0.0f, cfg->w_init);
  std::vector<float> w(S);
  for (float& x : w) x = wu(rng);
  std::vector<int32_t> internal_row(n, -1);
  for (int i = 0; i < n; ++i)
    if (g->role[i] == SNC_ROLE_INTERNAL) internal_row[i] = s->n_internal++;
  std::normal_distribution<float> bn(0.0f, cfg->feedback_scale);
  std::vector<float> Bfb((size_t)s->n_internal * s->classes);
  for (float& b : Bfb) b = bn(rng);

  // Incoming columns padded to 32 (the ordered-fold layout of the forward path).
  std::vector<int32_t> indeg(n, 0), src(S);
  for (int i = 0; i < n; ++i)
    for (int e = g->row_ptr[i]; e < g->row_ptr[i + 1]; ++e) {
      src[e] = i;
      indeg[g->post_ids[e]]++;
    }
  std::vector<int32_t> ccol((size_t)n + 1, 0);
  int64_t P = 0;
  for (int j = 0; j < n; ++j) {
    ccol[j] = (int32_t)P;
    P += ((int64_t)indeg[j] + 31) / 32 * 32;
  }
  if (P > 0x7fffff00LL) return fail(SNC_E_UNSUPPORTED, "graph too large");
  ccol[n] = (int32_t)P;
  s->P = P;
  // Column order: delay descending, then pre, then CSR position. One step's
  // arrivals at a target with one delay share one ring slot; with each
  // delay's entries contiguous, every slot is one run of the column, in the
  // (pre, CSR position) order the CPU ring adds them in.
  std::vector<int32_t> cpos(S), csc_e((size_t)P, -1), fillc(n, 0);
  std::vector<uint8_t> cdel((size_t)P, 0), sdel(S), role(g->role, g->role + n);
  {
    std::vector<int32_t> by_col(S);
    std::iota(by_col.begin(), by_col.end(), 0);
    std::stable_sort(by_col.begin(), by_col.end(), [&](int32_t x, int32_t y) {
      const int jx = g->post_ids[x], jy = g->post_ids[y];
      if (jx != jy) return jx < jy;
      return g->delays[x] > g->delays[y];
    });
    for (const int32_t e : by_col) {
      const int j = g->post_ids[e];
      const int32_t p = ccol[j] + fillc[j]++;
      cpos[e] = p;
      csc_e[p] = e;
      cdel[p] = (uint8_t)g->delays[e];
    }
  }
  for (int e = 0; e < S; ++e) sdel[e] = (uint8_t)g->delays[e];
  // Eligibility work units: each post's synapses in column order, cut into
  // runs of at most kEligPerThread; a lane of the update phase takes one run.
  if (n >= (1 << 26)) return fail(SNC_E_UNSUPPORTED, "too many neurons for eligibility units");
  std::vector<int2> eunit;
  std::vector<int32_t> ecol_pre, ecol_syn;
  ecol_pre.reserve(S);
  ecol_syn.reserve(S);
  // Within a post the synapses are dealt to the unit slots so that slot k of
  // unit i holds a pre with pre % 8 == (i + k) % 8 where one is left: the
  // eight lanes one LDS pass serves then read trace rows in different bank
  // groups. Each synapse's sum is unchanged; only which lane computes it.
  std::vector<std::vector<int32_t>> by_cls(8);
  for (int j = 0; j < n; ++j) {
    const int q0 = (int)ecol_pre.size();
    for (auto& v : by_cls) v.clear();
    int cnt = 0;
    for (int64_t p = ccol[j + 1] - 1; p >= ccol[j]; --p)  // popped back to front below
      if (csc_e[p] >= 0) {
        by_cls[src[csc_e[p]] & 7].push_back(csc_e[p]);
        ++cnt;
      }
    const int units = (cnt + kEligPerThread - 1) / kEligPerThread;
    std::vector<int32_t> slot((size_t)units * kEligPerThread, -1);
    auto take = [&](int cls) {
      for (int d = 0; d < 8; ++d) {
        auto& v = by_cls[(cls + d) & 7];
        if (!v.empty()) {
          const int32_t e = v.back();
          v.pop_back();
          return e;
        }
      }
      return (int32_t)-1;
    };
    int placed = 0;
    for (int k = 0; k < kEligPerThread && placed < cnt; ++k)
      for (int i = 0; i < units && placed < cnt; ++i) {
        // the last unit is the short one: it only takes slots below its length
        const int len_i = i + 1 < units ? kEligPerThread : cnt - (units - 1) * kEligPerThread;
        if (k >= len_i) continue;
        slot[(size_t)i * kEligPerThread + k] = take((i + k) & 7);
        ++placed;
      }
    for (int i = 0; i < units; ++i) {
      const int len_i = i + 1 < units ? kEligPerThread : cnt - (units - 1) * kEligPerThread;
      eunit.push_back(make_int2((j << 5) | (len_i - 1), (int)ecol_pre.size()));
      for (int k = 0; k < len_i; ++k) {
        const int32_t e = slot[(size_t)i * kEligPerThread + k];
        ecol_pre.push_back(src[e]);
        ecol_syn.push_back(e);
      }
    }
    (void)q0;
  }
  if ((int64_t)ecol_syn.size() != S) return fail(SNC_E_HIP, "eligibility units miss synapses");
  for (int i = 0; i < n; ++i)
    if (g->role[i] == SNC_ROLE_INPUT) s->h_inputs.push_back(i);
  s->h_src = src;

  int rc = SNC_OK;
  if (!rc) rc = tupload(&s->row_ptr, g->row_ptr, (size_t)n + 1);
  if (!rc) rc = tupload(&s->post, g->post_ids, (size_t)S);
  if (!rc) rc = tupload(&s->src, src.data(), (size_t)S);
  if (!rc) rc = tupload(&s->cpos, cpos.data(), (size_t)S);
  if (!rc) rc = tupload(&s->ccol, ccol.data(), ccol.size());
  if (!rc) rc = tupload(&s->csc_e, csc_e.data(), csc_e.size());
  if (!rc) rc = tupload(&s->cdel, cdel.data(), cdel.size());
  if (!rc) rc = tupload(&s->sdel, sdel.data(), sdel.size());
  if (!rc) rc = tupload(&s->role, role.data(), role.size());
  if (!rc) rc = tupload(&s->channel, g->channel, (size_t)n);
  if (!rc) rc = tupload(&s->internal_row, internal_row.data(), internal_row.size());
  if (!rc) rc = tupload(&s->inputs, s->h_inputs.data(), s->h_inputs.size());
  if (!rc) rc = tupload(&s->w, w.data(), w.size());
  if (!rc) rc = tupload(&s->eunit, eunit.data(), eunit.size());
  if (!rc) rc = tupload(&s->ecol_pre, ecol_pre.data(), ecol_pre.size());
  if (!rc) rc = tupload(&s->ecol_syn, ecol_syn.data(), ecol_syn.size());
  if (!rc) rc = talloc(&s->wq, 2 * (size_t)batch);
  if (!rc) rc = talloc(&s->Ls_g, (size_t)batch * n);
  if (!rc) rc = talloc(&s->wc, (size_t)P);
  if (!rc) rc = tupload(&s->Bfb, Bfb.data(), Bfb.size());
  const size_t esz = s->f64 ? sizeof(double) : sizeof(float);
  const int hn = (n + 3) / 4 * 4;
  if (!rc) rc = talloc(&s->hist, (size_t)batch * max_steps * 2 * hn);
  if (!rc) rc = talloc(&s->C, (size_t)batch * S);
  if (!rc) rc = talloc(&s->idx, (size_t)batch);
  if (!rc) rc = talloc(&s->logits, (size_t)batch * s->classes);
  if (!rc) rc = talloc(&s->loss, (size_t)batch);
  if (!rc) rc = talloc(&s->correct, (size_t)batch);
  if (!rc) rc = talloc(&s->blk_spikes, (size_t)batch);
  if (!rc) rc = talloc(&s->blk_events, (size_t)batch);
  if (!rc) rc = talloc(&s->neur_fire, (size_t)n);
  // plain-path state
  if (!rc) rc = talloc(&s->p_v, (size_t)n);
  if (!rc) rc = talloc(&s->p_tr, (size_t)n);
  if (!rc) rc = talloc(&s->p_psi, (size_t)n);
  if (!rc) rc = talloc(&s->p_ext, (size_t)n);
  if (!rc) rc = talloc(&s->p_ring, (size_t)s->L * n);
  if (!rc) rc = talloc(&s->p_logit, (size_t)s->classes);
  if (!rc) rc = talloc(&s->p_L, (size_t)n);
  if (!rc) rc = talloc(&s->p_refr, (size_t)n);
  if (!rc) rc = talloc(&s->p_fired, (size_t)n);
  if (!rc) rc = talloc(&s->p_nf, 1);
  if (!rc) rc = talloc(&s->p_fw, (size_t)(n + 31) / 32);
  if (!rc) rc = talloc(&s->p_bits, (size_t)(P / 32));
  if (!rc && hipMalloc(&s->p_E, esz * (size_t)std::max(1, S)) != hipSuccess)
    rc = fail(SNC_E_NOMEM, "eligibility allocation failed");
  if (!rc) rc = talloc(&s->p_spikes, 1);
  if (!rc) rc = talloc(&s->p_events, 1);
  if (!rc) rc = talloc(&s->p_cur, 4);
  if (!rc && cfg->reward_mode) {
    // The baseline starts at chance, 1 / classes in float (trainer.cpp:40).
    const float base0 = 1.0f / (float)s->classes;
    rc = talloc(&s->r_u, (size_t)batch);
    if (!rc) rc = tupload(&s->r_base, &base0, 1);
    if (!rc) rc = talloc(&s->r_delta, (size_t)batch * s->classes);
    if (!rc) rc = talloc(&s->r_L, (size_t)batch * n);
    if (!rc) rc = talloc(&s->r_adv, (size_t)batch);
    if (!rc) rc = talloc(&s->r_act, (size_t)batch);
    s->action_rng.seed(cfg->seed ^ 0xA5A5C0FFEEULL);
  }
  if (rc) return rc;
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
    return fail(SNC_E_HIP, "stream/event creation failed");
  rc = train_launch_gather(s->w, s->csc_e, s->wc, P, s->stream);
  if (!rc) rc = fill0(s->stream, s->neur_fire, sizeof(unsigned long long) * n);
  if (rc) return rc;

  TrainArgs& a = s->a;
  a.n = n;
  a.S = S;
  a.L = s->L;
  a.classes = s->classes;
  a.words = (n + 31) / 32;
  a.bits_words = (int)(P / 32);
  a.n_inputs = (int)s->h_inputs.size();
  a.readout_leaky = cfg->readout_mode == SNC_READOUT_LEAKY;
  a.readout_psi = cfg->readout_psi;
  a.lif = {cfg->lif.threshold, cfg->lif.decay, cfg->lif.reset, cfg->lif.refractory,
           cfg->lif.reset_mode == SNC_RESET_SUBTRACT ? 1 : 0};
  a.gamma = cfg->surrogate_scale;
  a.lr = cfg->lr;
  a.w_max = cfg->w_max;
  a.train_hidden = cfg->train_hidden;
  a.reward = cfg->reward_mode ? 1 : 0;
  a.encoder = cfg->encoder;
  a.gain = cfg->enc_gain;
  a.max_rate = cfg->enc_max_rate;
  a.enc_thr = cfg->enc_threshold;
  a.seed_lo = (uint32_t)cfg->seed;
  a.seed_hi = (uint32_t)(cfg->seed >> 32);
  a.T_max = max_steps;
  a.row_ptr = s->row_ptr;
  a.post = s->post;
  a.src = s->src;
  a.cpos = s->cpos;
  a.ccol = s->ccol;
  a.cdel = s->cdel;
  a.wc = s->wc;
  a.w_csr = s->w;
  a.sdel = s->sdel;
  a.role = s->role;
  a.channel = s->channel;
  a.internal_row = s->internal_row;
  a.Bfb = s->Bfb;
  a.inputs = s->inputs;
  a.idx = s->idx;
  a.hist = s->hist;
  a.hn = hn;
  a.eunit = s->eunit;
  a.n_units = (int)eunit.size();
  a.ecol_pre = s->ecol_pre;
  a.ecol_syn = s->ecol_syn;
  a.rounds = (a.n_units + kTrainThreads - 1) / kTrainThreads;
  a.wq = s->wq;
  a.Ls_g = s->Ls_g;
  a.logits = s->logits;
  a.loss = s->loss;
  a.correct = s->correct;
  a.blk_spikes = s->blk_spikes;
  a.blk_events = s->blk_events;

  // The delay ring goes to LDS when the whole per-sample state fits.
  hipDeviceProp_t prop;
  (void)hipGetDeviceProperties(&prop, 0);
  const size_t lds_max = prop.sharedMemPerBlock;
  s->n_cu = std::max(1, prop.multiProcessorCount);
  // A workgroup polls for unpublished eligibility work for at most 200 ms.
  int dev = 0, khz = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
    return fail(SNC_E_HIP, "wall clock rate unavailable");
  a.wait_ticks = (long long)khz * 200;
  s->lds_max = lds_max;
  // Leave room for the kernel's few bytes of static LDS.
  const size_t lds_budget = lds_max > 256 ? lds_max - 256 : 0;
  if (std::max(train_lds_bytes(a, false), train_elig_lds_bytes(n, 4)) > lds_budget)
    return fail(SNC_E_UNSUPPORTED, "per-sample state does not fit in LDS");
  // SNC_HIP_TRAIN_DELAY selects where the batched kernel keeps the deposits
  // in flight: "queues" (arrival queues) or "hbm" (the ring in HBM) even when
  // the ring would fit in LDS; unset or anything else, the ring in LDS when
  // it fits, else the queues, else the ring in HBM.
  const char* dm = std::getenv("SNC_HIP_TRAIN_DELAY");
  const bool want_queues = dm && std::strcmp(dm, "queues") == 0;
  const bool want_hbm = dm && std::strcmp(dm, "hbm") == 0;
  s->ring_lds = !want_queues && !want_hbm && train_lds_bytes(a, true) <= lds_budget;
  if (!s->ring_lds) {
    // Arrival queues (each sample's deposits, per ring slot, sorted when
    // their step comes) up to kArqBytesMax; beyond, the ring in HBM.
    const size_t arq_bytes = (size_t)batch * s->L * (size_t)std::max<int64_t>(1, s->S) * sizeof(uint2);
    a.arq_cap = std::max<int64_t>(1, s->S);
    a.arq = nullptr;
    if (!want_hbm && arq_bytes <= kArqBytesMax) {
      a.arq = reinterpret_cast<uint2*>(1);  // sizes the LDS of the queue mode below
      if (std::max(train_lds_bytes(a, false), train_elig_lds_bytes(n, 4)) <= lds_budget) {
        rc = talloc(&s->arq, (size_t)batch * s->L * (size_t)a.arq_cap);
        if (rc) return rc;
      }
      a.arq = s->arq;
    }
    if (!a.arq) {
      rc = talloc(&s->ring_g, (size_t)batch * s->L * n);
      if (rc) return rc;
    }
  }
  a.ring_g = s->ring_g;

  s->ps = {s->p_v, s->p_tr, s->p_psi, s->p_ext, s->p_ring, s->p_logit, s->p_L, s->p_refr,
           s->p_fired, s->p_nf, s->p_cur, s->p_fw, s->p_bits, s->p_E, s->p_spikes,
           s->p_events};
  if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(SNC_E_HIP, "init failed");
  *out = s.release();
  return SNC_OK;
}

void snc_hip_train_destroy(snc_hip_train* s) { delete s; }

int snc_hip_train_set_weights(snc_hip_train* s, const float* w, int64_t n) {
  if (!s || !w) return fail(SNC_E_INVALID, "null argument");
  if (n != s->S) return fail(SNC_E_INVALID, "weight count != num_synapses");
  if (hipMemcpy(s->w, w, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess)
    return fail(SNC_E_HIP, "weight upload failed");
  int rc = train_launch_gather(s->w, s->csc_e, s->wc, s->P, s->stream);
  if (!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(SNC_E_HIP, "gather failed");
  return rc;
}

int snc_hip_train_get_weights(const snc_hip_train* s, float* w, int64_t n) {
  if (!s || !w) return fail(SNC_E_INVALID, "null argument");
  if (n != s->S) return fail(SNC_E_INVALID, "weight count != num_synapses");
  if (hipStreamSynchronize(s->stream) != hipSuccess ||
      hipMemcpy(w, s->w, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(SNC_E_HIP, "weight readback failed");
  return SNC_OK;
}

int snc_hip_train_upload_features(snc_hip_train* s, const float* X, int32_t n, int32_t dim,
                                  const int32_t* labels, const int32_t* lengths) {
  if (!s || !X || !labels || n < 1 || dim < 1) return fail(SNC_E_INVALID, "bad dataset");
  std::vector<int32_t> len;
  int rc = check_lengths(s, lengths, n, len);
  if (!rc) rc = check_labels(s, labels, n);
  if (rc) return rc;
  tfree(s->X);
  tfree(s->labels);
  tfree(s->lengths);
  tfree(s->ev_base);
  tfree(s->ev_step_ptr);
  tfree(s->ev_nid);
  tfree(s->ev_amp);
  if (!rc) rc = tupload(&s->X, X, (size_t)n * dim);
  if (!rc) rc = tupload(&s->labels, labels, (size_t)n);
  if (!rc) rc = tupload(&s->lengths, len.data(), len.size());
  if (rc) return rc;
  s->n_data = n;
  s->dim = dim;
  s->drop_graphs();  // captured chains hold the old dataset pointers
  s->h_labels.assign(labels, labels + n);
  s->h_lengths = len;
  s->a.X = s->X;
  s->a.dim = dim;
  s->a.labels = s->labels;
  s->a.lengths = s->lengths;
  s->a.ev_base = nullptr;
  s->a.max_inj_per_step = 0;
  return SNC_OK;
}

int snc_hip_train_upload_events(snc_hip_train* s, const snc_event* ev, int64_t n_events,
                                int32_t n, const int32_t* labels, const int32_t* lengths) {
  if (!s || (n_events > 0 && !ev) || !labels || n < 1) return fail(SNC_E_INVALID, "bad dataset");
  std::vector<int32_t> len;
  int rc = check_lengths(s, lengths, n, len);
  if (!rc) rc = check_labels(s, labels, n);
  if (rc) return rc;
  // channel -> INPUT neurons (cpu_backend.cpp:50-55) and per-(sample, step)
  // injections folded per neuron in event order.
  std::vector<int32_t> role_ch(s->n);
  std::vector<uint8_t> role(s->n);
  (void)hipMemcpy(role.data(), s->role, s->n, hipMemcpyDeviceToHost);
  (void)hipMemcpy(role_ch.data(), s->channel, sizeof(int32_t) * s->n, hipMemcpyDeviceToHost);
  int n_in_ch = 0;
  for (int i = 0; i < s->n; ++i)
    if (role[i] == SNC_ROLE_INPUT) n_in_ch = std::max(n_in_ch, role_ch[i] + 1);
  std::vector<std::vector<int32_t>> by_ch(n_in_ch);
  for (int i = 0; i < s->n; ++i)
    if (role[i] == SNC_ROLE_INPUT && role_ch[i] >= 0) by_ch[role_ch[i]].push_back(i);
  std::vector<int32_t> base(n + 1, 0);
  for (int i = 0; i < n; ++i) base[i + 1] = base[i] + len[i] + 1;
  std::vector<std::vector<std::pair<int32_t, float>>> lists((size_t)base[n]);
  for (int64_t k = 0; k < n_events; ++k) {
    const snc_event& e = ev[k];
    if (e.sample_id < 0 || e.sample_id >= n) continue;
    if (e.time < 0 || e.time >= len[e.sample_id]) continue;
    if (e.channel_id < 0 || e.channel_id >= n_in_ch) continue;
    for (int nid : by_ch[e.channel_id])
      lists[(size_t)base[e.sample_id] + e.time].push_back({nid, e.amplitude});
  }
  std::vector<int32_t> step_ptr((size_t)base[n] + 1, 0), nid;
  std::vector<float> amp;
  int64_t mx = 0;
  for (size_t q = 0; q < lists.size(); ++q) {
    auto& l = lists[q];
    std::stable_sort(l.begin(), l.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    for (size_t k = 0; k < l.size();) {
      float acc = 0.0f;
      size_t r = k;
      for (; r < l.size() && l[r].first == l[k].first; ++r) acc += l[r].second;
      nid.push_back(l[k].first);
      amp.push_back(acc);
      k = r;
    }
    step_ptr[q + 1] = (int32_t)nid.size();
    mx = std::max<int64_t>(mx, step_ptr[q + 1] - step_ptr[q]);
  }
  s->drop_graphs();  // captured chains hold the old dataset pointers
  tfree(s->X);
  tfree(s->labels);
  tfree(s->lengths);
  tfree(s->ev_base);
  tfree(s->ev_step_ptr);
  tfree(s->ev_nid);
  tfree(s->ev_amp);
  if (!rc) rc = tupload(&s->labels, labels, (size_t)n);
  if (!rc) rc = tupload(&s->lengths, len.data(), len.size());
  if (!rc) rc = tupload(&s->ev_base, base.data(), (size_t)n);
  if (!rc) rc = tupload(&s->ev_step_ptr, step_ptr.data(), step_ptr.size());
  if (!rc) rc = tupload(&s->ev_nid, nid.data(), nid.size());
  if (!rc) rc = tupload(&s->ev_amp, amp.data(), amp.size());
  if (rc) return rc;
  s->n_data = n;
  s->dim = 0;
  s->h_labels.assign(labels, labels + n);
  s->h_lengths = len;
  s->a.X = nullptr;
  s->a.dim = 0;
  s->a.labels = s->labels;
  s->a.lengths = s->lengths;
  s->a.ev_base = s->ev_base;
  s->a.ev_step_ptr = s->ev_step_ptr;
  s->a.ev_nid = s->ev_nid;
  s->a.ev_amp = s->ev_amp;
  s->a.max_inj_per_step = mx;
  return SNC_OK;
}

int snc_hip_train_step(snc_hip_train* s, const int32_t* idx, int32_t B, int32_t epoch,
                       int32_t mode, int32_t apply, float* logits_out, uint32_t* raster_out,
                       snc_train_stats* st) {
  if (!s || !idx || B < 1 || B > s->batch) return fail(SNC_E_INVALID, "bad minibatch");
  hip_clear();
  if (!s->n_data) return fail(SNC_E_INVALID, "no dataset uploaded");
  for (int b = 0; b < B; ++b)
    if (idx[b] < 0 || idx[b] >= s->n_data) return fail(SNC_E_INVALID, "sample index out of range");
  hipStream_t sm = s->stream;
  int rc = SNC_OK;
  if (hipMemcpyAsync(s->idx, idx, sizeof(int32_t) * B, hipMemcpyHostToDevice, sm) != hipSuccess)
    return fail(SNC_E_HIP, "index upload failed");
  const bool reward = apply && s->cfg.reward_mode;
  if (reward) {
    // One action draw per trained sample, in minibatch order (trainer.cpp:171-173).
    std::vector<float> u(B);
    std::uniform_real_distribution<float> ud(0.0f, 1.0f);
    for (float& x : u) x = ud(s->action_rng);
    if (hipMemcpyAsync(s->r_u, u.data(), sizeof(float) * B, hipMemcpyHostToDevice, sm) !=
            hipSuccess ||
        hipStreamSynchronize(sm) != hipSuccess)
      return fail(SNC_E_HIP, "action draw upload failed");
  }
  TrainArgs a = s->a;
  a.epoch = epoch;
  a.C = apply ? s->C : nullptr;
  a.neur_fire = apply ? s->neur_fire : nullptr;
  a.raster = nullptr;
  if (raster_out) {
    const size_t words = (size_t)B * s->max_steps * ((s->n + 31) / 32);
    if (words > s->raster_words) {
      tfree(s->raster);
      rc = talloc(&s->raster, words);
      if (rc) return rc;
      s->raster_words = words;
    }
    rc = fill0(sm, s->raster, words * 4);
    if (rc) return rc;
    a.raster = s->raster;
  }
  (void)hipEventRecord(s->ev0, sm);
  if (mode == SNC_TRAIN_BATCHED) {
    if (a.C) rc = fill0(sm, s->wq, sizeof(int) * 2 * (size_t)B);
    if (!rc)
      rc = train_launch_batched(a, B, train_batched_grid(a, B, s->n_cu), s->f64, s->ring_lds,
                                s->lds_max, sm);
  } else if (mode == SNC_TRAIN_SERIAL) {
    // One launch of the batched kernel per sample, in minibatch order. Each
    // launch sees sample b's rows of every per-minibatch array as a
    // minibatch of one, so the contributions land in row b as the batched
    // launch writes them, and the update below sums the same rows in the
    // same order.
    if (a.C) rc = fill0(sm, s->wq, sizeof(int) * 2 * (size_t)B);
    for (int b = 0; b < B && !rc; ++b) {
      TrainArgs ab = a;
      ab.idx = a.idx + b;
      ab.hist = a.hist + (size_t)b * a.T_max * 2 * a.hn;
      if (a.C) ab.C = a.C + (size_t)b * a.S;
      if (a.ring_g) ab.ring_g = a.ring_g + (size_t)b * a.L * a.n;
      if (a.arq) ab.arq = a.arq + (size_t)b * a.L * a.arq_cap;
      ab.logits = a.logits + (size_t)b * a.classes;
      ab.loss = a.loss + b;
      ab.correct = a.correct + b;
      ab.blk_spikes = a.blk_spikes + b;
      ab.blk_events = a.blk_events + b;
      if (a.raster) ab.raster = a.raster + (size_t)b * a.T_max * a.words;
      ab.wq = a.wq + 2 * (size_t)b;
      ab.Ls_g = a.Ls_g + (size_t)b * a.n;
      rc = train_launch_batched(ab, 1, train_batched_grid(ab, 1, s->n_cu), s->f64, s->ring_lds,
                                s->lds_max, sm);
    }
  } else if (mode == SNC_TRAIN_PLAIN) {
    const size_t esz = s->f64 ? sizeof(double) : sizeof(float);
    rc = fill0(sm, s->p_spikes, 8);
    if (!rc) rc = fill0(sm, s->p_events, 8);
    // One sample's chain: clear its state (fill kernels), then T steps of
    // six kernels and the end-of-sample kernels.
    auto chain = [&](int T) {
      int r = fill0(sm, s->p_v, 4 * (size_t)s->n);
      if (!r) r = fill0(sm, s->p_tr, 4 * (size_t)s->n);
      if (!r) r = fill0(sm, s->p_ext, 4 * (size_t)s->n);
      if (!r) r = fill0(sm, s->p_refr, 4 * (size_t)s->n);
      if (!r) r = fill0(sm, s->p_ring, 4 * (size_t)s->L * s->n);
      if (!r) r = fill0(sm, s->p_logit, 4 * (size_t)s->classes);
      if (!r) r = fill0(sm, s->p_bits, 4 * (size_t)(s->P / 32));
      if (!r) r = fill0(sm, s->p_nf, 4);
      if (!r) r = fill0(sm, s->p_E, esz * (size_t)s->S);
      if (!r) r = train_launch_plain_sample(a, s->ps, T, s->f64, sm);
      return r;
    };
    // The chain is captured into a hipGraph per (chain length, update) and
    // replayed; rasters (a per-call buffer) and SNC_HIP_GRAPH=0 run eagerly.
    // A sample of length T replays the chain of length T rounded up to a
    // multiple of plain_granule() (its step kernels past T return at once),
    // so at most 2 * ceil(max_steps / granule) <= 66 graphs are ever live.
    const char* ge = std::getenv("SNC_HIP_GRAPH");
    const bool use_graph = !a.raster && !(ge && ge[0] == '0');
    const int gran = plain_granule(s->max_steps);
    for (int b = 0; b < B && !rc; ++b) {
      const int sample = idx[b];
      const int T = s->h_lengths[sample];
      rc = train_launch_plain_begin(s->ps, b, sample, epoch, T, sm);
      if (rc) break;
      if (!use_graph) {
        rc = chain(T);
        continue;
      }
      const int Tc = std::min(s->max_steps, (T + gran - 1) / gran * gran);
      const auto key = std::make_pair(Tc, apply ? 1 : 0);
      auto it = s->plain_graphs.find(key);
      if (it == s->plain_graphs.end()) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal) != hipSuccess)
          return fail(SNC_E_HIP, "hipStreamBeginCapture failed");
        const int crc = chain(Tc);
        const hipError_t e = hipStreamEndCapture(sm, &graph);
        if (crc || e != hipSuccess || !graph) {
          if (graph) (void)hipGraphDestroy(graph);
          return crc ? crc : fail(SNC_E_HIP, "hipStreamEndCapture failed");
        }
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return fail(SNC_E_HIP, "hipGraphInstantiate failed");
        it = s->plain_graphs.emplace(key, exec).first;
      }
      if (hipGraphLaunch(it->second, sm) != hipSuccess) rc = fail(SNC_E_HIP, "hipGraphLaunch failed");
    }
  } else {
    return fail(SNC_E_INVALID, "unknown training mode");
  }
  if (!rc && reward)
    rc = train_launch_reward(a, {s->r_u, s->r_base, s->r_delta, s->r_L, s->r_act, s->r_adv,
                                 s->cfg.reward_baseline_decay},
                             B, sm);
  if (!rc && apply) rc = train_launch_apply(a, s->w, B, s->csc_e, s->wc, s->P, sm);
  if (rc) return rc;
  (void)hipEventRecord(s->ev1, sm);
  if (hipStreamSynchronize(sm) != hipSuccess) return fail(SNC_E_HIP, "training step failed");
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, s->ev0, s->ev1);

  std::vector<double> loss(B);
  std::vector<int32_t> corr(B);
  std::vector<unsigned long long> sp(B), evn(B);
  if (hipMemcpy(loss.data(), s->loss, 8 * B, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(corr.data(), s->correct, 4 * B, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(SNC_E_HIP, "result readback failed");
  if (logits_out &&
      hipMemcpy(logits_out, s->logits, sizeof(float) * B * s->classes, hipMemcpyDeviceToHost) !=
          hipSuccess)
    return fail(SNC_E_HIP, "logit readback failed");
  if (raster_out &&
      hipMemcpy(raster_out, s->raster, sizeof(uint32_t) * B * s->max_steps * ((s->n + 31) / 32),
                hipMemcpyDeviceToHost) != hipSuccess)
    return fail(SNC_E_HIP, "raster readback failed");
  if (st) {
    std::memset(st, 0, sizeof(*st));
    if (mode != SNC_TRAIN_PLAIN) {
      (void)hipMemcpy(sp.data(), s->blk_spikes, 8 * B, hipMemcpyDeviceToHost);
      (void)hipMemcpy(evn.data(), s->blk_events, 8 * B, hipMemcpyDeviceToHost);
      for (int b = 0; b < B; ++b) {
        st->spikes += (int64_t)sp[b];
        st->synaptic_events += (int64_t)evn[b];
      }
    } else {
      unsigned long long a1 = 0, a2 = 0;
      (void)hipMemcpy(&a1, s->p_spikes, 8, hipMemcpyDeviceToHost);
      (void)hipMemcpy(&a2, s->p_events, 8, hipMemcpyDeviceToHost);
      st->spikes = (int64_t)a1;
      st->synaptic_events = (int64_t)a2;
    }
    for (int b = 0; b < B; ++b) {
      st->loss_sum += loss[b];
      st->correct += corr[b];
    }
    st->samples = B;
    st->ms = ms;
  }
  return SNC_OK;
}

int snc_hip_train_epoch(snc_hip_train* s, const int32_t* idx, int32_t n, int32_t epoch,
                        int32_t mode, snc_train_stats* st) {
  if (!s || !s->n_data) return fail(SNC_E_INVALID, "no dataset uploaded");
  if (idx && n < 0) return fail(SNC_E_INVALID, "negative sample count");
  hip_clear();
  std::vector<int32_t> order;
  if (idx) {
    order.assign(idx, idx + n);
  } else {
    order.resize(s->n_data);
    std::iota(order.begin(), order.end(), 0);
  }
  std::mt19937_64 rng(0x9e3779b97f4a7c15ULL * (uint64_t)(epoch + 1));
  std::shuffle(order.begin(), order.end(), rng);
  const int total = (int)order.size();
  snc_train_stats tot{};
  for (int off = 0; off < total; off += s->batch) {
    const int B = std::min(s->batch, total - off);
    snc_train_stats one{};
    int rc = snc_hip_train_step(s, order.data() + off, B, epoch, mode, 1, nullptr, nullptr, &one);
    if (rc) return rc;
    tot.spikes += one.spikes;
    tot.synaptic_events += one.synaptic_events;
    tot.loss_sum += one.loss_sum;
    tot.correct += one.correct;
    tot.samples += one.samples;
    tot.ms += one.ms;
  }
  if (st) *st = tot;
  return SNC_OK;
}

int snc_hip_train_evaluate(snc_hip_train* s, const int32_t* idx, int32_t n, int32_t mode,
                           snc_train_stats* st) {
  if (!s || !idx || n < 0) return fail(SNC_E_INVALID, "bad arguments");
  hip_clear();
  snc_train_stats tot{};
  for (int off = 0; off < n; off += s->batch) {
    const int B = std::min(s->batch, n - off);
    snc_train_stats one{};
    int rc = snc_hip_train_step(s, idx + off, B, 0, mode, 0, nullptr, nullptr, &one);
    if (rc) return rc;
    tot.spikes += one.spikes;
    tot.synaptic_events += one.synaptic_events;
    tot.loss_sum += one.loss_sum;
    tot.correct += one.correct;
    tot.samples += one.samples;
    tot.ms += one.ms;
  }
  if (st) *st = tot;
  return SNC_OK;
}

int snc_hip_train_profile(snc_hip_train* s, const int32_t* idx, int32_t B, int32_t epoch,
                          double* mean_us, double* max_us, double* kernel_ms) {
  if (!s || !idx || B < 1 || B > s->batch || !mean_us || !max_us)
    return fail(SNC_E_INVALID, "bad arguments");
  static_assert(kTrainPhases == SNC_TRAIN_PHASES, "phase count mismatch");
  hip_clear();
  if (!s->n_data) return fail(SNC_E_INVALID, "no dataset uploaded");
  for (int b = 0; b < B; ++b)
    if (idx[b] < 0 || idx[b] >= s->n_data) return fail(SNC_E_INVALID, "sample index out of range");
  int dev = 0, khz = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
    return fail(SNC_E_HIP, "wall clock rate unavailable");
  TrainArgs a = s->a;
  a.epoch = epoch;
  a.C = s->C;  // contributions are computed but not applied
  a.neur_fire = nullptr;
  a.raster = nullptr;
  const int G = train_batched_grid(a, B, s->n_cu);
  unsigned long long* prof = nullptr;
  int rc = talloc(&prof, (size_t)G * kTrainPhases);
  if (rc) return rc;
  a.prof = prof;
  hipStream_t sm = s->stream;
  if (hipMemcpyAsync(s->idx, idx, sizeof(int32_t) * B, hipMemcpyHostToDevice, sm) != hipSuccess)
    rc = fail(SNC_E_HIP, "index upload failed");
  if (!rc) rc = fill0(sm, prof, sizeof(unsigned long long) * G * kTrainPhases);
  if (!rc) rc = fill0(sm, s->wq, sizeof(int) * 2 * (size_t)B);
  if (!rc) (void)hipEventRecord(s->ev0, sm);
  if (!rc) rc = train_launch_batched(a, B, G, s->f64, s->ring_lds, s->lds_max, sm);
  if (!rc) (void)hipEventRecord(s->ev1, sm);
  std::vector<unsigned long long> h((size_t)G * kTrainPhases);
  if (!rc && (hipStreamSynchronize(sm) != hipSuccess ||
              hipMemcpy(h.data(), prof, sizeof(unsigned long long) * h.size(),
                        hipMemcpyDeviceToHost) != hipSuccess))
    rc = fail(SNC_E_HIP, "profiled minibatch failed");
  tfree(prof);
  if (rc) return rc;
  const double us_per_tick = 1e3 / (double)khz;
  for (int k = 0; k < kTrainPhases; ++k) {
    double sum = 0.0, mx = 0.0;
    for (int b = 0; b < G; ++b) {
      const double us = (double)h[(size_t)b * kTrainPhases + k] * us_per_tick;
      sum += us;
      mx = std::max(mx, us);
    }
    mean_us[k] = sum / B;  // per sample: eligibility rounds may run on any workgroup
    max_us[k] = mx;
  }
  if (kernel_ms) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, s->ev0, s->ev1);
    *kernel_ms = ms;
  }
  return SNC_OK;
}

int snc_hip_train_reset_stats(snc_hip_train* s) {
  if (!s) return fail(SNC_E_INVALID, "null session");
  int rc = fill0(s->stream, s->neur_fire, sizeof(unsigned long long) * s->n);
  if (!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(SNC_E_HIP, "reset failed");
  return rc;
}

// Per-synapse deliveries equal the fires of the synapse's pre neuron: in a
// frozen graph every spike is delivered once on every outgoing synapse
// (trainer.cpp:142-150), so they are derived from the per-neuron counts.
int snc_hip_train_collect_stats(const snc_hip_train* s, int64_t* syn, int64_t* neur) {
  if (!s) return fail(SNC_E_INVALID, "null session");
  std::vector<unsigned long long> nf(s->n);
  if (hipStreamSynchronize(s->stream) != hipSuccess ||
      hipMemcpy(nf.data(), s->neur_fire, 8 * (size_t)s->n, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(SNC_E_HIP, "stats readback failed");
  if (neur)
    for (int i = 0; i < s->n; ++i) neur[i] = (int64_t)nf[i];
  if (syn)
    for (int e = 0; e < s->S; ++e) syn[e] = (int64_t)nf[s->h_src[e]];
  return SNC_OK;
}

}  // extern "C"
