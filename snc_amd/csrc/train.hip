// e-prop training on gfx950: the batched persistent path and the plain
// per-sample path, sharing one definition of the math.
//
// Per sample and step (CPU Trainer order, src/training/trainer.cpp:91-151):
//   inject -> LIF + surrogate psi -> trace -> eligibility E_s += psi_post tr_pre
//   -> delivery of fired rows with the trainer's signed weights.
// Delivery is the ordered bucket fold of the forward path: fired rows set
// presence bits over each target's incoming column (delay descending, pre,
// CSR position), then every target adds its bits in column order, so rasters
// are the CPU's bit for bit. The batched kernel also stages each arrival's
// weight and delay in LDS at its rank among the step's set bits (one
// popcount scan of the bitmap), so the fold reads LDS instead of gathering
// column entries from HBM. Eligibility is accumulated per (sample, synapse)
// in double (the CPU's exact sums) or float.
//
// After a sample's T_b steps: delta = softmax(logits) - onehot(label)
// (trainer.cpp:158-180), learning signal L (outputs: delta, hidden: B delta,
// trainer.cpp:182-195) and the per-synapse contribution
// C_{b,s} = (lr / T_b * L_post) * float(E_{b,s}). A separate kernel sums C over
// the minibatch in sample order and applies w <- clamp(w - sum / B), the
// minibatch-mean rule of the reference GPU trainer (cuda_trainer.cu:163-172).
// In reward mode (trainer.cpp:165-179) a sample's delta depends on the
// baseline left by the samples before it, so the sample kernels store
// float(E_{b,s}) in C and three small kernels finish the minibatch: the
// ordered action/advantage/baseline scan, the learning signals, and the
// contributions.
//
// Batched (SNC_TRAIN_BATCHED): one workgroup per sample runs all of its T_b
// steps in a single launch -- neuron state, fired bitmap, delivery buckets and
// (when it fits) the delay ring live in LDS; uneven lengths need no masking.
// Plain (SNC_TRAIN_PLAIN): the same phases as grid-wide kernels, one sample
// at a time, six launches per step -- the shape of the reference's trainer.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "common.hpp"
#include "hip_err.hpp"
#include "fold.hpp"
#include "forward_plan.hpp"
#include "rng.hpp"
#include "train.hpp"
#include "wave.hpp"

namespace snc_amd {


// ---------------------------------------------------------------------------
// shared device pieces
// ---------------------------------------------------------------------------

// Drive of input neuron j at step t of dataset sample `sample`.
// Time-to-first-spike step of a feature (encoders.cpp:56-66): brighter is
// earlier, lround((T - 1) * (1 - min(1, x))) clamped to [0, T - 1], in the
// host's float arithmetic (one product, rounded half away from zero).
__device__ __forceinline__ int latency_step(float x, int T) {
  const float xc = fminf(1.0f, x);
  const int ts = (int)lroundf(__fmul_rn((float)(T - 1), __fsub_rn(1.0f, xc)));
  return min(max(ts, 0), T - 1);
}

__device__ __forceinline__ float input_drive(const TrainArgs& a, int j, int sample, int t,
                                             int epoch) {
  const int c = a.channel[j];
  if (c < 0 || c >= a.dim) return 0.0f;
  const float x = a.X[(size_t)sample * a.dim + c];
  if (!(x > a.enc_thr)) return 0.0f;
  if (a.encoder == SNC_ENC_DIRECT) return x * a.gain;
  if (a.encoder == SNC_ENC_LATENCY) return t == latency_step(x, a.lengths[sample]) ? a.gain : 0.0f;
  const Philox4 r = philox4x32_10((uint32_t)(c >> 2), (uint32_t)t, (uint32_t)sample,
                                  (uint32_t)epoch, a.seed_lo, a.seed_hi);
  const float p = fminf(1.0f, x * a.max_rate);
  return r.v[c & 3] < prob_to_u32((double)p) ? a.gain : 0.0f;
}

// LIF + surrogate of one neuron (trainer.cpp:96-125 plus the reset-mode and
// leaky-readout extensions). Returns fired; writes psi.
__device__ __forceinline__ bool train_lif(const TrainArgs& a, uint8_t role, float& v, int& refr,
                                          float ext, float slot, float& psi) {
  const float thr = a.lif.threshold;
  if (a.readout_leaky && role == SNC_ROLE_OUTPUT) {
    v = __fadd_rn(__fadd_rn(__fmul_rn(a.lif.decay, v), ext), slot);
    psi = a.readout_psi;
    return false;
  }
  float vpre;
  bool fired = false, in_refr = false;
  if (refr > 0) {
    --refr;
    in_refr = true;
    if (!a.lif.subtract) v = a.lif.reset;
    vpre = a.lif.subtract ? v : a.lif.reset;
  } else {
    v = __fadd_rn(__fadd_rn(__fmul_rn(a.lif.decay, v), ext), slot);
    vpre = v;
    if (v >= thr) {
      v = a.lif.subtract ? __fsub_rn(v, thr) : a.lif.reset;
      refr = a.lif.refractory;
      fired = true;
    }
  }
  // |vpre - thr| / thr; a threshold of 1 (x / 1 = x exactly) skips the division
  const float dist = fabsf(__fsub_rn(vpre, thr));
  const float d = thr == 1.0f ? dist : __fdiv_rn(dist, thr > 0.0f ? thr : 1.0f);
  psi = (a.lif.subtract && in_refr) ? 0.0f
                                    : (d < 1.0f ? __fmul_rn(a.gamma, __fsub_rn(1.0f, d)) : 0.0f);
  return fired;
}

__device__ __forceinline__ float trace_step(float decay, float tr, bool fired) {
  return __fadd_rn(__fmul_rn(decay, tr), __fmul_rn(__fsub_rn(1.0f, decay), fired ? 1.0f : 0.0f));
}

template <typename ET>
__device__ __forceinline__ void elig_add(ET* E, size_t k, float psi, float tr);
template <>
__device__ __forceinline__ void elig_add<double>(double* E, size_t k, float psi, float tr) {
  // psi * tr is exact in double, so one rounding: the CPU's double sum.
  E[k] = __fma_rn((double)psi, (double)tr, E[k]);
}
template <>
__device__ __forceinline__ void elig_add<float>(float* E, size_t k, float psi, float tr) {
  E[k] = __fmaf_rn(psi, tr, E[k]);
}

// Learning signal of neuron j from the output deltas (trainer.cpp:182-195):
// an output neuron takes its channel's delta, a hidden one sum_k B[j,k] delta_k
// (k ascending, float), every other neuron 0.
__device__ __forceinline__ float learning_signal(const TrainArgs& a, const float* delta, int j) {
  const int K = a.classes;
  const int c = a.channel[j];
  if (a.role[j] == SNC_ROLE_OUTPUT && c >= 0 && c < K) return delta[c];
  if (a.train_hidden && a.internal_row[j] >= 0) {
    const float* br = a.Bfb + (size_t)a.internal_row[j] * K;
    float s = 0.0f;
    for (int k = 0; k < K; ++k) s = __fadd_rn(s, __fmul_rn(br[k], delta[k]));
    return s;
  }
  return 0.0f;
}

// End of sample: delta, loss, prediction, learning signal into Ls[n].
// Runs on the whole block; s_delta holds `classes` floats.
__device__ void sample_epilogue(const TrainArgs& a, int b, const float* logits_in, float* Ls,
                                float* s_delta, int label, int n, int tid, int nthreads) {
  const int K = a.classes;
  if (tid == 0) {
    float mx = logits_in[0];
    for (int k = 1; k < K; ++k) mx = fmaxf(mx, logits_in[k]);
    double sum = 0.0;
    for (int k = 0; k < K; ++k) sum += (double)expf(__fsub_rn(logits_in[k], mx));
    for (int k = 0; k < K; ++k) {
      const float p = __fdiv_rn(expf(__fsub_rn(logits_in[k], mx)), (float)sum);
      s_delta[k] = __fsub_rn(p, k == label ? 1.0f : 0.0f);
    }
    double ls = 0.0;
    for (int k = 0; k < K; ++k) ls += exp((double)__fsub_rn(logits_in[k], mx));
    int pred = 0;
    for (int k = 1; k < K; ++k)
      if (logits_in[k] > logits_in[pred]) pred = k;
    a.loss[b] = -(double)__fsub_rn(logits_in[label], mx) + log(ls);
    a.correct[b] = pred == label ? 1 : 0;
    for (int k = 0; k < K; ++k) a.logits[(size_t)b * K + k] = logits_in[k];
  }
  __syncthreads();
  for (int j = tid; j < n; j += nthreads) Ls[j] = learning_signal(a, s_delta, j);
}

// ---------------------------------------------------------------------------
// batched persistent kernel: blockIdx.x = position in the minibatch
// ---------------------------------------------------------------------------

constexpr int kFoldBatch = 8;       // staged arrivals a fold round loads together
constexpr int kDeposit = 4;         // deposits a lane loads before queueing them
constexpr int kHeld = 6;            // arrivals a thread keeps in registers while sorting
constexpr int kRowsHeld = 4;        // fired rows a wave keeps in registers from bits to staging

// Ordered fold of target j's arrivals into its ring slots. Columns are
// ordered delay descending, so one step's arrivals for one ring slot are one
// contiguous run of the column, added in (pre, CSR position) order onto the
// slot's previous value -- the CPU's sum bit for bit -- and each slot is read
// and written once; the ring values of kFoldBatch arrivals are therefore
// loaded together before their ordered adds.
//
// fold_staged: the step's arrivals are staged in LDS in column order (rank
// among the set bits of the bitmap), so target j's are s_cval[r0, r1).
__device__ __forceinline__ void fold_staged(const TrainArgs& a, int j, int r0, int r1,
                                            const uint2* s_cval, float* ring, int n, int tm) {
  const int L = a.L;
  if (r0 == r1) return;
  // Arrivals with one delay (first and last equal: the column is ordered by
  // delay) all land in one slot: one ring read, the adds in order, one write.
  if (s_cval[r0].y == s_cval[r1 - 1].y) {
    int s = tm + (int)s_cval[r0].y;
    if (s >= L) s -= L;
    float* slot = ring + (size_t)s * n + j;
    float acc = *slot;
    int r = r0;
    for (; r + 4 <= r1; r += 4) {
      const float v0 = __uint_as_float(s_cval[r].x), v1 = __uint_as_float(s_cval[r + 1].x);
      const float v2 = __uint_as_float(s_cval[r + 2].x), v3 = __uint_as_float(s_cval[r + 3].x);
      acc = __fadd_rn(__fadd_rn(__fadd_rn(__fadd_rn(acc, v0), v1), v2), v3);
    }
    for (; r < r1; ++r) acc = __fadd_rn(acc, __uint_as_float(s_cval[r].x));
    *slot = acc;
    return;
  }
  int cur = -1;
  float acc = 0.0f;
  for (int r = r0; r < r1; r += kFoldBatch) {
    int sl[kFoldBatch];
    float val[kFoldBatch], old[kFoldBatch];
#pragma unroll
    for (int u = 0; u < kFoldBatch; ++u) {
      sl[u] = -1;
      val[u] = 0.0f;
      if (r + u < r1) {
        const uint2 cv = s_cval[r + u];
        val[u] = __uint_as_float(cv.x);
        const int s = tm + (int)cv.y;
        sl[u] = s >= L ? s - L : s;
      }
    }
#pragma unroll
    for (int u = 0; u < kFoldBatch; ++u)
      old[u] = sl[u] >= 0 && sl[u] != cur ? ring[(size_t)sl[u] * n + j] : 0.0f;
#pragma unroll
    for (int u = 0; u < kFoldBatch; ++u) {
      if (sl[u] < 0) break;
      if (sl[u] != cur) {
        if (cur >= 0) ring[(size_t)cur * n + j] = acc;
        cur = sl[u];
        acc = old[u];
      }
      acc = __fadd_rn(acc, val[u]);
    }
  }
  if (cur >= 0) ring[(size_t)cur * n + j] = acc;
}

// fold_gather: a step with more arrivals than the staging holds walks the set
// bits of target j's column words [w0, w1) (clearing them) and gathers each
// arrival's weight and delay from the column arrays in HBM.
__device__ __forceinline__ void fold_gather(const TrainArgs& a, int j, int w0, int w1,
                                            uint32_t* s_bits, float* ring, int n, int tm) {
  const int L = a.L;
  int cur = -1;
  float acc = 0.0f;
  for (int w = w0; w < w1; ++w) {
    uint32_t bw = s_bits[w];
    if (!bw) continue;
    s_bits[w] = 0u;
    while (bw) {
      const int p = (w << 5) | __builtin_ctz(bw);
      bw &= bw - 1u;
      int s = tm + (int)a.cdel[p];
      if (s >= L) s -= L;
      if (s != cur) {
        if (cur >= 0) ring[(size_t)cur * n + j] = acc;
        cur = s;
        acc = ring[(size_t)s * n + j];
      }
      acc = __fadd_rn(acc, a.wc[p]);
    }
  }
  if (cur >= 0) ring[(size_t)cur * n + j] = acc;
}

// Phase timer of the profiled instantiation (PROF): after each barrier that
// ends a phase, thread 0 adds the wall-clock ticks since the previous mark to
// that phase's total; the totals go to a.prof [B][kTrainPhases].
template <bool PROF>
__device__ __forceinline__ void phase_mark(unsigned long long* s_ph, long long* s_last, int k,
                                           int tid) {
  if constexpr (PROF) {
    if (tid == 0) {
      const long long now = wall_clock64();
      s_ph[k] += (unsigned long long)(now - *s_last);
      *s_last = now;
    }
  }
}

// Eligibility and contributions of one round of sample b (work units u0 ..
// u0 + kTrainThreads - 1) after its T steps, from the psi and trace histories
// the step loop wrote:
//   E_s = sum_t psi_t[post_s] * tr_t[pre_s],   t ascending, one fma per step,
// exactly the per-step accumulation of trainer.cpp:132-139 (the product of
// two floats is exact in double, so fma == add of the exact product).
// The histories stream through LDS in chunks of CH steps, transposed in
// groups of four steps: float4 q * n + j holds steps 4q .. 4q + 3 of neuron
// j's psi (and, in a second array, its trace), so one 16-byte load gives a
// lane four steps of one neuron, and that load's LDS bank group (16-byte
// granule mod 8) is j mod 8 whatever CH is. The units deal their synapses
// so that the eight lanes one LDS pass serves read pres of different
// classes mod 8, so those gathers do not conflict; the transposing writes
// (consecutive j) do not either. Each lane takes one work unit: up to kEligPerThread synapses of
// one post, in column order (a.eunit), with their sums in registers; it reads
// the post's four psi values once and reuses them for every synapse of the
// unit, so a fused multiply-add costs about 4 bytes of LDS reads (the
// synapse's trace) instead of 8. Steps past T in the last chunk are zero and
// add +0 to a nonnegative sum, changing nothing.
template <typename ET, int CH, bool PROF>
__device__ void elig_contrib(const TrainArgs& a, int b, int T, float scale, const float* Ls,
                             float* s_chunk, int tid, int u0, unsigned long long* s_ph,
                             long long* s_last) {
  static_assert(CH % 4 == 0, "chunk rows are read four steps at a time");
  const int hn = a.hn, n = a.n;
  float4* ps_t = reinterpret_cast<float4*>(s_chunk);                   // [CH / 4][n]
  float4* tr_t = reinterpret_cast<float4*>(s_chunk + (size_t)n * CH);  // [CH / 4][n]
  const float* hb = a.hist + (size_t)b * a.T_max * 2 * hn;
  {
    const int u = u0 + tid;
    int post = 0, len = 0, q0 = 0;
    if (u < a.n_units) {
      const int2 unit = a.eunit[u];
      post = unit.x >> 5;
      len = (unit.x & 31) + 1;
      q0 = unit.y;
    }
    // the unit's pres; a unit shorter than kEligPerThread repeats its last
    // pre, so the sum loop has no per-synapse branches (the extra sums are
    // never stored)
    int tro[kEligPerThread];
    ET E[kEligPerThread];
#pragma unroll
    for (int k = 0; k < kEligPerThread; ++k) {
      tro[k] = len > 0 ? a.ecol_pre[q0 + min(k, len - 1)] : 0;
      E[k] = (ET)0;
    }
    for (int t0 = 0; t0 < T; t0 += CH) {
      const int tc = min(CH, T - t0);
      __syncthreads();  // previous chunk consumed
      phase_mark<PROF>(s_ph, s_last, 7, tid);
      // transpose in groups of 4 steps: 8 coalesced loads in flight per lane
      for (int c = 0; c < CH; c += 4) {
        for (int j = tid; j < n; j += kTrainThreads) {
          float pv[4], tv[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float* row = hb + (size_t)(t0 + c + v) * 2 * hn;
            pv[v] = c + v < tc ? row[j] : 0.0f;
            tv[v] = c + v < tc ? row[hn + j] : 0.0f;
          }
          ps_t[(size_t)(c / 4) * n + j] = make_float4(pv[0], pv[1], pv[2], pv[3]);
          tr_t[(size_t)(c / 4) * n + j] = make_float4(tv[0], tv[1], tv[2], tv[3]);
        }
      }
      __syncthreads();
      phase_mark<PROF>(s_ph, s_last, 8, tid);
      if (len > 0) {
        constexpr int kBatch = 4;  // trace rows loaded together before their sums (8 spilled VGPRs)
#pragma unroll 1
        for (int q = 0; q < CH / 4; ++q) {
          const float4 p = ps_t[(size_t)q * n + post];
#pragma unroll
          for (int k0 = 0; k0 < kEligPerThread; k0 += kBatch) {
            float4 r[kBatch];
#pragma unroll
            for (int k = 0; k < kBatch; ++k) r[k] = tr_t[(size_t)q * n + tro[k0 + k]];
#pragma unroll
            for (int k = 0; k < kBatch; ++k) {
              ET e[1] = {E[k0 + k]};
              elig_add<ET>(e, 0, p.x, r[k].x);
              elig_add<ET>(e, 0, p.y, r[k].y);
              elig_add<ET>(e, 0, p.z, r[k].z);
              elig_add<ET>(e, 0, p.w, r[k].w);
              E[k0 + k] = e[0];
            }
          }
        }
      }
    }
    if constexpr (PROF) {
      __syncthreads();
      phase_mark<PROF>(s_ph, s_last, 7, tid);
    }
    // C is stored in unit order (each post's synapses in column order), so a
    // lane writes its unit's contributions to consecutive addresses.
    if (a.C && len > 0) {
      const float ls = __fmul_rn(scale, Ls[post]);
      float* cb = a.C + (size_t)b * a.S + q0;
#pragma unroll
      for (int k = 0; k < kEligPerThread; ++k)
        if (k < len) cb[k] = a.reward ? (float)E[k] : __fmul_rn(ls, (float)E[k]);
    }
    if constexpr (PROF) {
      __syncthreads();
      phase_mark<PROF>(s_ph, s_last, 9, tid);
    }
  }
}

// Largest chunk length (steps) whose two transposed arrays fit in `bytes`.
__host__ __device__ inline int elig_chunk_steps(int n, size_t bytes) {
  const size_t per_step = sizeof(float) * 2 * (size_t)n;
  const int fit = per_step ? (int)(bytes / per_step) : 0;
  return fit >= 28 ? 28 : fit >= 20 ? 20 : fit >= 12 ? 12 : fit >= 8 ? 8 : fit >= 4 ? 4 : 0;
}

// Steps of minibatch sample b (workgroup b): the whole step loop, the
// end-of-sample epilogue, and -- when contributions are wanted -- the
// learning signals to HBM and the sample's publication to the eligibility
// work queue.
template <typename ET, bool RING_LDS, bool PROF>
__device__ void train_sample(const TrainArgs& a, unsigned char* smem, int b,
                             unsigned long long* s_ph, long long* s_last_p) {
  const int n = a.n, L = a.L, tid = threadIdx.x;
  long long& s_last = *s_last_p;
  float* s_v = (float*)smem;
  float* s_tr = s_v + n;
  float* s_ext = s_tr + n;
  int* s_refr = (int*)(s_ext + n);
  int* s_fired = s_refr + n;                      // fired list
  uint32_t* s_fw = (uint32_t*)(s_fired + n);      // fired bitmap [words]
  uint32_t* s_bits = s_fw + a.words;              // delivery buckets [bits_words]
  float* s_logit = (float*)(s_bits + a.bits_words);  // [classes]
  float* s_delta = s_logit + a.classes;              // [classes]
  uint32_t* s_rand = (uint32_t*)(s_delta + a.classes);  // [rand_words] Philox words of one step
  const int rand_words = train_rand_words(a);
  const size_t state_bytes = train_state_bytes(n, a.words, a.bits_words, a.classes, rand_words);
  const size_t const_bytes = train_const_bytes(n, a.n_inputs);
  int32_t* s_rowp = (int32_t*)(smem + state_bytes);  // [n + 1] row_ptr
  int32_t* s_ccol = s_rowp + n + 1;                  // [n + 1] column starts
  int32_t* s_chan = s_ccol + n + 1;                  // [n] channel
  int32_t* s_inp = s_chan + n;                       // [n_inputs] input neuron ids
  float* s_inpx = (float*)(s_inp + a.n_inputs);      // [n_inputs] drive value / probability
  uint8_t* s_role = (uint8_t*)(s_inpx + a.n_inputs); // [n] role
  float* s_ring = (float*)(smem + state_bytes + const_bytes);  // [L][n] when RING_LDS
  // Arrival queues instead of a ring in HBM (block-uniform): the queues of
  // this sample in HBM, their fill counts in LDS, and each target's sum of
  // the coming step's arrivals in LDS where the ring would start.
  const bool queue = !RING_LDS && a.arq != nullptr;
  uint2* arq = queue ? a.arq + (size_t)b * L * a.arq_cap : nullptr;
  float* s_in = s_ring;  // [n] (queue)
  __shared__ int s_qcnt[256];
  // Fold staging, in the bytes the history chunks take after the step loop:
  // the bitmap's exclusive popcount prefix per word, then one {weight, delay}
  // (one weight, with the queues) per arrival of the step, at its rank; `cap`
  // arrivals fit.
  const size_t fold_base =
      state_bytes + const_bytes +
      (RING_LDS ? (sizeof(float) * (size_t)L * n + 15) / 16 * 16 : queue ? train_in_bytes(n) : 0);
  // Word prefixes are 16-bit: a staged step has at most 65535 arrivals.
  const size_t pref_bytes = (sizeof(uint16_t) * ((size_t)a.bits_words + 1) + 15) / 16 * 16;
  uint16_t* s_pref = (uint16_t*)(smem + fold_base);
  uint2* s_cval = (uint2*)(smem + fold_base + pref_bytes);
  float* s_val = (float*)s_cval;  // (queue)
  const size_t stage_elem = queue ? sizeof(float) : sizeof(uint2);
  const int cap = (size_t)a.lds_bytes > fold_base + pref_bytes
                      ? (int)min(((size_t)a.lds_bytes - fold_base - pref_bytes) / stage_elem,
                                 (size_t)65535)
                      : 0;
  __shared__ int s_nf, s_arr;
  __shared__ int s_wsum[kTrainThreads / kWave];
  __shared__ unsigned long long s_ev, s_sp;

  const int sample = a.idx[b];
  const int T = a.lengths[sample];
  const int label = a.labels[sample];
  float* ring = RING_LDS ? s_ring : a.ring_g + (size_t)b * L * n;
  float* hist = a.hist + (size_t)b * a.T_max * 2 * a.hn;
  const int ev0 = a.ev_base ? a.ev_base[sample] : 0;

  for (int j = tid; j < n; j += kTrainThreads) {
    s_v[j] = 0.0f;
    s_tr[j] = 0.0f;
    s_ext[j] = 0.0f;
    s_refr[j] = 0;
    s_chan[j] = a.channel[j];
    s_role[j] = a.role[j];
  }
  for (int j = tid; j <= n; j += kTrainThreads) {
    s_rowp[j] = a.row_ptr[j];
    s_ccol[j] = a.ccol[j];
  }
  // Input drive of this sample, from the features once: the direct value,
  // the Poisson spike probability (0: the channel never fires), or the
  // latency spike step as int bits (-1: never), or for Poisson the spike
  // threshold on a Philox word as uint bits (0: never).
  if (!a.ev_base)
    for (int q = tid; q < a.n_inputs; q += kTrainThreads) {
      const int j = a.inputs[q];
      const int c = a.channel[j];
      const bool lat = a.encoder == SNC_ENC_LATENCY;
      float val = lat ? __int_as_float(-1) : 0.0f;
      if (c >= 0 && c < a.dim) {
        const float x = a.X[(size_t)sample * a.dim + c];
        if (x > a.enc_thr)
          val = a.encoder == SNC_ENC_DIRECT ? x * a.gain
                : lat                       ? __int_as_float(latency_step(x, T))
                                            : __uint_as_float(prob_to_u32(
                                                  (double)fminf(1.0f, x * a.max_rate)));
      }
      s_inp[q] = j;
      s_inpx[q] = val;
    }
  for (int k = tid; k < a.bits_words; k += kTrainThreads) s_bits[k] = 0u;
  for (int k = tid; k < a.classes; k += kTrainThreads) s_logit[k] = 0.0f;
  if (queue) {
    for (int j = tid; j < n; j += kTrainThreads) s_in[j] = 0.0f;
    for (int k = tid; k < L; k += kTrainThreads) s_qcnt[k] = 0;
  } else {
    for (size_t k = tid; k < (size_t)L * n; k += kTrainThreads) ring[k] = 0.0f;
  }
  if (tid == 0) {
    s_nf = 0;
    s_arr = 0;
    s_ev = 0;
    s_sp = 0;
  }
  __syncthreads();

  // injections of step t into s_ext (entries are pre-folded per neuron)
  auto inject = [&](int t) {
    if (a.ev_base) {
      for (int k = a.ev_step_ptr[ev0 + t] + tid; k < a.ev_step_ptr[ev0 + t + 1];
           k += kTrainThreads)
        s_ext[a.ev_nid[k]] = a.ev_amp[k];
    } else if (a.encoder == SNC_ENC_DIRECT) {
      for (int q = tid; q < a.n_inputs; q += kTrainThreads)
        if (s_inpx[q] != 0.0f) s_ext[s_inp[q]] = s_inpx[q];
    } else if (a.encoder == SNC_ENC_LATENCY) {
      for (int q = tid; q < a.n_inputs; q += kTrainThreads)
        if (__float_as_int(s_inpx[q]) == t) s_ext[s_inp[q]] = a.gain;
    } else {
      // Poisson: the words of step t, staged in s_rand by fill_rand(t)
      for (int q = tid; q < a.n_inputs; q += kTrainThreads) {
        const uint32_t thr = __float_as_uint(s_inpx[q]);
        if (thr == 0u) continue;
        const int j = s_inp[q];
        if (s_rand[s_chan[j]] < thr) s_ext[j] = a.gain;
      }
    }
  };
  // The Philox block of channel group g = c / 4 at step t (counter
  // {g, t, sample, epoch}), computed once per group instead of once per
  // channel; the highest threads take it, since the LIF phase gives the
  // lowest ones a second neuron when n > kTrainThreads.
  auto fill_rand = [&](int t) {
    for (int g = kTrainThreads - 1 - tid; g < rand_words / 4; g += kTrainThreads) {
      const Philox4 r = philox4x32_10((uint32_t)g, (uint32_t)t, (uint32_t)sample,
                                      (uint32_t)a.epoch, a.seed_lo, a.seed_hi);
#pragma unroll
      for (int i = 0; i < 4; ++i) s_rand[4 * g + i] = r.v[i];
    }
  };
  if (T > 0 && rand_words) {
    fill_rand(0);
    __syncthreads();
  }
  if (T > 0) inject(0);
  __syncthreads();
  phase_mark<PROF>(s_ph, &s_last, 0, tid);

  int tm = 0;
  for (int t = 0; t < T; ++t) {
    // (A) LIF + psi, trace, histories, fired bitmap and list
    float* slot = ring + (size_t)tm * n;
    float* hrow = hist + (size_t)t * 2 * a.hn;
    for (int j0 = 0; j0 < n; j0 += kTrainThreads) {
      const int j = j0 + tid;
      bool f = false;
      if (j < n) {
        float v = s_v[j];
        int r = s_refr[j];
        float psi;
        const uint8_t role = s_role[j];
        f = train_lif(a, role, v, r, s_ext[j], queue ? s_in[j] : slot[j], psi);
        s_v[j] = v;
        s_refr[j] = r;
        s_ext[j] = 0.0f;
        if (queue) s_in[j] = 0.0f;
        else slot[j] = 0.0f;
        const float tr = trace_step(a.lif.decay, s_tr[j], f);
        s_tr[j] = tr;
        if (a.C) {  // histories feed the update only
          hrow[j] = psi;
          hrow[a.hn + j] = tr;
        }
        const int ch = s_chan[j];
        if (f) {
          if (a.neur_fire) atomicAdd(a.neur_fire + j, 1ull);
          if (!a.readout_leaky && role == SNC_ROLE_OUTPUT && ch >= 0) atomicAdd(&s_logit[ch], 1.0f);
        }
        if (a.readout_leaky && role == SNC_ROLE_OUTPUT && ch >= 0)
          s_logit[ch] = __fadd_rn(s_logit[ch], v);  // one readout per channel
      }
      const uint64_t m = ballot(f);
      if (lane_id() == 0 && j - lane_id() < n) {
        const int w0 = (j - lane_id()) >> 5;
        s_fw[w0] = (uint32_t)m;
        if (w0 + 1 < a.words) s_fw[w0 + 1] = (uint32_t)(m >> 32);
      }
      if (m) {
        int pos = 0;
        if (lane_id() == 0) pos = atomicAdd(&s_nf, __popcll(m));
        pos = __shfl(pos, 0, kWave);
        if (f) s_fired[pos + rank_below(m)] = j;
      }
    }
    // next step's Philox words: inject(t) has read this step's (it ran
    // before the barrier that ended the previous step)
    if (rand_words && t + 1 < T) fill_rand(t + 1);
    __syncthreads();
    phase_mark<PROF>(s_ph, &s_last, 1, tid);
    // (B) raster row; fired rows set bucket bits
    if (a.raster)
      for (int w = tid; w < a.words; w += kTrainThreads)
        a.raster[((size_t)b * a.T_max + t) * a.words + w] = s_fw[w];
    const int nf = s_nf;
    const int wave = tid >> 6, lane = lane_id();
    constexpr int kWaves = kTrainThreads / kWave;
    unsigned long long ev = 0;
    // Bucket bits (no queues): the first 64 entries of the wave's first
    // kRowsHeld fired rows keep their column position, weight and delay in
    // registers until the staging below, which then reads no row twice.
    int hp[kRowsHeld];
    uint2 hv[kRowsHeld];
    if (queue) {
      // each deposit joins the queue of the slot it arrives at; a lane loads
      // kDeposit deposits before it reserves and stores any
      for (int k = wave; k < nf; k += kWaves) {
        const int i = s_fired[k];
        const int rb = s_rowp[i], re = s_rowp[i + 1];
        for (int e0 = rb + lane; e0 < re; e0 += kDeposit * kWave) {
          int sl[kDeposit];
          uint2 rec[kDeposit];
#pragma unroll
          for (int u = 0; u < kDeposit; ++u) {
            const int e = e0 + u * kWave;
            sl[u] = -1;
            if (e < re) {
              sl[u] = tm + (int)a.sdel[e];
              rec[u] = make_uint2((uint32_t)a.cpos[e], __float_as_uint(a.w_csr[e]));
            }
          }
#pragma unroll
          for (int u = 0; u < kDeposit; ++u)
            if (sl[u] >= 0) {
              if (sl[u] >= L) sl[u] -= L;
              const int q = atomicAdd(&s_qcnt[sl[u]], 1);
              arq[(size_t)sl[u] * a.arq_cap + q] = rec[u];
            }
        }
        ev += (unsigned long long)(re - rb);
      }
    } else {
#pragma unroll
      for (int r = 0; r < kRowsHeld; ++r) {
        const int k = wave + r * kWaves;
        hp[r] = -1;
        if (k < nf) {
          const int i = s_fired[k];
          const int e = s_rowp[i] + lane;
          if (e < s_rowp[i + 1]) {
            hp[r] = a.cpos[e];
            hv[r] = make_uint2(__float_as_uint(a.w_csr[e]), (uint32_t)a.sdel[e]);
          }
        }
      }
      asm volatile("" ::: "memory");  // every held row's loads in flight before any is used
#pragma unroll
      for (int r = 0; r < kRowsHeld; ++r)
        if (hp[r] >= 0) atomicOr(&s_bits[hp[r] >> 5], 1u << (hp[r] & 31));
      for (int k = wave; k < nf; k += kWaves) {
        const int i = s_fired[k];
        const int rb = s_rowp[i], re = s_rowp[i + 1];
        const bool held = k < wave + kRowsHeld * kWaves;
        for (int e = rb + lane + (held ? kWave : 0); e < re; e += kWave) {
          const int p = a.cpos[e];
          atomicOr(&s_bits[p >> 5], 1u << (p & 31));
        }
        ev += (unsigned long long)(re - rb);
      }
    }
    if (lane == 0 && ev) {
      atomicAdd(&s_ev, ev);
      atomicAdd(&s_arr, (int)ev);
    }
    __syncthreads();
    phase_mark<PROF>(s_ph, &s_last, 2, tid);
    if (queue) {
      // The arrivals of step t + 1, sorted by column position (presence
      // bits, their prefix: each record's rank), each target's run added
      // from zero in column order -- the ring slot sum of the CPU path.
      const int s1 = tm + 1 == L ? 0 : tm + 1;
      const int count = t + 1 < T ? s_qcnt[s1] : 0;
      const uint2* qa = arq + (size_t)s1 * a.arq_cap;
      // the first kHeld records of each thread stay in registers for the
      // staging below
      uint2 held[kHeld];
#pragma unroll
      for (int u = 0; u < kHeld; ++u) {
        const int k = tid + u * kTrainThreads;
        held[u] = k < count ? qa[k] : make_uint2(0u, 0u);
      }
#pragma unroll
      for (int u = 0; u < kHeld; ++u)
        if (tid + u * kTrainThreads < count) {
          const int p = (int)held[u].x;
          atomicOr(&s_bits[p >> 5], 1u << (p & 31));
        }
      for (int k = tid + kHeld * kTrainThreads; k < count; k += kTrainThreads) {
        const int p = (int)qa[k].x;
        atomicOr(&s_bits[p >> 5], 1u << (p & 31));
      }
      __syncthreads();
      if (count > 0 && count <= cap) {
        const int W = a.bits_words;
        const int per = (W + kTrainThreads - 1) / kTrainThreads;
        const int lo = min(W, tid * per), hi = min(W, lo + per);
        int cnt = 0;
        for (int w = lo; w < hi; ++w) cnt += __popc(s_bits[w]);
        const int incl = wave_inclusive_scan(cnt);
        if (lane == kWave - 1) s_wsum[wave] = incl;
        __syncthreads();
        int run = incl - cnt;
        for (int k = 0; k < wave; ++k) run += s_wsum[k];
        for (int w = lo; w < hi; ++w) {
          s_pref[w] = (uint16_t)run;
          run += __popc(s_bits[w]);
        }
        if (hi == W) s_pref[W] = (uint16_t)run;
        __syncthreads();
        phase_mark<PROF>(s_ph, &s_last, 3, tid);
        auto stage = [&](uint2 e) {
          const int p = (int)e.x, q = p >> 5;
          const int rk = (int)s_pref[q] + __popc(s_bits[q] & ((1u << (p & 31)) - 1u));
          s_val[rk] = __uint_as_float(e.y);
        };
#pragma unroll
        for (int u = 0; u < kHeld; ++u)
          if (tid + u * kTrainThreads < count) stage(held[u]);
        for (int k = tid + kHeld * kTrainThreads; k < count; k += kTrainThreads) stage(qa[k]);
        __syncthreads();
        phase_mark<PROF>(s_ph, &s_last, 4, tid);
        for (int j = tid; j < n; j += kTrainThreads) {
          const int r0 = (int)s_pref[s_ccol[j] >> 5], r1 = (int)s_pref[s_ccol[j + 1] >> 5];
          float acc = 0.0f;
          int r = r0;
          for (; r + 4 <= r1; r += 4)
            acc = __fadd_rn(__fadd_rn(__fadd_rn(__fadd_rn(acc, s_val[r]), s_val[r + 1]),
                                      s_val[r + 2]),
                            s_val[r + 3]);
          for (; r < r1; ++r) acc = __fadd_rn(acc, s_val[r]);
          s_in[j] = acc;
        }
#pragma unroll
        for (int u = 0; u < kHeld; ++u)
          if (tid + u * kTrainThreads < count) s_bits[(int)held[u].x >> 5] = 0u;
        for (int k = tid + kHeld * kTrainThreads; k < count; k += kTrainThreads)
          s_bits[(int)qa[k].x >> 5] = 0u;
      } else if (count > 0) {
        // more arrivals than the staging holds: each target walks its
        // column's set bits in order, weights from the column-ordered copy
        for (int j = tid; j < n; j += kTrainThreads) {
          float acc = 0.0f;
          for (int w = s_ccol[j] >> 5; w < (s_ccol[j + 1] >> 5); ++w) {
            uint32_t bw = s_bits[w];
            if (!bw) continue;
            s_bits[w] = 0u;
            while (bw) {
              const int p = (w << 5) | __builtin_ctz(bw);
              bw &= bw - 1u;
              acc = __fadd_rn(acc, a.wc[p]);
            }
          }
          s_in[j] = acc;
        }
      }
      if (tid == 0 && t + 1 < T) s_qcnt[s1] = 0;  // every thread read the count before a barrier
    }
    const bool staged = !queue && s_arr <= cap;  // block-uniform
    if (staged) {
      // (B2) exclusive popcount prefix of the bitmap words (block scan)
      const int W = a.bits_words;
      const int per = (W + kTrainThreads - 1) / kTrainThreads;
      const int lo = min(W, tid * per), hi = min(W, lo + per);
      int cnt = 0;
      for (int w = lo; w < hi; ++w) cnt += __popc(s_bits[w]);
      const int incl = wave_inclusive_scan(cnt);
      if (lane == kWave - 1) s_wsum[wave] = incl;
      __syncthreads();
      int run = incl - cnt;
      for (int k = 0; k < wave; ++k) run += s_wsum[k];
      for (int w = lo; w < hi; ++w) {
        s_pref[w] = (uint16_t)run;
        run += __popc(s_bits[w]);
      }
      if (hi == W) s_pref[W] = (uint16_t)run;  // the total (every such thread writes the same)
      __syncthreads();
      phase_mark<PROF>(s_ph, &s_last, 3, tid);
      // (B3) every arrival's weight and delay at its rank: the held entries
      // from registers, the rest from the fired rows read in CSR order
      // (coalesced)
#pragma unroll
      for (int r = 0; r < kRowsHeld; ++r)
        if (hp[r] >= 0) {
          const int p = hp[r], q = p >> 5;
          const int rk = (int)s_pref[q] + __popc(s_bits[q] & ((1u << (p & 31)) - 1u));
          s_cval[rk] = hv[r];
        }
      for (int k = wave; k < nf; k += kWaves) {
        const int i = s_fired[k];
        const int rb = s_rowp[i], re = s_rowp[i + 1];
        const bool held = k < wave + kRowsHeld * kWaves;
        for (int e = rb + lane + (held ? kWave : 0); e < re; e += kWave) {
          const int p = a.cpos[e];
          const float wv = a.w_csr[e];
          const uint32_t d = a.sdel[e];
          const int q = p >> 5;
          const int rk = (int)s_pref[q] + __popc(s_bits[q] & ((1u << (p & 31)) - 1u));
          s_cval[rk] = make_uint2(__float_as_uint(wv), d);
        }
      }
      __syncthreads();
      phase_mark<PROF>(s_ph, &s_last, 4, tid);
    }
    // (C) ordered fold into the ring (staged: the bitmap is only cleared);
    // next step's injections
    if (queue) {
    } else if (staged) {
      for (int j = tid; j < n; j += kTrainThreads)
        fold_staged(a, j, (int)s_pref[s_ccol[j] >> 5], (int)s_pref[s_ccol[j + 1] >> 5], s_cval,
                    ring, n, tm);
      for (int w = tid; w < a.bits_words; w += kTrainThreads) s_bits[w] = 0u;
    } else {
      for (int j = tid; j < n; j += kTrainThreads)
        fold_gather(a, j, s_ccol[j] >> 5, s_ccol[j + 1] >> 5, s_bits, ring, n, tm);
    }
    if (t + 1 < T) inject(t + 1);
    if (tid == 0) {
      s_sp += (unsigned long long)nf;
      s_nf = 0;
      s_arr = 0;
    }
    tm = tm + 1 == L ? 0 : tm + 1;
    __syncthreads();
    phase_mark<PROF>(s_ph, &s_last, 5, tid);
  }

  if (a.readout_leaky && tid < a.classes)
    s_logit[tid] = __fdiv_rn(s_logit[tid], (float)max(1, T));
  __syncthreads();
  // The state arrays are dead from here: the learning signal takes s_v's
  // place and the rest of LDS holds history chunks.
  float* Ls = s_v;
  sample_epilogue(a, b, s_logit, Ls, s_delta, label, n, tid, kTrainThreads);
  if constexpr (PROF) {
    __syncthreads();
    phase_mark<PROF>(s_ph, &s_last, 6, tid);
  }
  if (tid == 0) {
    a.blk_spikes[b] = s_sp;
    a.blk_events[b] = s_ev;
  }
  if (a.C) {
    for (int j = tid; j < n; j += kTrainThreads) a.Ls_g[(size_t)b * n + j] = Ls[j];
    // Publish: every thread's history and learning-signal stores are made
    // visible device-wide (all XCDs) before the sample enters the queue.
    __threadfence();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&a.wq[b], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The eligibility work queue. Sample b's eligibility is `rounds` items
// (round r: work units r * kTrainThreads ...). wq[b] is 1 once sample b is
// published (its histories and learning signals are in HBM), wq[B + b]
// counts the claims of its rounds. Wave 0 of a workgroup scans the samples 64
// at a time, starting from its own position, and claims a round of a
// published sample with one atomic add, so claims of different samples never
// meet on one address; a claim beyond `rounds` is simply ignored. An item is
// only claimed once its sample is published, so a claimed item never waits.
// The scan returns -1 when every round of every sample is claimed. While some
// sample is still stepping it polls, and gives up after `wait_ticks` of wall
// clock: a workgroup that gives up leaves its work to the others, and each
// sample's own workgroup claims its sample's rounds right after publishing
// it, so every item is processed exactly once. The polls are relaxed loads
// (an acquire on every poll would invalidate this XCD's L2 each time and
// evict the data of the samples still stepping); one acquire fence follows a
// successful claim, and the workgroup barrier after it orders the other
// waves' loads of the item's data behind it.
__device__ int claim_item(const TrainArgs& a, int start) {
  const int B = a.n_samples, R = a.rounds, lane = lane_id();
  const long long t0 = wall_clock64();
  for (;;) {
    bool unfinished = false;
    for (int c0 = 0; c0 < B; c0 += kWave) {
      const int k = c0 + lane;
      const int c = k < B ? (start + k) % B : 0;
      int pub = 0, cl = R;
      if (k < B) {
        cl = __hip_atomic_load(&a.wq[B + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        pub = __hip_atomic_load(&a.wq[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (ballot(cl < R)) unfinished = true;
      uint64_t avail = ballot(pub != 0 && cl < R);
      while (avail) {
        const int l = __builtin_ctzll(avail);
        avail &= avail - 1;
        int r = 0;
        if (lane == l) r = atomicAdd(&a.wq[B + c], 1);
        r = __shfl(r, l, kWave);
        if (r < R) {
          const int cc = __shfl(c, l, kWave);
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          return cc * R + r;
        }
      }
    }
    if (!unfinished) return -1;
    if (wall_clock64() - t0 > a.wait_ticks) return -1;
    __builtin_amdgcn_s_sleep(127);
  }
}

template <typename ET, bool PROF>
__device__ void elig_queue(const TrainArgs& a, unsigned char* smem, unsigned long long* s_ph,
                           long long* s_last) {
  const int n = a.n, tid = threadIdx.x;
  __shared__ int s_item;
  const size_t ls_bytes = (sizeof(float) * (size_t)n + 15) / 16 * 16;
  float* Ls = (float*)smem;
  float* s_chunk = (float*)(smem + ls_bytes);
  const int ch = elig_chunk_steps(n, (size_t)a.lds_bytes - ls_bytes);  // the host checked >= 4
  for (;;) {
    __syncthreads();  // the previous item's LDS is consumed
    if (tid < kWave) {
      const int it = claim_item(a, (int)blockIdx.x % a.n_samples);
      if (tid == 0) s_item = it;
    }
    __syncthreads();
    phase_mark<PROF>(s_ph, s_last, 10, tid);
    const int it = s_item;
    if (it < 0) break;
    const int b = it / a.rounds, r = it - b * a.rounds;
    const int T = a.lengths[a.idx[b]];
    const float scale = __fdiv_rn(a.lr, (float)max(1, T));
    for (int j = tid; j < n; j += kTrainThreads) Ls[j] = a.Ls_g[(size_t)b * n + j];
    const int u0 = r * kTrainThreads;
    switch (ch) {
      case 28: elig_contrib<ET, 28, PROF>(a, b, T, scale, Ls, s_chunk, tid, u0, s_ph, s_last); break;
      case 20: elig_contrib<ET, 20, PROF>(a, b, T, scale, Ls, s_chunk, tid, u0, s_ph, s_last); break;
      case 12: elig_contrib<ET, 12, PROF>(a, b, T, scale, Ls, s_chunk, tid, u0, s_ph, s_last); break;
      case 8: elig_contrib<ET, 8, PROF>(a, b, T, scale, Ls, s_chunk, tid, u0, s_ph, s_last); break;
      default: elig_contrib<ET, 4, PROF>(a, b, T, scale, Ls, s_chunk, tid, u0, s_ph, s_last); break;
    }
  }
}

// One launch per minibatch: workgroups 0 .. B-1 each run one sample's steps
// (train_sample); then every workgroup, including the helpers B .. grid-1,
// takes eligibility rounds from the queue until it is empty.
template <typename ET, bool RING_LDS, bool PROF>
__global__ __launch_bounds__(kTrainThreads) void k_train_batched(TrainArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  __shared__ unsigned long long s_ph[kTrainPhases];
  __shared__ long long s_last;
  if constexpr (PROF) {
    if (tid < kTrainPhases) s_ph[tid] = 0ull;
    if (tid == 0) s_last = wall_clock64();
  }
  if ((int)blockIdx.x < a.n_samples)
    train_sample<ET, RING_LDS, PROF>(a, smem, (int)blockIdx.x, s_ph, &s_last);
  if (a.C) elig_queue<ET, PROF>(a, smem, s_ph, &s_last);
  if constexpr (PROF) {
    __syncthreads();
    if (tid < kTrainPhases) a.prof[(size_t)blockIdx.x * kTrainPhases + tid] = s_ph[tid];
  }
}

// ---------------------------------------------------------------------------
// plain path: grid-wide kernels, one sample at a time
// ---------------------------------------------------------------------------

// The plain kernels read the sample being run from ps.cur = {position b in
// the minibatch, dataset index, epoch, length T}, so one captured chain of
// some length serves every sample that is not longer: the step kernels of
// steps t >= T return at once.
__global__ void k_p_begin(PlainState ps, int b, int sample, int epoch, int T) {
  ps.cur[0] = b;
  ps.cur[1] = sample;
  ps.cur[2] = epoch;
  ps.cur[3] = T;
}

__global__ void k_p_inject(TrainArgs a, PlainState ps, int t) {
  if (t >= ps.cur[3]) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int sample = ps.cur[1];
  if (a.ev_base) {
    const int e0 = a.ev_base[sample];
    const int k = a.ev_step_ptr[e0 + t] + q;
    if (k < a.ev_step_ptr[e0 + t + 1]) ps.ext[a.ev_nid[k]] = a.ev_amp[k];
  } else if (q < a.n_inputs) {
    const int j = a.inputs[q];
    const float x = input_drive(a, j, sample, t, ps.cur[2]);
    if (x != 0.0f) ps.ext[j] = x;
  }
}

__global__ void k_p_lif(TrainArgs a, PlainState ps, int t, int tm) {
  if (t >= ps.cur[3]) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  bool f = false;
  if (j < a.n) {
    float v = ps.v[j];
    int r = ps.refr[j];
    float psi;
    float* slot = ps.ring + (size_t)tm * a.n + j;
    f = train_lif(a, a.role[j], v, r, ps.ext[j], *slot, psi);
    ps.v[j] = v;
    ps.refr[j] = r;
    ps.psi[j] = psi;
    ps.ext[j] = 0.0f;
    *slot = 0.0f;
    if (f) {
      if (a.neur_fire) atomicAdd(a.neur_fire + j, 1ull);
      if (!a.readout_leaky && a.role[j] == SNC_ROLE_OUTPUT && a.channel[j] >= 0)
        atomicAdd(&ps.logit[a.channel[j]], 1.0f);
    }
    if (a.readout_leaky && a.role[j] == SNC_ROLE_OUTPUT && a.channel[j] >= 0)
      ps.logit[a.channel[j]] = __fadd_rn(ps.logit[a.channel[j]], v);
  }
  const uint64_t m = ballot(f);
  if (lane_id() == 0 && j - lane_id() < a.n) {
    const int w0 = (j - lane_id()) >> 5;
    ps.fw[w0] = (uint32_t)m;
    if (w0 + 1 < a.words) ps.fw[w0 + 1] = (uint32_t)(m >> 32);
  }
  if (m) {
    int pos = 0;
    if (lane_id() == 0) pos = atomicAdd(ps.nf, __popcll(m));
    pos = __shfl(pos, 0, kWave);
    if (f) ps.fired[pos + rank_below(m)] = j;
  }
}

__global__ void k_p_trace(TrainArgs a, PlainState ps, int t) {
  if (t >= ps.cur[3]) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < a.n) ps.tr[j] = trace_step(a.lif.decay, ps.tr[j], (ps.fw[j >> 5] >> (j & 31)) & 1u);
  if (a.raster && j < a.words)
    a.raster[((size_t)ps.cur[0] * a.T_max + t) * a.words + j] = ps.fw[j];
}

template <typename ET>
__global__ void k_p_elig(TrainArgs a, PlainState ps, int t) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ps.cur[3] || s >= a.S) return;
  const float tp = ps.tr[a.src[s]];
  if (tp != 0.0f) elig_add<ET>((ET*)ps.E, s, ps.psi[a.post[s]], tp);
}

__global__ void k_p_deliver(TrainArgs a, PlainState ps, int t) {
  if (t >= ps.cur[3]) return;
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int wave = gtid >> 6, lane = lane_id(), nw = (gridDim.x * blockDim.x) >> 6;
  const int nf = *ps.nf;
  unsigned long long ev = 0;
  for (int k = wave; k < nf; k += nw) {
    const int i = ps.fired[k];
    const int rb = a.row_ptr[i], re = a.row_ptr[i + 1];
    for (int e = rb + lane; e < re; e += kWave) {
      const int p = a.cpos[e];
      atomicOr(ps.bits + (p >> 5), 1u << (p & 31));
    }
    ev += (unsigned long long)(re - rb);
  }
  if (lane == 0 && ev) atomicAdd(ps.events, ev);
}

__global__ void k_p_fold(TrainArgs a, PlainState ps, int t, int tm) {
  if (t >= ps.cur[3]) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0) {
    *ps.spikes += (unsigned long long)*ps.nf;
    *ps.nf = 0;  // the fold runs after delivery consumed the list
  }
  if (j >= a.n) return;
  fold_light(j, a.n, a.L, tm, a.ccol, a.wc, a.cdel, ps.bits, ps.ring);
}

__global__ __launch_bounds__(kTrainThreads) void k_p_finish(TrainArgs a, PlainState ps) {
  __shared__ float s_delta[64];
  __shared__ float s_logit[64];
  const int tid = threadIdx.x;
  const int b = ps.cur[0];
  const int T = ps.cur[3];
  if (tid < a.classes)
    s_logit[tid] = a.readout_leaky ? __fdiv_rn(ps.logit[tid], (float)max(1, T)) : ps.logit[tid];
  __syncthreads();
  sample_epilogue(a, b, s_logit, ps.L, s_delta, a.labels[ps.cur[1]], a.n, tid, kTrainThreads);
}

// Contributions in unit order (q: each post's synapses in column order; the
// synapse is ecol_syn[q]), the layout the batched kernel stores.
template <typename ET>
__global__ void k_p_contrib(TrainArgs a, PlainState ps) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.S) return;
  const int s = a.ecol_syn[q];
  const float scale = __fdiv_rn(a.lr, (float)max(1, ps.cur[3]));
  const float e = (float)((const ET*)ps.E)[s];
  a.C[(size_t)ps.cur[0] * a.S + q] = a.reward ? e : __fmul_rn(__fmul_rn(scale, ps.L[a.post[s]]), e);
}

// ---------------------------------------------------------------------------
// reward mode (trainer.cpp:165-179)
// ---------------------------------------------------------------------------

// One block. Each sample's softmax p (the epilogue's arithmetic), its action
// a ~ p by the first k with u_b <= cdf_k (cdf summed in float, the last class
// when none), and its reward R = 1[a == label]; then one lane walks the
// minibatch in order: adv_b = R_b - baseline, baseline <- decay * baseline +
// (1 - decay) * R_b; then delta_b[k] = adv_b * (p_b[k] - onehot(a_b)[k]).
__global__ __launch_bounds__(256) void k_reward_delta(TrainArgs a, RewardArgs r, int B) {
  const int K = a.classes, tid = threadIdx.x;
  for (int b = tid; b < B; b += blockDim.x) {
    const float* lg = a.logits + (size_t)b * K;
    float* p = r.delta + (size_t)b * K;  // p until the last phase
    float mx = lg[0];
    for (int k = 1; k < K; ++k) mx = fmaxf(mx, lg[k]);
    double sum = 0.0;
    for (int k = 0; k < K; ++k) sum += (double)expf(__fsub_rn(lg[k], mx));
    const float u = r.u[b];
    float cdf = 0.0f;
    int act = -1;
    for (int k = 0; k < K; ++k) {
      p[k] = __fdiv_rn(expf(__fsub_rn(lg[k], mx)), (float)sum);
      cdf = __fadd_rn(cdf, p[k]);
      if (act < 0 && u <= cdf) act = k;
    }
    r.act[b] = act < 0 ? K - 1 : act;
  }
  __syncthreads();
  if (tid == 0) {
    float base = *r.base;
    const float keep = __fsub_rn(1.0f, r.decay);
    for (int b = 0; b < B; ++b) {
      const float R = r.act[b] == a.labels[a.idx[b]] ? 1.0f : 0.0f;
      r.adv[b] = __fsub_rn(R, base);
      base = __fadd_rn(__fmul_rn(r.decay, base), __fmul_rn(keep, R));
    }
    *r.base = base;
  }
  __syncthreads();
  for (int q = tid; q < B * K; q += blockDim.x) {
    const int b = q / K, k = q - b * K;
    const float p = r.delta[q];
    r.delta[q] = __fmul_rn(r.adv[b], __fsub_rn(p, k == r.act[b] ? 1.0f : 0.0f));
  }
}

// Learning signals of every sample: Ls[b][j] from delta_b.
__global__ void k_reward_signal(TrainArgs a, RewardArgs r) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (j >= a.n) return;
  r.Ls[(size_t)b * a.n + j] = learning_signal(a, r.delta + (size_t)b * a.classes, j);
}

// C_{b,s} = (lr / T_b * L_b[post_s]) * float(E_{b,s}), in place over float(E)
// (unit order: entry q is synapse ecol_syn[q]).
__global__ void k_reward_contrib(TrainArgs a, RewardArgs r) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (q >= a.S) return;
  const int T = a.lengths[a.idx[b]];
  const float scale = __fdiv_rn(a.lr, (float)max(1, T));
  float* c = a.C + (size_t)b * a.S + q;
  *c = __fmul_rn(__fmul_rn(scale, r.Ls[(size_t)b * a.n + a.post[a.ecol_syn[q]]]), *c);
}

// ---------------------------------------------------------------------------
// update
// ---------------------------------------------------------------------------

// w_s <- clamp(w_s - (sum_b C_{b,s}) / B, +-w_max), b ascending; C in unit
// order, so entry q updates synapse ecol_syn[q] (each synapse exactly once).
__global__ void k_apply(TrainArgs a, float* w, int B) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.S) return;
  float acc = 0.0f;
  for (int b = 0; b < B; ++b) acc = __fadd_rn(acc, a.C[(size_t)b * a.S + q]);
  const int s = a.ecol_syn[q];
  const float nw = __fsub_rn(w[s], __fdiv_rn(acc, (float)B));
  w[s] = fminf(fmaxf(nw, -a.w_max), a.w_max);
}

// Trainer weights in padded column order for the ordered fold.
__global__ void k_gather_wc(const float* __restrict__ w, const int32_t* __restrict__ csc_e,
                            float* __restrict__ wc, int64_t P) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int e = csc_e[p];
  wc[p] = e >= 0 ? w[e] : 0.0f;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline int blocks_for(int64_t n, int t) { return (int)std::max<int64_t>(1, (n + t - 1) / t); }

size_t train_lds_bytes(const TrainArgs& a, bool ring_lds) {
  size_t b = train_state_bytes(a.n, a.words, a.bits_words, a.classes, train_rand_words(a)) +
             train_const_bytes(a.n, a.n_inputs);
  if (ring_lds) b += sizeof(float) * (size_t)a.L * a.n;
  else if (a.arq) b += train_in_bytes(a.n);
  return b;
}

int train_launch_batched(const TrainArgs& a, int B, int grid, bool f64, bool ring_lds,
                         size_t lds_max, hipStream_t st) {
  const bool prof = a.prof != nullptr;
  const void* fn =
      prof ? (f64 ? (ring_lds ? (const void*)k_train_batched<double, true, true>
                              : (const void*)k_train_batched<double, false, true>)
                  : (ring_lds ? (const void*)k_train_batched<float, true, true>
                              : (const void*)k_train_batched<float, false, true>))
           : (f64 ? (ring_lds ? (const void*)k_train_batched<double, true, false>
                              : (const void*)k_train_batched<double, false, false>)
                  : (ring_lds ? (const void*)k_train_batched<float, true, false>
                              : (const void*)k_train_batched<float, false, false>));
  // The kernel's own static LDS counts against the same per-block limit.
  hipFuncAttributes attr{};
  const hipError_t eg = hipFuncGetAttributes(&attr, fn);
  if (eg != hipSuccess) return hip_fail(eg, "hipFuncGetAttributes of the batched trainer");
  if (attr.sharedSizeBytes >= lds_max) return fail(SNC_E_UNSUPPORTED, "no dynamic LDS left");
  const size_t dyn_max = lds_max - attr.sharedSizeBytes;
  // The step loop's state (+ ring); after it, the learning signal and two
  // transposed history chunks of at least 4 and ideally 28 steps.
  size_t lds = std::max(train_lds_bytes(a, ring_lds), train_elig_lds_bytes(a.n, 4));
  lds = std::max(lds, std::min(dyn_max, train_elig_lds_bytes(a.n, 28)));
  // Room to stage every arrival of a step when all synapses deliver at once
  // (the kernel falls back to HBM gathers for a step that does not fit).
  const size_t fold_full = (train_lds_bytes(a, ring_lds) + 15) / 16 * 16 +
                           (sizeof(uint16_t) * ((size_t)a.bits_words + 1) + 15) / 16 * 16 +
                           sizeof(uint2) * std::min<size_t>((size_t)a.S, 65535);
  lds = std::max(lds, std::min(dyn_max, fold_full));
  if (lds > dyn_max) return fail(SNC_E_UNSUPPORTED, "batched trainer state exceeds LDS");
  const hipError_t ea = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (ea != hipSuccess) {
    hip_fail(ea, "LDS request of the batched trainer");
    return SNC_E_UNSUPPORTED;
  }
  TrainArgs aa = a;
  aa.lds_bytes = (int)lds;
  aa.n_samples = B;
  void* args[] = {(void*)&aa};
  const hipError_t e = hipLaunchKernel(fn, dim3(std::max(grid, B)), dim3(kTrainThreads), args,
                                       lds, st);
  if (e != hipSuccess) return hip_fail(e, "batched trainer launch");
  return SNC_OK;
}

int train_launch_plain_begin(const PlainState& ps, int b, int sample, int epoch, int T,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_p_begin, dim3(1), dim3(1), 0, st, ps, b, sample, epoch, T);
  return hip_last("plain begin launch failed");
}

int train_launch_plain_sample(const TrainArgs& a, const PlainState& ps, int T, bool f64,
                              hipStream_t st) {
  const int n = a.n, L = a.L;
  const int gn = blocks_for(n, 256), gs = blocks_for(a.S, 256);
  const int gi = blocks_for(std::max<int64_t>(a.n_inputs, a.max_inj_per_step), 256);
  int tm = 0;
  for (int t = 0; t < T; ++t) {
    hipLaunchKernelGGL(k_p_inject, dim3(gi), dim3(256), 0, st, a, ps, t);
    hipLaunchKernelGGL(k_p_lif, dim3(gn), dim3(256), 0, st, a, ps, t, tm);
    hipLaunchKernelGGL(k_p_trace, dim3(std::max(gn, blocks_for(a.words, 256))), dim3(256), 0, st,
                       a, ps, t);
    if (f64)
      hipLaunchKernelGGL(k_p_elig<double>, dim3(gs), dim3(256), 0, st, a, ps, t);
    else
      hipLaunchKernelGGL(k_p_elig<float>, dim3(gs), dim3(256), 0, st, a, ps, t);
    hipLaunchKernelGGL(k_p_deliver, dim3(std::min(gn, 1024)), dim3(256), 0, st, a, ps, t);
    hipLaunchKernelGGL(k_p_fold, dim3(gn), dim3(256), 0, st, a, ps, t, tm);
    tm = tm + 1 == L ? 0 : tm + 1;
  }
  hipLaunchKernelGGL(k_p_finish, dim3(1), dim3(kTrainThreads), 0, st, a, ps);
  if (a.C) {
    if (f64)
      hipLaunchKernelGGL(k_p_contrib<double>, dim3(gs), dim3(256), 0, st, a, ps);
    else
      hipLaunchKernelGGL(k_p_contrib<float>, dim3(gs), dim3(256), 0, st, a, ps);
  }
  return hip_last("plain trainer launch failed");
}

int train_launch_reward(const TrainArgs& a, const RewardArgs& r, int B, hipStream_t st) {
  hipLaunchKernelGGL(k_reward_delta, dim3(1), dim3(256), 0, st, a, r, B);
  hipLaunchKernelGGL(k_reward_signal, dim3(blocks_for(a.n, 256), B), dim3(256), 0, st, a, r);
  hipLaunchKernelGGL(k_reward_contrib, dim3(blocks_for(a.S, 256), B), dim3(256), 0, st, a, r);
  return hip_last("reward launch failed");
}

int train_launch_apply(const TrainArgs& a, float* w, int B, const int32_t* csc_e, float* wc,
                       int64_t P, hipStream_t st) {
  hipLaunchKernelGGL(k_apply, dim3(blocks_for(a.S, 256)), dim3(256), 0, st, a, w, B);
  hipLaunchKernelGGL(k_gather_wc, dim3(blocks_for(P, 256)), dim3(256), 0, st, w, csc_e, wc, P);
  return hip_last("update launch failed");
}

int train_launch_gather(const float* w, const int32_t* csc_e, float* wc, int64_t P,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_gather_wc, dim3(blocks_for(P, 256)), dim3(256), 0, st, w, csc_e, wc, P);
  return hip_last("gather launch failed");
}

}  // namespace snc_amd
