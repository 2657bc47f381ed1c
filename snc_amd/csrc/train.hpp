// e-prop training: kernel argument blocks and launchers (train.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "forward_plan.hpp"
#include "snc_hip.h"

namespace snc_amd {

struct TrainArgs {
  int n, S, L, classes, words, bits_words, n_inputs, dim;
  int readout_leaky;
  float readout_psi;
  LifK lif;
  float gamma, lr, w_max;
  int train_hidden;
  int reward;  // reward mode: C receives float(E) and the reward kernels scale it
  int encoder;
  float gain, max_rate, enc_thr;
  uint32_t seed_lo, seed_hi;
  int epoch, T_max;
  int64_t max_inj_per_step;
  // graph
  const int32_t* row_ptr;
  const int32_t* post;
  const int32_t* src;  // pre of each synapse (CSR order)
  const int32_t* cpos;
  const int32_t* ccol;
  const uint8_t* cdel;
  const float* wc;     // trainer weights in padded column order
  const float* w_csr;  // trainer weights, CSR order
  const uint8_t* sdel; // delays, CSR order
  const uint8_t* role;
  const int32_t* channel;
  const int32_t* internal_row;
  const float* Bfb;    // [n_internal][classes]
  const int32_t* inputs;  // INPUT neuron ids
  // dataset
  const float* X;          // [n_data][dim] or null
  const int32_t* labels;   // [n_data]
  const int32_t* lengths;  // [n_data]
  const int32_t* ev_base;      // [n_data] offset into ev_step_ptr, or null (features)
  const int32_t* ev_step_ptr;  // per (sample, step) injection ranges
  const int32_t* ev_nid;
  const float* ev_amp;
  // batch
  const int32_t* idx;  // [B] dataset indices
  float* hist;         // [B][T_max][2][hn]: psi then trace of every neuron, per step
  int hn;              // history row length: n rounded up to a multiple of 4
  int lds_bytes;       // dynamic LDS of the batched launch
  float* C;            // [B][S] contributions in unit order (entry q: synapse ecol_syn[q]), or null
  float* ring_g;       // [B][L][n] when the ring is not in LDS and no arrival queues
  // Arrival queues (the batched kernel when the ring does not fit in LDS):
  // per sample and ring slot, arq_cap records {column position, weight bits}
  // of the deposits that arrive at that slot's step; null: ring_g instead.
  uint2* arq;
  int64_t arq_cap;
  float* logits;       // [B][classes]
  double* loss;        // [B]
  int32_t* correct;    // [B]
  unsigned long long* blk_spikes;  // [B]
  unsigned long long* blk_events;  // [B]
  unsigned long long* neur_fire;   // [n] or null
  uint32_t* raster;                // [B][T_max][words] or null
  unsigned long long* prof;        // [B][kTrainPhases] wall-clock ticks per phase, or null
  // eligibility work units: {post << 5 | (count - 1), first index into ecol_*},
  // each up to kEligPerThread synapses of one post in column order
  const int2* eunit;
  int n_units;
  const int32_t* ecol_pre;  // [S] pre of each synapse, unit order
  const int32_t* ecol_syn;  // [S] its CSR position
  // eligibility work queue of the batched launch (see claim_item)
  int n_samples;            // B: workgroups 0 .. B-1 run one sample each
  int rounds;               // work items per sample: ceil(n_units / kTrainThreads)
  long long wait_ticks;     // wall-clock ticks a workgroup polls for unpublished work
  int* wq;                  // [2B]: published flag, then claimed rounds, per sample
  float* Ls_g;              // [B][n] learning signals of each sample
};

// Largest arrival-queue storage a training session allocates.
constexpr size_t kArqBytesMax = (size_t)32 << 30;

// Threads of a batched-trainer workgroup (one sample per workgroup).
constexpr int kTrainThreads = 1024;

// Synapses of one post that one lane's eligibility unit holds in registers.
constexpr int kEligPerThread = 16;

// Phases of the batched kernel timed by its profiled instantiation:
// 0 set-up, 1 LIF + trace + histories, 2 delivery bits, 3 bitmap prefix scan,
// 4 arrival staging, 5 ordered fold + next injections, 6 end-of-sample
// epilogue, 7 eligibility sums, 8 history chunks streamed into LDS,
// 9 contribution stores, 10 claiming (and waiting for) eligibility work.
constexpr int kTrainPhases = 11;

// Device state of the plain path (one sample at a time).
struct PlainState {
  float *v, *tr, *psi, *ext, *ring, *logit, *L;
  int32_t *refr, *fired, *nf;
  int32_t* cur;  // [4]: position in the minibatch, dataset index, epoch, length
  uint32_t *fw, *bits;
  void* E;  // [S]
  unsigned long long *spikes, *events;
};

// LDS bytes of the batched kernel's per-sample state (membrane, trace,
// injection, refractory, fired list, fired bitmap, buckets, logits, delta),
// rounded up to 16; the ring (when in LDS) and later the history chunks
// start there.
__host__ __device__ inline size_t train_state_bytes(int n, int words, int bits_words,
                                                    int classes, int rand_words) {
  const size_t b = sizeof(float) * 3 * (size_t)n + sizeof(int) * 2 * (size_t)n +
                   sizeof(uint32_t) * ((size_t)words + bits_words + rand_words) +
                   sizeof(float) * 2 * (size_t)classes;
  return (b + 15) / 16 * 16;
}

// Philox words the batched kernel stages per step (feature inputs with the
// Poisson encoder: one word per channel, channels rounded up to 4), else 0.
__host__ __device__ inline int train_rand_words(const TrainArgs& a) {
  return !a.ev_base && a.X && a.encoder == SNC_ENC_POISSON ? (a.dim + 3) / 4 * 4 : 0;
}

// LDS bytes of the per-sample copies of read-only graph data the step loop
// touches every step (row pointers, column starts, channels, roles, input
// drive), rounded up to 16; they follow the state.
__host__ __device__ inline size_t train_const_bytes(int n, int n_inputs) {
  const size_t b = sizeof(int32_t) * (2 * ((size_t)n + 1) + (size_t)n + (size_t)n_inputs) +
                   sizeof(float) * (size_t)n_inputs + (size_t)n;
  return (b + 15) / 16 * 16;
}

// Reward-mode state of a session (device buffers).
struct RewardArgs {
  const float* u;  // [B] action draws of this minibatch, in minibatch order
  float* base;     // [1] reward baseline, carried across minibatches
  float* delta;    // [B][classes]
  float* Ls;       // [B][n] learning signals
  int32_t* act;    // [B] sampled action
  float* adv;      // [B] reward advantage
  float decay;
};

// LDS bytes of the queue mode's per-target arrival sums (n floats, rounded
// up to 16), where the ring would start.
__host__ __device__ inline size_t train_in_bytes(int n) {
  return (sizeof(float) * (size_t)n + 15) / 16 * 16;
}

// LDS of the update phase: the learning signal (n floats, rounded up to 16
// bytes), then the transposed psi and trace chunks of `steps` steps.
__host__ __device__ inline size_t train_elig_lds_bytes(int n, int steps) {
  return (sizeof(float) * (size_t)n + 15) / 16 * 16 + sizeof(float) * 2 * (size_t)n * steps;
}

size_t train_lds_bytes(const TrainArgs& a, bool ring_lds);
// Workgroups of a batched launch of B samples: B, plus (when contributions
// are computed) helpers on the CUs the samples leave free, no more than there
// are eligibility items beyond one per sample.
inline int train_batched_grid(const TrainArgs& a, int B, int n_cu) {
  if (!a.C) return B;
  const int helpers = std::max(0, std::min(n_cu - B, B * a.rounds - B));
  return B + helpers;
}
// a.wq must be zero (2 * B ints) when a.C is set.
int train_launch_batched(const TrainArgs& a, int B, int grid, bool f64, bool ring_lds,
                         size_t lds_max, hipStream_t st);
int train_launch_plain_begin(const PlainState& ps, int b, int sample, int epoch, int T,
                             hipStream_t st);
// The step kernels of T steps and the end-of-sample kernels. The sample's own
// length (set by train_launch_plain_begin) may be shorter: its step kernels
// of later steps return at once.
int train_launch_plain_sample(const TrainArgs& a, const PlainState& ps, int T, bool f64,
                              hipStream_t st);
// Reward mode, after every sample of the minibatch has run: actions,
// advantages and the baseline (one ordered scan), learning signals, and the
// contributions C_{b,s} = (lr / T_b * L_b[post_s]) * float(E_{b,s}).
int train_launch_reward(const TrainArgs& a, const RewardArgs& r, int B, hipStream_t st);
int train_launch_apply(const TrainArgs& a, float* w, int B, const int32_t* csc_e, float* wc,
                       int64_t P, hipStream_t st);
int train_launch_gather(const float* w, const int32_t* csc_e, float* wc, int64_t P,
                        hipStream_t st);

}  // namespace snc_amd
