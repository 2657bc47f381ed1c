/*
 * snc_hip.h -- C ABI of the MI355X (gfx950) backend for SNC's event-driven
 * spiking path.
 *
 * Plain C: POD structs, raw pointers and sizes, integer status codes. No C++
 * types, exceptions or torch types cross this boundary. Every entry point is
 * synchronous with respect to the host unless stated otherwise and returns
 * SNC_OK (0) on success or a negative SNC_E_* code; snc_hip_last_error()
 * returns a human-readable message for the last failure on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   snc_hip_available       <- snc::cuda::available()       include/snc/runtime.hpp:61-64,
 *                                                             src/runtime/cuda_backend.cu:181-184
 *   snc_hip_forward         <- snc::cuda::forward(...)       include/snc/runtime.hpp:65-69,
 *                                                             src/runtime/cuda_backend.cu:186-361
 *   snc_hip_plan_*          <- the per-call upload + step loop of cuda::forward, split so the
 *                              graph and inputs stay resident in HBM across calls
 *                              (cuda_backend.cu:236-336 re-uploads on every call)
 *   snc_hip_train_*         <- snc::cudatrain::* session     include/snc/cuda_trainer.hpp:24-57
 *   snc_hip_mlp_*           <- no counterpart: learned grow proposal that plugs into
 *                              Connectome::structural_update's grow step
 *                              (src/structure/connectome.cpp:164-211)
 *
 * Not thread-safe: one host thread per plan/session, as in the reference.
 */
#ifndef SNC_HIP_H_
#define SNC_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define SNC_OK 0
#define SNC_E_INVALID (-1)     /* bad argument / graph fails validation      */
#define SNC_E_NODEVICE (-2)    /* no usable gfx950 device                     */
#define SNC_E_HIP (-3)         /* a HIP runtime call failed                   */
#define SNC_E_NOMEM (-4)       /* device allocation failed                    */
#define SNC_E_UNSUPPORTED (-5) /* variant / config not supported by this build */

/* ---- delivery variants (values follow snc::Backend, runtime.hpp:21-27,
 *      extended with the hip-* strategies) ---------------------------------- */
#define SNC_VARIANT_ATOMIC 5 /* hip-atomic: float atomics into the ring      */
#define SNC_VARIANT_BUCKET 6 /* hip-bucket: per-target ordered fold, bit-exact */
#define SNC_VARIANT_SORT 7   /* hip-sort: per-segment counting sort of each step's events */
#define SNC_VARIANT_AUTO 8   /* hip-auto: autotuner picks per graph           */
#define SNC_VARIANT_AUTO_EXACT 9 /* hip-auto-exact: autotuner over the bit-exact
                                    strategies (bucket, sort; pull for
                                    batched plans) only                       */
#define SNC_VARIANT_PULL 10  /* hip-pull: batched gather over the incoming
                                columns, all samples of a plan (up to 64) in
                                one pass, bit-exact (hub targets walk their
                                chunks' arrival lists in column order) */

/* ---- graph roles (GraphRole, snn_graph.hpp:29) ---------------------------- */
#define SNC_ROLE_INTERNAL 0
#define SNC_ROLE_INPUT 1
#define SNC_ROLE_OUTPUT 2

/* ---- reset modes --------------------------------------------------------- */
#define SNC_RESET_TO_VALUE 0 /* v = reset on a spike (cpu_backend.cpp:135-138)  */
#define SNC_RESET_SUBTRACT 1 /* v -= threshold on a spike; v held while refractory */

/* Read-only view over SNNGraph's arrays (snn_graph.hpp:43-77). CSR over
 * OUTGOING synapses, rows sorted by pre id. */
typedef struct snc_graph_view {
  int32_t num_neurons;
  int32_t num_synapses;
  const int32_t* row_ptr;  /* [N+1]                                     */
  const int32_t* post_ids; /* [S]                                       */
  const float* weights;    /* [S]  (signed values are allowed)          */
  const int32_t* delays;   /* [S]  >= 1                                 */
  const int8_t* sign;      /* [N]  +1 / -1 (applied on the pre)         */
  const uint8_t* role;     /* [N]  SNC_ROLE_*                           */
  const int32_t* channel;  /* [N]  input/output channel or -1           */
} snc_graph_view;

/* LIFParams (runtime.hpp:34-39) plus the reset mode. */
typedef struct snc_lif {
  float threshold;
  float decay;
  float reset;
  int32_t refractory;
  int32_t reset_mode; /* SNC_RESET_* */
} snc_lif;

/* InputEvent (encoders.hpp:21-26), identical layout. */
typedef struct snc_event {
  int32_t sample_id;
  int32_t channel_id;
  int32_t time;
  float amplitude;
} snc_event;

/* RunStats (runtime.hpp:41-46) plus the variant that actually ran. */
typedef struct snc_forward_stats {
  int64_t spikes;
  int64_t synaptic_events;
  int32_t overflow; /* events dropped for lack of buffer space (always 0 here) */
  int32_t used_variant;
  double ms; /* device time of the step loop only (HIP events) */
} snc_forward_stats;

/* Optional outputs of a forward run; any pointer may be NULL. */
typedef struct snc_forward_capture {
  float* logits;               /* [num_output_channels] summed output spikes  */
  int32_t* neuron_spikes;      /* [N] per-neuron spike counts                 */
  uint32_t* raster;            /* [num_steps][ceil(N/32)] fired bitmap        */
  float* final_v;              /* [N] membrane after the last step            */
  int64_t* step_spikes;        /* [num_steps] spikes per step                 */
} snc_forward_capture;

typedef struct snc_hip_graph snc_hip_graph; /* device-resident graph      */
typedef struct snc_hip_plan snc_hip_plan;   /* graph + inputs + state      */

const char* snc_hip_last_error(void);

/* 1 when a gfx950 device is usable, else 0 (cuda_backend.cu:181-184). */
int snc_hip_available(void);

/* Upload a graph once; it stays resident until snc_hip_graph_free. The graph
 * is validated as SNNGraph::validate does (snn_graph.cpp:30-56). */
int snc_hip_graph_upload(const snc_graph_view* g, snc_hip_graph** out);
void snc_hip_graph_free(snc_hip_graph* g);

/* Static properties of an uploaded graph used by the autotuner. */
typedef struct snc_graph_profile {
  int32_t num_neurons;
  int64_t num_synapses;
  int32_t max_delay;
  int32_t ring_len;
  int32_t max_in_degree;
  int32_t max_out_degree;
  double mean_in_degree;
  double heavy_fraction;  /* fraction of synapses onto heavy targets */
  int32_t heavy_threshold; /* in-degree above which a target is heavy (hub) */
  int32_t num_heavy;       /* number of heavy targets */
} snc_graph_profile;
int snc_hip_graph_profile_get(const snc_hip_graph* g, snc_graph_profile* out);

/* Build a forward plan: the sample's events are resolved to per-step
 * injections (build_schedule, cpu_backend.cpp:47-65) and uploaded, and the
 * device state is allocated. `variant` is SNC_VARIANT_*: an explicit
 * strategy runs with its default launch configuration; SNC_VARIANT_AUTO /
 * SNC_VARIANT_AUTO_EXACT take the autotuner's configuration for this graph
 * (searched on the first such call, then cached with the graph). */
int snc_hip_plan_create(snc_hip_graph* g, const snc_event* events, int64_t n_events,
                        int32_t num_steps, const snc_lif* lif, int32_t variant,
                        snc_hip_plan** out);

/* Launch configuration of a forward plan: the autotuner's search space. */
typedef struct snc_plan_config {
  int32_t variant;         /* SNC_VARIANT_ATOMIC / _BUCKET / _SORT                 */
  int32_t block_neurons;   /* neurons per block of the step kernel: 128, 256, 512  */
  int32_t heavy_threshold; /* hip-bucket: targets with in-degree above this are
                              hubs stepped by their own block (bucket capacity of
                              the per-lane fold); at most 512                       */
  int32_t persistent;      /* 1: the whole run as one cooperative launch
                              (hip-atomic / hip-bucket)                            */
} snc_plan_config;
int snc_hip_plan_create_config(snc_hip_graph* g, const snc_event* events, int64_t n_events,
                               int32_t num_steps, const snc_lif* lif,
                               const snc_plan_config* cfg, snc_hip_plan** out);
int snc_hip_plan_get_config(const snc_hip_plan* p, snc_plan_config* out);

/* The autotuner's search for a graph: the candidates it derived from the
 * graph's in-degree statistics, the measured device time per step of each on
 * a trial window of the first trial_steps steps of the input that triggered
 * the search, and the one chosen (the fastest). exact_only selects the
 * SNC_VARIANT_AUTO_EXACT search, batched the search of batched plans (kept
 * apart from single-sample plans; the latest one, as a batched search is
 * redone when a plan has another sample count). Fails when no such search
 * has run yet. */
#define SNC_TUNE_MAX_CANDIDATES 24
#define SNC_TUNE_ERROR_LEN 120
typedef struct snc_tune_report {
  int32_t n_candidates;
  int32_t chosen;
  int32_t trial_steps;
  double search_ms; /* wall time of the whole search */
  snc_plan_config cand[SNC_TUNE_MAX_CANDIDATES];
  double ms_per_step[SNC_TUNE_MAX_CANDIDATES]; /* -1: the candidate could not run */
  int32_t status[SNC_TUNE_MAX_CANDIDATES];     /* SNC_OK or the SNC_E_* it failed with */
  char error[SNC_TUNE_MAX_CANDIDATES][SNC_TUNE_ERROR_LEN]; /* its error text, or "" */
} snc_tune_report;
int snc_hip_graph_tune_report(const snc_hip_graph* g, int32_t exact_only, int32_t batched,
                              snc_tune_report* out);

/* Batched forward: num_samples independent samples of the same graph in one
 * plan, stepped together (one launch per step covers every sample). An event
 * belongs to sample event.sample_id in [0, num_samples); others are ignored.
 * No reference counterpart: snc_bench evaluates samples one forward() at a
 * time (src/bench/snc_bench.cpp:185-211); each sample's results equal that
 * forward's. cfg (may be NULL) fixes the launch configuration, else variant
 * selects it as in snc_hip_plan_create (hip-sort and persistent launches are
 * single-sample only). Capture arrays of a batched run carry a leading sample
 * dimension: logits [B][channels], neuron_spikes and final_v [B][N], raster
 * [B][T][ceil(N/32)], step_spikes [B][T]; stats are totals over the samples. */
int snc_hip_plan_create_batch(snc_hip_graph* g, const snc_event* events, int64_t n_events,
                              int32_t num_samples, int32_t num_steps, const snc_lif* lif,
                              int32_t variant, const snc_plan_config* cfg, snc_hip_plan** out);
int32_t snc_hip_plan_num_samples(const snc_hip_plan* p);
/* Spikes and synaptic events of each sample ([num_samples] each) in the last run. */
int snc_hip_plan_sample_stats(const snc_hip_plan* p, int64_t* spikes, int64_t* events);
/* Run the plan from a zero state. Inputs are already resident; stats->ms is
 * the device time of the step loop. */
int snc_hip_plan_run(snc_hip_plan* p, snc_forward_capture* cap, snc_forward_stats* stats);
/* Same as plan_run but without synchronising or reading results back: the
 * caller synchronises the device itself. For throughput timing. */
int snc_hip_plan_launch(snc_hip_plan* p);
/* The variant a plan resolved to (useful after SNC_VARIANT_AUTO). */
int32_t snc_hip_plan_variant(const snc_hip_plan* p);
/* Average device duration (ms) of the dominant kernel over the last run, and
 * the number of its launches. */
int snc_hip_plan_kernel_time(const snc_hip_plan* p, double* avg_ms, int64_t* launches);
void snc_hip_plan_free(snc_hip_plan* p);

/* Per-kernel device time of one instrumented run: eager launches with a HIP
 * event around every kernel on the plan's stream. kernel_ms_total[k] and
 * launches[k] are filled for each class k = SNC_KERNEL_*:
 *   HUB_STEP    kept for compatibility; hub targets are stepped inside k_step
 *               (its leading blocks), so this class reports no launches
 *   STEP        k_step: fold, integrate/fire and expansion of every neuron,
 *               one launch per step (hip-atomic and hip-bucket)
 *   SORT_STEP   the whole kernel chain of one hip-sort step
 *   PERSISTENT  k_persist, the whole run in one cooperative launch */
#define SNC_KERNEL_HUB_STEP 0
#define SNC_KERNEL_STEP 1
#define SNC_KERNEL_SORT_STEP 2
#define SNC_KERNEL_PERSISTENT 3
#define SNC_KERNEL_CLASSES 4
int snc_hip_plan_profile(snc_hip_plan* p, double* kernel_ms_total, int64_t* launches);

/* Wait until all work the library queued on the device has finished. */
int snc_hip_synchronize(void);

/* Device memory in bytes held by a graph (its arrays, plus hip-pull's light
 * columns and halos once a hip-pull plan built them) and by a plan (state, schedule,
 * scratch, the raster once captured). No counterpart in the reference: the
 * memory fields its observability plan asks for (docs/new-plan.md:727-741). */
int snc_hip_graph_memory(const snc_hip_graph* g, int64_t* bytes);
int snc_hip_plan_memory(const snc_hip_plan* p, int64_t* bytes);

/* Device ceiling for scattered delivery (no counterpart in the reference: a
 * measurement the benchmark's roofline uses). Times `reps` graph-launched
 * launches, each doing `ops` accesses of one kind at hashed addresses over a
 * fresh buffer of target_bytes, from `blocks` blocks of 512 threads (the
 * fused step kernel's shape), and returns the best average microseconds per
 * launch. The step kernel's delivery does one such access per synaptic event,
 * so this is the time floor of a step with that many events. */
#define SNC_SCATTER_ATOMIC_ADD 0 /* float atomicAdd (hip-atomic delivery)    */
#define SNC_SCATTER_ATOMIC_OR 1  /* u32 atomicOr (hip-bucket delivery)       */
#define SNC_SCATTER_STORE 2      /* plain 4-byte store                       */
#define SNC_SCATTER_LOAD 3       /* plain 4-byte load (hip-bucket's gathers) */
int snc_hip_measure_scatter(int32_t kind, int64_t target_bytes, int64_t ops, int32_t blocks,
                            int32_t reps, double* us_per_launch);

/* The hip-atomic step kernel's delivery alone, for one step's fired rows (no
 * counterpart in the reference; a benchmark measurement): replays the
 * expansion of rows `fired` of graph g exactly as the fused step kernel
 * places it (the rows of each 512-neuron block on that block's waves, one
 * wave per row, lanes striding the row) into a scratch ring with float
 * atomics, graph-launched `reps` times; returns the best average microseconds
 * per launch -- the time floor of a step with these events. */
int snc_hip_measure_expansion(const snc_hip_graph* g, const int32_t* fired, int32_t n_fired,
                              int32_t reps, double* us_per_launch);

/* One-shot forward, the shape of cuda::forward: upload, run, read back, free. */
int snc_hip_forward(const snc_graph_view* g, const snc_event* events, int64_t n_events,
                    int32_t num_steps, const snc_lif* lif, int32_t variant,
                    snc_forward_capture* cap, snc_forward_stats* stats);

/* ---- e-prop training (cudatrain::*, include/snc/cuda_trainer.hpp:24-57) ---- */

#define SNC_READOUT_SPIKING 0 /* logits = output spike counts (trainer.cpp:122-123)   */
#define SNC_READOUT_LEAKY 1   /* outputs never fire; logits = mean membrane over T_b   */

#define SNC_ENC_DIRECT 0
#define SNC_ENC_POISSON 1 /* on-device Philox Bernoulli, see snc_hip_train_batch */
#define SNC_ENC_LATENCY 2 /* one spike of `gain` per channel at step
                             lround((T_i - 1) * (1 - min(1, x))), encoders.cpp:56-66,
                             T_i the sample's length */

/* TrainConfig (include/snc/trainer.hpp:30-48) plus the readout mode. */
typedef struct snc_train_config {
  int32_t num_steps; /* T when per-sample lengths are not given              */
  snc_lif lif;
  float lr;
  float w_init;          /* weights start uniform [0, w_init] (trainer.cpp:15-18) */
  float w_max;
  float surrogate_scale; /* gamma of the triangular surrogate                */
  float feedback_scale;  /* sd of the random feedback matrix B               */
  int32_t train_hidden;
  int32_t readout_mode;  /* SNC_READOUT_*                                    */
  float readout_psi;     /* surrogate of a leaky readout neuron (constant)   */
  int32_t encoder;       /* SNC_ENC_* for feature inputs                     */
  float enc_gain;
  float enc_max_rate;
  float enc_threshold;
  uint64_t seed;
  /* Three-factor reward rule (trainer.cpp:165-179): per trained sample, in
   * minibatch order, an action a ~ softmax(logits) drawn with one
   * uniform_real_distribution<float> value of mt19937_64(seed ^ 0xA5A5C0FFEE),
   * reward R = 1[a == label], delta = (R - baseline) * (p - onehot(a)), then
   * baseline <- decay * baseline + (1 - decay) * R (baseline starts at
   * 1 / classes). 0 = the supervised delta p - onehot(label). */
  int32_t reward_mode;
  float reward_baseline_decay;
} snc_train_config;

#define SNC_TRAIN_PLAIN 0   /* one sample at a time, per-step kernel launches       */
#define SNC_TRAIN_BATCHED 1 /* whole minibatch in one persistent launch, one
                               workgroup per sample running its own T_b steps    */
#define SNC_TRAIN_SERIAL 2  /* one sample at a time, each one launch of the
                               persistent kernel as a minibatch of one (its
                               eligibility rounds on otherwise idle CUs); the
                               update is the batched one, so the weights are
                               bit-identical to SNC_TRAIN_BATCHED's. No
                               counterpart in the reference: it is the
                               per-sample baseline of the batched launch      */

#define SNC_ELIG_F32 0 /* eligibility accumulated in float (fused multiply-add)    */
#define SNC_ELIG_F64 1 /* in double, exactly the CPU trainer's sums (trainer.cpp:137) */

typedef struct snc_hip_train snc_hip_train;

/* cudatrain::create (cuda_trainer.cu:190-273): device session for a frozen
 * graph. Weights start uniform [0, w_init], then the feedback matrix
 * B ~ N(0, feedback_scale), both from one mt19937_64(seed) stream as the CPU
 * Trainer draws them (trainer.cpp:10-41). batch = largest minibatch;
 * max_steps = longest sample. elig = SNC_ELIG_*. */
int snc_hip_train_create(const snc_graph_view* g, const snc_train_config* cfg, int32_t batch,
                         int32_t max_steps, int32_t elig, snc_hip_train** out);
void snc_hip_train_destroy(snc_hip_train* s);

/* cudatrain::set_weights / get_weights (signed, CSR order). Unlike the
 * reference, a size mismatch is reported (SNC_E_INVALID), not ignored. */
int snc_hip_train_set_weights(snc_hip_train* s, const float* w, int64_t n);
int snc_hip_train_get_weights(const snc_hip_train* s, float* w, int64_t n);

/* Upload a whole dataset once; later steps index it. Features: [n][dim]
 * values in [0, 1] encoded on the device (cfg.encoder: Philox Poisson with
 * counter (channel / 4, step, sample index, epoch), Direct, or Latency). Events:
 * per-sample InputEvent lists (sample_id = dataset index). lengths[i] is T_i
 * (NULL: cfg.num_steps for every sample). */
int snc_hip_train_upload_features(snc_hip_train* s, const float* X, int32_t n, int32_t dim,
                                  const int32_t* labels, const int32_t* lengths);
int snc_hip_train_upload_events(snc_hip_train* s, const snc_event* ev, int64_t n_events,
                                int32_t n, const int32_t* labels, const int32_t* lengths);

typedef struct snc_train_stats {
  int64_t spikes;
  int64_t synaptic_events;
  double loss_sum; /* sum over samples of -log softmax(logits)[label]          */
  int32_t correct;
  int32_t samples;
  double ms;       /* device time of the step's launches                          */
} snc_train_stats;

/* One minibatch: the samples idx[0..B) of the uploaded dataset run from the
 * current weights; when apply is set the weights then move by the
 * minibatch-mean e-prop update (cuda_trainer.cu:163-172, with lr / T_b per
 * sample): w <- clamp(w - sum_b (lr / T_b * L_b[post]) * E_b / B, +-w_max).
 * In reward mode the samples draw their actions and advance the baseline in
 * idx order, so B = 1 is the CPU Trainer's per-sample step exactly; steps
 * without apply draw nothing (the reference draws only in apply_update).
 * logits (optional) receives [B][classes]; raster (optional) [B][T_max][ceil(N/32)]. */
int snc_hip_train_step(snc_hip_train* s, const int32_t* idx, int32_t B, int32_t epoch,
                       int32_t mode, int32_t apply, float* logits, uint32_t* raster,
                       snc_train_stats* st);

/* cudatrain::train_epoch / evaluate (cuda_trainer.cu:374-435) over the
 * uploaded dataset: shuffled minibatches (mt19937_64 seeded with
 * 0x9e3779b97f4a7c15 * (epoch + 1)); evaluate runs without updates.
 * train_epoch covers the samples idx[0..n) (the training split), or the whole
 * uploaded pool when idx is NULL. */
int snc_hip_train_epoch(snc_hip_train* s, const int32_t* idx, int32_t n, int32_t epoch,
                        int32_t mode, snc_train_stats* st);
int snc_hip_train_evaluate(snc_hip_train* s, const int32_t* idx, int32_t n, int32_t mode,
                           snc_train_stats* st);

/* Where the batched kernel's time goes (no counterpart in the reference: an
 * observability measurement). Runs one batched minibatch of dataset samples
 * idx[0..B) with a timed build of the kernel, computing the contributions but
 * applying no update (weights and statistics are unchanged). For each of the
 * SNC_TRAIN_PHASES phases, mean_us[k] is the time all workgroups together
 * spent in it divided by B (per sample; eligibility rounds run on whichever
 * workgroup claims them) and max_us[k] the largest time one workgroup spent in
 * it: 0 set-up, 1 LIF + trace + histories, 2 delivery bits, 3 bitmap prefix
 * scan, 4 arrival staging, 5 ordered fold + next injections, 6 end-of-sample
 * epilogue, 7 eligibility sums, 8 history chunks streamed into LDS,
 * 9 contribution stores, 10 claiming (and waiting for) eligibility work.
 * *kernel_ms (may be NULL) is the launch's device time by HIP events. */
#define SNC_TRAIN_PHASES 11
int snc_hip_train_profile(snc_hip_train* s, const int32_t* idx, int32_t B, int32_t epoch,
                          double* mean_us, double* max_us, double* kernel_ms);

/* cudatrain::reset_stats / collect_stats: per-synapse deliveries and
 * per-neuron fires accumulated over training steps (not evaluation). */
int snc_hip_train_reset_stats(snc_hip_train* s);
int snc_hip_train_collect_stats(const snc_hip_train* s, int64_t* syn_deliveries,
                                int64_t* neuron_fires);

/* ---- learned grow proposal (no reference counterpart) ---------------------
 * A 16 -> 64 -> 64 -> 1 MLP with BF16 weights and FP32 accumulation scores
 * candidate synapses for the grow step of Connectome::structural_update
 * (src/structure/connectome.cpp:164-211); the heuristic stays as fallback.
 * Parameters: FP32 master vector of SNC_MLP_PARAMS floats laid out as
 * W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3. */
#define SNC_MLP_IN 16
#define SNC_MLP_PARAMS 5313

typedef struct snc_hip_mlp snc_hip_mlp;

/* He-uniform weights, zero biases, from SplitMix64(seed). */
int snc_hip_mlp_create(uint64_t seed, snc_hip_mlp** out);
void snc_hip_mlp_destroy(snc_hip_mlp* m);
int snc_hip_mlp_get_params(const snc_hip_mlp* m, float* params);
int snc_hip_mlp_set_params(snc_hip_mlp* m, const float* params);

/* Logits of M candidates (features [M][16], host arrays). *kernel_ms (may be
 * NULL) is the device time of the scoring kernel alone, inputs resident. */
int snc_hip_mlp_score(snc_hip_mlp* m, const float* features, int64_t M, float* scores,
                      double* kernel_ms);

/* Mean binary cross-entropy and its gradient w.r.t. the FP32 master
 * parameters over M labelled examples (labels 0/1). */
int snc_hip_mlp_grad(snc_hip_mlp* m, const float* features, const float* labels, int64_t M,
                     float* grad, double* loss);

/* Allocation-free form of snc_hip_mlp_grad. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight training or serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as snc_hip_mlp_grad for at most
 * SNC_MLP_NOALLOC_MAX_EXAMPLES examples; a larger M is rejected
 * (SNC_E_INVALID). It calls no hipMalloc or hipFree: the staged examples and
 * the per-example gradients and losses live in one fixed device array of
 * about 21.8 MB per GPU, held from the first model created on that GPU until
 * the process ends. Every call on a GPU shares that array, so calls that
 * overlap on one GPU (from several threads or models) race. */
#define SNC_MLP_NOALLOC_MAX_EXAMPLES 1024
int snc_hip_mlp_grad_noalloc(snc_hip_mlp* m, const float* features, const float* labels,
                             int64_t M, float* grad, double* loss);

/* Minibatch Adam on the GPU: `epochs` shuffled passes (SplitMix64(seed)). */
int snc_hip_mlp_train(snc_hip_mlp* m, const float* features, const float* labels, int64_t M,
                      int32_t epochs, int32_t batch, float lr, uint64_t seed, double* final_loss);

#ifdef __cplusplus
}
#endif
#endif /* SNC_HIP_H_ */
