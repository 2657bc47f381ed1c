/*
 * snc_gen.h -- C ABI of the host-side generators and input encoders that feed
 * the HIP backend: the graph and input formats on either side of the spiking
 * path. Deterministic for a fixed seed on any x86-64 host (SplitMix64 and
 * Philox4x32-10; no libstdc++ distributions except where the reference's own
 * encoder semantics require them, see snc_encode).
 *
 * Reference interfaces mirrored (paths relative to the reference tree):
 *   snc_gen_static_snc   <- snc::make_static_snc      src/graph/snn_graph.cpp:185-236
 *   snc_gen_random_sparse<- snc::make_random_sparse   src/graph/snn_graph.cpp:288-326
 *   snc_gen_dense        <- snc::make_dense           src/graph/snn_graph.cpp:267-286
 *   snc_graph_validate   <- SNNGraph::validate        src/graph/snn_graph.cpp:30-56
 *   snc_encode           <- snc::encode               src/training/encoders.cpp:25-71
 *   snc_make_synthetic   <- snc::make_synthetic       src/training/datasets.cpp:11-30
 * The sheet / hub / recurrent generators and the image, drive and event-stream
 * inputs have no counterpart in the reference: they build the benchmark
 * workloads (2-D locality, power-law hubs, random recurrence, Poisson drive).
 */
#ifndef SNC_GEN_H_
#define SNC_GEN_H_

#include <stdint.h>

#include "snc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct snc_graph_buf snc_graph_buf; /* owns SNNGraph-shaped arrays */
typedef struct snc_event_buf snc_event_buf; /* owns an InputEvent array    */

/* Pointers stay valid until snc_graph_buf_free. */
int snc_graph_buf_view(const snc_graph_buf* b, snc_graph_view* out);
void snc_graph_buf_free(snc_graph_buf* b);
/* Build a buffer from caller arrays (copied); validates like SNNGraph::validate. */
int snc_graph_buf_from_view(const snc_graph_view* g, snc_graph_buf** out);

int snc_event_buf_data(const snc_event_buf* b, const snc_event** data, int64_t* n);
void snc_event_buf_free(snc_event_buf* b);

/* 1 if valid; otherwise 0 and a message in snc_hip_last_error(). */
int snc_graph_validate(const snc_graph_view* g);

/* Layered graphs on a 2-D sheet. Every layer is laid out as a near-square grid
 * stretched over a sheet x sheet square (input grid spacing = 1 unit). Each
 * neuron of layer l >= 1 draws `fan_in` distinct presynaptic partners from
 * layer l-1 with probability proportional to exp(-d / radius); hidden layers
 * additionally draw `recurrent_fan_in` partners from their own layer (no
 * self-synapses). With dense_readout the last layer is fully connected from
 * the one before. Weights ~ N(0, w_std) (signed, sign = +1); delays are
 * `delay`, or 1 + round((max_delay-1) * d / d_max) when max_delay > 0.
 * Layer 0 is INPUT (channel = index), the last layer OUTPUT. */
typedef struct snc_sheet_params {
  int32_t n_layers;
  const int32_t* layers;
  int32_t fan_in;
  int32_t recurrent_fan_in;
  int32_t dense_readout;
  float radius;
  float sheet;
  float w_std;
  int32_t delay;
  int32_t max_delay;
  uint64_t seed;
} snc_sheet_params;
int snc_gen_sheet(const snc_sheet_params* p, snc_graph_buf** out);

/* Delivery stress graph: width x height LIF sheet, every neuron with exactly
 * `out_degree` outgoing synapses. A hub_fraction of neurons are hubs whose
 * in-degree follows a power law (exponent hub_alpha on [hub_min, hub_max]);
 * hub synapses come from uniformly random sources, the rest of each row from
 * 2-D locality (radial density ~ d exp(-d/radius)). Delays are
 * 1 + round((max_delay-1) * d / d_max). Weights ~ N(0, w_std). When all_input
 * is set every neuron is INPUT with channel = id, so a per-neuron background
 * drive can be given as ordinary input events. */
typedef struct snc_hub_params {
  int32_t width;
  int32_t height;
  int32_t out_degree;
  float radius;
  float hub_fraction;
  float hub_alpha;
  int32_t hub_min;
  int32_t hub_max;
  float w_std;
  int32_t max_delay;
  int32_t all_input;
  uint64_t seed;
} snc_hub_params;
int snc_gen_hub_sheet(const snc_hub_params* p, snc_graph_buf** out);

/* Random-sparse recurrent network: n_in INPUT -> n_hidden recurrent LIF ->
 * n_out OUTPUT. Input->hidden and hidden->hidden edges kept independently
 * with density_in / density_rec (no self-synapses), hidden->output dense.
 * Delays uniform in [delay_min, delay_max] (clamped to >= 1) for recurrent
 * edges, 1 elsewhere. Weights ~ N(0, w_std). */
int snc_gen_random_recurrent(int32_t n_in, int32_t n_hidden, int32_t n_out, float density_in,
                             float density_rec, int32_t delay_min, int32_t delay_max,
                             float w_std, uint64_t seed, snc_graph_buf** out);

/* Reference generators, same algorithm and libstdc++ draws as the reference. */
int snc_gen_static_snc(const int32_t* layers, int32_t n_layers, int32_t budget, float weight,
                       int32_t delay, float inhib, uint64_t seed, snc_graph_buf** out);
int snc_gen_random_sparse(const int32_t* layers, int32_t n_layers, int32_t budget,
                          float weight, int32_t delay, float inhib, uint64_t seed,
                          snc_graph_buf** out);
int snc_gen_dense(const int32_t* layers, int32_t n_layers, float weight, int32_t delay,
                  snc_graph_buf** out);

/* Background drive: every neuron of [0, n) receives an event of `amplitude`
 * on each step with probability p, from Philox4x32-10 keyed by seed with
 * counter (neuron, step). Events are ordered by step, then neuron. */
int snc_gen_background(int32_t n, int32_t steps, double p, float amplitude, uint64_t seed,
                       snc_event_buf** out);

/* n synthetic width x height u8 images: each pixel is 0 with probability
 * zero_frac, else uniform in 1..255. */
void snc_gen_images(int32_t n, int32_t pixels, float zero_frac, uint64_t seed, uint8_t* out);

/* Labelled images: `classes` prototypes drawn as snc_gen_images; sample i has
 * label i % classes and each pixel replaced by a fresh random pixel with
 * probability noise. */
void snc_gen_prototype_images(int32_t n, int32_t pixels, int32_t classes, float zero_frac,
                              float noise, uint64_t seed, uint8_t* X, int32_t* y);

/* Sparse event streams: sample i has label i % classes, length
 * T_i ~ U[t_min, t_max], ~events_per_sample events with times ~ U[0, T_i).
 * Channel ids follow a mixture of n_centres Gaussians (sd) whose centre order
 * over time is a class-specific permutation. Writes lengths[n], labels[n];
 * events carry sample_id = i and amplitude 1, sorted by (sample, time). */
int snc_gen_event_streams(int32_t n, int32_t channels, int32_t classes, int32_t n_centres,
                          float sd, int32_t events_per_sample, int32_t t_min, int32_t t_max,
                          uint64_t seed, int32_t* lengths, int32_t* labels,
                          snc_event_buf** out);

/* encode (encoders.cpp:25-71): enc 0 direct, 1 poisson (mt19937_64 +
 * uniform_real_distribution<float>, the reference's draws), 2 latency. */
int snc_encode(int32_t enc, const float* features, int32_t n_channels, int32_t num_steps,
               float gain, float max_rate, float threshold, uint64_t seed, int32_t sample_id,
               snc_event_buf** out);

/* make_synthetic (datasets.cpp:11-30). X is [n][dim]. */
void snc_make_synthetic(int32_t n, int32_t dim, int32_t classes, float noise, uint64_t seed,
                        float* X, int32_t* y);

/* ---- SNCG flat graph files (the reference's PyTorch-bridge format) ----
 * Layout, little-endian int32 (src/bench/snc_export.cpp:10-12, 107-123;
 * python/snc/graph.py:11-24): magic 0x534E4347 'SNCG', N, S,
 * num_input_channels, num_output_channels, pre[S], post[S], delays[S],
 * role[N], channel[N]. Weights and signs are not stored. */

/* num_input_channels / num_output_channels as SNNGraph::rebuild_io_index
 * computes them (src/graph/snn_graph.cpp:13-28). */
int snc_graph_io_channels(const snc_graph_view* g, int32_t* n_in, int32_t* n_out);
/* snc_export's file write (snc_export.cpp:107-123); the graph is validated first. */
int snc_graph_write_sncg(const snc_graph_view* g, const char* path);
/* load_graph (python/snc/graph.py:11-24) into a CSR buffer: rows in file order
 * per pre, every weight = `weight`, every sign +1. Rejects a bad magic, a size
 * that does not match N and S, ids out of range, delays < 1 and unknown roles.
 * n_in / n_out (nullable) receive the header's channel counts. */
int snc_graph_read_sncg(const char* path, float weight, snc_graph_buf** out, int32_t* n_in,
                        int32_t* n_out);

/* snc_export's heterogeneous delays (snc_export.cpp:81-101), applied in place
 * when delay_max > 1:
 *   SNC_DELAYS_RANDOM    uniform in [1, delay_max] from mt19937_64(seed ^ 0xDE1A0)
 *   SNC_DELAYS_DISTANCE  1 + round(|pre - (post - A)| / (max(A, B) - 1) * (delay_max - 1)),
 *                        A, B = layers[0], layers[1]; needs n_layers >= 2, else the
 *                        random rule applies, as in the exporter */
#define SNC_DELAYS_RANDOM 0
#define SNC_DELAYS_DISTANCE 1
int snc_graph_spread_delays(snc_graph_buf* b, int32_t mode, int32_t delay_max,
                            const int32_t* layers, int32_t n_layers, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif /* SNC_GEN_H_ */
