/*
 * snc_struct.h -- C ABI of the structural slow path with the learned grow
 * proposal.
 *
 * Reference interfaces mirrored (paths relative to the reference tree):
 *   snc_conn_layered_local       <- Connectome::layered_local   src/structure/connectome.cpp:15-64
 *   snc_conn_to_graph            <- Connectome::to_graph        src/structure/connectome.cpp:78-102
 *   snc_conn_edge_weights / set  <- edge_weights / set_edge_weights  connectome.cpp:104-119
 *   snc_conn_structural_update   <- Connectome::structural_update    connectome.cpp:121-219
 * With grow_mode SNC_GROW_HEURISTIC the update is the reference's (prune the
 * K weakest |w| among prune-eligible synapses, regrow each freed post from
 * its most-firing unconnected local pre, fallback tournament onto
 * under-driven posts). SNC_GROW_LEARNED replaces the 1:1 regrowth choice by
 * the highest MLP score among the same candidates (snc_hip_mlp_*); the
 * fallback tournament stays heuristic. Every grown synapse can be logged with
 * its 16 features and later labelled by whether it survived the first prune
 * it was eligible for.
 */
#ifndef SNC_STRUCT_H_
#define SNC_STRUCT_H_

#include <stdint.h>

#include "snc_gen.h"
#include "snc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SNC_GROW_HEURISTIC 0
#define SNC_GROW_LEARNED 1
/* Learned, over the whole candidate pool: every window candidate of every
 * non-input post is scored each step and the highest-scoring unconnected
 * pairs regrow as many synapses as were pruned, wherever their posts are
 * (the fallback tournament fills any shortfall). */
#define SNC_GROW_LEARNED_POOL 2

/* StructConfig (include/snc/connectome.hpp:33-48). */
typedef struct snc_struct_config {
  int32_t grow_per_epoch;
  int32_t delay;
  float w_grow_init; /* negative: the surviving synapses' median |w| at each step
                       (no reference counterpart; the reference's value is fixed) */
  float locality_window;
  int32_t protect_rounds;
  int32_t max_attempts_mult;
  int32_t reward_modulated;
  float reward_chance;
  float reward_floor;
  uint64_t seed;
  /* Heuristic mode with grow logging only: probability that a freed post is
   * regrown from a uniformly drawn unconnected window candidate instead of the
   * most-firing one, so the log also holds outcomes of non-greedy choices.
   * Drawn from a generator of its own; 0 leaves the heuristic unchanged. */
  float explore;
  /* Sheet connectomes only: select the prune set on the GPU, and in learned
   * mode generate the window candidates, their features and scores there
   * (candidates never leave the device before scoring). Same result as the
   * host path, which restates it. */
  int32_t on_device;
} snc_struct_config;

/* StructReport (connectome.hpp:50-54) plus the candidates scored. */
typedef struct snc_struct_report {
  int32_t pruned;
  int32_t grown;
  int32_t synapses;
  int64_t candidates_scored;
  double score_ms; /* device time of the MLP scoring kernel */
} snc_struct_report;

typedef struct snc_connectome snc_connectome;

int snc_conn_layered_local(const int32_t* layers, int32_t n_layers, int32_t synapse_budget,
                           int32_t delay, float w_init, uint64_t seed, snc_connectome** out);
/* A layered connectome on 2-D sheets with recurrent hidden layers (no
 * counterpart in the reference, whose Connectome is layered feed-forward with
 * 1-D positions, connectome.cpp:30-62). Every layer is a side x side grid over
 * the unit square (side = ceil(sqrt(count))); each neuron of layer l >= 1 draws
 * fan_in partners from layer l-1 and, in a hidden layer, rec_fan_in from its
 * own layer (no self-synapse), with probability proportional to
 * exp(-d / radius), SplitMix64 draws. Grow windows are discs of radius
 * locality_window around the post over layer l-1 and (hidden posts) layer l.
 * Pruning removes the K smallest (|w|, CSR position). */
int snc_conn_sheet_recurrent(const int32_t* layers, int32_t n_layers, int32_t fan_in,
                             int32_t rec_fan_in, float radius, int32_t delay, float w_init,
                             uint64_t seed, snc_connectome** out);
/* Thins a connectome to a synapse budget (no reference counterpart: it sets
 * up a co-training base "under an N-synapse budget"): removes the synapses
 * of smallest (|w|, CSR position) until at most `budget` remain; ages and
 * the grow log of the survivors are kept. */
int snc_conn_prune_to(snc_connectome* c, int32_t budget);
void snc_conn_free(snc_connectome* c);
int32_t snc_conn_num_synapses(const snc_connectome* c);
int snc_conn_to_graph(const snc_connectome* c, snc_graph_buf** out);
int snc_conn_edge_weights(const snc_connectome* c, float* w, int64_t n);
int snc_conn_set_edge_weights(snc_connectome* c, const float* w, int64_t n);

/* One structural step from the inner loop's activity statistics
 * (ActivityStats, connectome.hpp:28-31). activity_norm divides fire counts
 * into rates for the features (e.g. samples * steps of the inner loop).
 * model may be NULL (then learned mode falls back to the heuristic). When
 * log_grown is set, grown synapses are appended to the grow log. */
int snc_conn_structural_update(snc_connectome* c, const int64_t* syn_deliveries, int64_t n_syn,
                               const int64_t* neuron_fires, int64_t n_neur,
                               const snc_struct_config* cfg, float reward, int32_t grow_mode,
                               snc_hip_mlp* model, double activity_norm, int32_t log_grown,
                               snc_struct_report* rep);

/* Grow log: n entries of 16 features and a label (1 survived its first
 * eligible prune, 0 pruned by it, -1 not yet decided). */
int64_t snc_conn_grow_log_size(const snc_connectome* c);
int snc_conn_grow_log(const snc_connectome* c, float* features, int32_t* labels, int64_t n);
/* Graded outcome of each grow-log entry (no reference counterpart; a
 * training target for the learned proposal): the percentile of the edge's
 * |w| among all prune-eligible edges at its first eligible prune (fraction
 * below it, ties counted half), -1 while undecided. */
int snc_conn_grow_log_rank(const snc_connectome* c, float* rank, int64_t n);

/* The 16 features of candidate synapses (pre, post) under the given
 * statistics, as the learned grow step computes them. */
int snc_conn_candidate_features(const snc_connectome* c, const int32_t* pre, const int32_t* post,
                                int64_t n, const int64_t* neuron_fires, int64_t n_neur,
                                double activity_norm, int32_t delay, float* features);

#ifdef __cplusplus
}
#endif
#endif /* SNC_STRUCT_H_ */
