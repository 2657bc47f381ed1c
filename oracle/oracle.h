/*
 * oracle.h -- C ABI of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * A C++ restatement of the reference's CPU path, compiled with
 * -ffp-contract=off so every float expression is evaluated in ISO order.
 * It is the parity checker for the HIP backend and the timed CPU baseline.
 * Only tests, the smoke check and the benchmark's CPU-baseline leg may load
 * it; the product never calls it.
 */
#ifndef SNC_ORACLE_H_
#define SNC_ORACLE_H_

#include <stdint.h>

#include "snc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* run_cpu (src/runtime/cpu_backend.cpp:67-191): serial when use_omp == 0,
 * the OpenMP form otherwise (delivery order then depends on scheduling). */
int oracle_forward(const snc_graph_view* g, const snc_event* events, int64_t n_events,
                   int32_t num_steps, const snc_lif* lif, int32_t use_omp,
                   snc_forward_capture* cap, snc_forward_stats* stats);

/* encode (src/training/encoders.cpp:25-71). enc: 0 direct, 1 poisson,
 * 2 latency. Writes at most `cap` events; *n_out is the full count. */
int oracle_encode(int32_t enc, const float* features, int32_t n_channels, int32_t num_steps,
                  float gain, float max_rate, float threshold, uint64_t seed,
                  int32_t sample_id, snc_event* out, int64_t cap, int64_t* n_out);

/* make_synthetic (src/training/datasets.cpp:11-30): X is [n][dim], y is [n]. */
void oracle_make_synthetic(int32_t n, int32_t dim, int32_t classes, float noise,
                           uint64_t seed, float* X, int32_t* y);

/* Philox4x32-10 (Salmon et al., SC'11): one block for (counter, key). */
void oracle_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* SplitMix64 (Steele, Lea, Flood, OOPSLA'14): next value, advancing *state. */
uint64_t oracle_splitmix64(uint64_t* state);

/* ---- e-prop trainer (oracle_train.cpp) ---------------------------------- */
void* oracle_trainer_create(const snc_graph_view* g, const snc_train_config* cfg);
void oracle_trainer_free(void* h);
void oracle_trainer_get_weights(void* h, float* w);
void oracle_trainer_set_weights(void* h, const float* w);
void oracle_trainer_get_feedback(void* h, float* B);
void oracle_trainer_reset_stats(void* h);
void oracle_trainer_stats(void* h, int64_t* syn_deliv, int64_t* neuron_fires);
int64_t oracle_trainer_run_sample(void* h, const snc_event* ev, int64_t n_ev, int32_t sample,
                                  int32_t T, float* logits, double* elig, uint32_t* raster,
                                  int64_t* events, int32_t record_stats);
double oracle_trainer_train_epoch_ref(void* h, const float* X, const int32_t* y, int32_t n,
                                      int32_t dim, int32_t epoch, double* loss, int64_t* spikes,
                                      int64_t* events);
void oracle_trainer_minibatch(void* h, const snc_event* ev, int64_t n_ev, const int32_t* lengths,
                              const int32_t* labels, int32_t B, int32_t apply, float* logits_out,
                              uint32_t* raster_out, int32_t t_max, int64_t* spikes,
                              int64_t* events);
int64_t oracle_encode_philox(const float* x, int32_t n_ch, int32_t T, float gain, float max_rate,
                             float thr, uint64_t seed, int32_t sample, int32_t epoch,
                             snc_event* out, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif /* SNC_ORACLE_H_ */
