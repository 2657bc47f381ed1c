// Device half of the sheet connectome's structural step (conn_gpu.hip):
// the prune selection and the learned grow's candidate generation, features
// and scoring. connectome.cpp is its host side and its restatement.
#pragma once

#include <cstdint>
#include <vector>

#include "snc_hip.h"

namespace snc_amd {

// Everything the candidate kernels read, from the host connectome and its
// per-neuron feature statistics (computed on the host exactly as the host
// path computes them, so every feature matches bit for bit).
struct ConnSheetView {
  int n = 0, num_layers = 0;
  std::vector<int32_t> layer, layer_begin, layer_count, side;
  std::vector<float> px, py;
  std::vector<double> fire;  // fire count of each neuron (0 past the statistics)
  std::vector<double> rate;  // fire / activity_norm
  // per-neuron feature columns: f5 in-degree / mean, f6 out-degree / mean,
  // f7 mean |w| in, f8 mean |w| out, f9 layer position, f10 log1p(fire),
  // f14 fire / layer mean fire
  std::vector<float> f5, f6, f7, f8, f9, f10, f14;
  float f4 = 0.0f;  // delay / 32
};

// Marks (mask[e] = 1) the K smallest keys (|w[e]|, e) among synapses with
// age[e] >= protect, in the order of the arrays (CSR order).
int conn_gpu_prune_select(const float* w, const int32_t* age, int64_t E, int protect, int K,
                          uint8_t* mask);

// Window candidates of every freed post (disc of radius `window` over the
// previous and, for hidden posts, the own layer; by layer then id), their
// 16 features (on the device, bf16) and MLP scores. offs[u]..offs[u+1] index
// post u's candidates in all_q / rank / score; rank is each candidate's
// window rank (feature 13) and wmax[u] post u's window maximum fire count,
// the two window statistics of the features.
int conn_gpu_candidates(const ConnSheetView& v, const std::vector<int>& freed, float window,
                        snc_hip_mlp* model, std::vector<int>& offs, std::vector<int>& all_q,
                        std::vector<float>& rank, std::vector<double>& wmax,
                        std::vector<float>& score, double* score_ms);

}  // namespace snc_amd
