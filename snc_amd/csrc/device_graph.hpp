// Device-resident SNNGraph plus the derived arrays the delivery strategies use.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "snc_hip.h"

namespace snc_amd {

// Targets whose in-degree exceeds this are folded by a whole wave instead of
// one lane (see k_fold_heavy); the autotuner may move it.
constexpr int kDefaultHeavyInDegree = 512;

struct DevGraph {
  int n = 0;
  int64_t S = 0;
  int max_delay = 1;
  int L = 2;  // ring length = max_delay + 1 (cpu_backend.cpp:75-80)

  // CSR over outgoing synapses (rows sorted by pre).
  int32_t* row_ptr = nullptr;  // [n+1]
  int32_t* post = nullptr;     // [S]
  float* wsig = nullptr;       // [S] float(sign[pre]) * weight, the value each event adds
  uint8_t* delay = nullptr;    // [S]

  // Incoming view for the ordered fold. Column j holds j's incoming synapses
  // in (delay descending, pre ascending, CSR position ascending) order -- the
  // order in which the CPU path adds the deposits that arrive together --
  // padded to a multiple of 32 so every 32-bit word of the presence bitmap
  // belongs to exactly one target.
  int32_t* cpos = nullptr;   // [S] padded column position of synapse e
  int32_t* ccol = nullptr;   // [n+1] padded column starts (multiples of 32)
  float* cval = nullptr;     // [P] wsig in column order (0 in padding)
  uint8_t* cdel = nullptr;   // [P] delay in column order
  int64_t P = 0;             // padded column length total
  int heavy_threshold = kDefaultHeavyInDegree;
  int32_t* heavy_ids = nullptr;  // targets with in-degree > heavy_threshold
  uint8_t* is_heavy = nullptr;   // [n]
  int n_heavy = 0;

  // Host copies used for decoding and the autotuner.
  std::vector<uint8_t> role;
  std::vector<int32_t> channel;
  std::vector<int32_t> in_degree, out_degree;
  int n_out_channels = 0;
  std::vector<std::vector<int32_t>> chan_to_inputs;  // INPUT neurons per channel, ascending
  std::vector<int32_t> heavy_host;                   // heavy target ids, ascending
  snc_graph_profile profile{};
  int64_t dev_bytes = 0;  // device memory of the arrays above

  ~DevGraph();
};

}  // namespace snc_amd

namespace snc_amd {
struct PullGraph;
void pull_free(PullGraph* pg);
}  // namespace snc_amd

struct snc_hip_graph {
  snc_amd::DevGraph g;
  std::vector<snc_amd::PullGraph*> pulls;  // hip-pull light columns and halos, per (light cap, block size)
  ~snc_hip_graph() {
    for (snc_amd::PullGraph* p : pulls) snc_amd::pull_free(p);
  }
  // Autotuner results: [exact_only + 2 * batched].
  bool tuned[4] = {false, false, false, false};
  snc_tune_report tune[4] = {};
  int tuned_samples[4] = {0, 0, 0, 0};  // the sample count each search ran with
};
