// hip-pull: batched forward by gathering, for many samples at once.
//
// Every neuron's firing of step t over the plan's B <= 32 samples is one
// 32-bit word F[t % L][i] (bit b: sample b fired). A target's input at step t
// is the sum, over its incoming synapses (pre i, delay d, value w) in column
// order, of w for every sample whose bit is set in F[(t - d) % L][i]: the
// deposits the CPU ring holds for slot t, gathered instead of scattered. No
// atomics are needed and one pass over the synapses serves all B samples.
//
// Light targets (in-degree up to the plan's cap) sit in 64-target slices,
// stored sliced (SELL-64): entry k of every target of a slice side by side,
// so a wave reads 64 consecutive words per step k. Each lane walks its own
// column sequentially and adds into its own LDS row, so a light target's sum
// is the CPU's, bit for bit. Hub targets (in-degree above the cap) are split
// into chunks of kPullHubChunk synapses; a chunk block sums per-sample
// partials in a fixed order and a second launch adds a hub's partials in
// chunk order and steps the hub, so hub sums are deterministic but
// associated differently from the CPU's (the atomic-class tolerance).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "forward_plan.hpp"

namespace snc_amd {

constexpr int kPullBlock = 256;      // targets per light block: 4 slices of 64
constexpr int kPullMaxSamples = 32;  // bits of one F word
constexpr int kPullHubChunk = 2048;  // hub synapses per chunk block
constexpr size_t kPullMaxLds = 150 * 1024;  // LDS a light block may take (pool included)
constexpr int kPullMaxCap = 1024;    // largest light in-degree cap

// Sliced incoming columns of one graph at one light cap (device arrays).
struct PullGraph {
  int cap = -1;
  int n_slices = 0;
  int n_hub = 0, n_chunks = 0;
  uint32_t* slab = nullptr;    // packed (pre << 8) | delay, 0 = padding
  float* slab_w = nullptr;     // value of each entry
  int64_t* slice_ptr = nullptr;  // [n_slices] first entry of slice s
  int32_t* slice_len = nullptr;  // [n_slices] entries per lane
  int32_t* hub_ids = nullptr;    // [n_hub]
  uint32_t* hub_e = nullptr;     // hub columns, packed like slab, column order
  float* hub_w = nullptr;
  int32_t* chunk_hub = nullptr;  // [n_chunks] hub index of a chunk
  int64_t* chunk_beg = nullptr;  // [n_chunks + 1] chunk bounds in hub_e
  int32_t* hub_chunk_ptr = nullptr;  // [n_hub + 1] chunks of hub h
  uint8_t* is_hub = nullptr;     // [n]
  std::vector<int32_t> hub_host;
  int64_t entries = 0;
  int64_t dev_bytes = 0;  // device memory of these arrays
  void release();
  ~PullGraph() { release(); }
};

struct PullArgs {
  int n, L, B, words, n_light_blocks;
  int pool;  // injection entries a light block stages per step (the plan's largest (step, block))
  LifK lif;
  uint32_t sample_mask;
  // graph
  const int32_t* row_ptr;
  const uint32_t* slab;
  const float* slab_w;
  const int64_t* slice_ptr;
  const int32_t* slice_len;
  const uint8_t* is_hub;
  const int32_t* hub_ids;
  const uint32_t* hub_e;
  const float* hub_w;
  const int32_t* chunk_hub;
  const int64_t* chunk_beg;
  const int32_t* hub_chunk_ptr;
  int n_hub, n_chunks;
  // injections: per (sample, step, light block) lists and dense hub table
  const int32_t* inj_blk_ptr;  // [B][T][n_light_blocks + 1]
  const int32_t* blk_nid;
  const float* blk_amp;
  const float* inj_hub;        // [B][T][n_hub]
  int64_t inj_stride, hinj_stride;
  // state
  uint32_t* F;      // [L][n]
  float* hub_part;  // [n_chunks][kPullMaxSamples]
  float* v;         // [B][n]
  int32_t* refr;    // [B][n]
  int32_t* nspk;    // [B][n]
  unsigned long long* blk_spikes;  // [B][blk_stride]
  unsigned long long* blk_events;
  int64_t blk_stride;
  uint32_t* raster;  // [B][T][words] or null
  int64_t raster_stride;
};

// Builds the sliced columns of `g` at light cap `cap` (nothing if pg already
// holds that cap).
int pull_prepare(const DevGraph& g, int cap, PullGraph& pg);
// Allows the step kernel the LDS of B samples and a pool of `pool` entries.
int pull_set_lds(int B, int pool);
// One step: the light/hub-chunk launch, then the hub launch.
int pull_launch_step(const PullArgs& a, int t, hipStream_t st);

}  // namespace snc_amd
