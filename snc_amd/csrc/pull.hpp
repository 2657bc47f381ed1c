// hip-pull: batched forward by gathering, for up to 64 samples at once.
//
// Every neuron's firing of step t over the plan's B samples is one word
// F[t % L][i] (bit b: sample b fired; 32-bit words for B <= 32, 64-bit for
// B <= 64). A target's input at step t is the sum, over its incoming
// synapses (pre i, delay d, value w) in column order, of w for every sample
// whose bit is set in F[(t - d) % L][i]: the deposits the CPU ring holds for
// slot t, gathered instead of scattered. No atomics are needed and one pass
// over the synapses serves all B samples.
//
// Light targets (in-degree up to the plan's cap). Their column entries are
// stored back to back in column order as a value (fp32) and a 16-bit halo
// index. One wave steps one target at a time with lane b = sample b: the
// wave loads 64 column entries at once, looks up each entry's
// fired word in LDS, and walks the entries whose word is nonzero in column
// order; lane b adds the entry's value when bit b is set (and +0 otherwise,
// which leaves a sum started from +0 unchanged). Every sample's sum is thus
// the CPU's sequential slot sum, bit for bit. The fired words a block needs
// are its halo: the distinct (pre, delay) pairs of its light targets'
// columns, staged in LDS once per step (sorted by delay, then pre, so the
// loads coalesce); entries beyond the halo capacity read F directly.
//
// Hub targets (in-degree above the cap) keep their entries in column order
// too -- (delay descending, pre, CSR position), so consecutive gathers of a
// chunk share F's cache lines -- cut into chunks of at most kPullHubChunk
// entries. A chunk is one wave's work in a launch of its own before the step
// launch: it gathers the chunk's fired words from F and, through the same bit
// transpose, lists every sample's arriving values in column order (hub_list:
// entry slot i of the chunk, sample b at (chunk start + i) * lstride + b;
// their count per chunk and sample in hub_cnt). In the step launch the wave
// that steps the hub's id range walks those lists chunk after chunk, lane b
// adding sample b's values one by one from +0: the CPU's slot sum bit for
// bit, as for light targets. A chunk gathers only the words whose bit is set
// in the fired summary: one bit per neuron and step, set when the neuron
// fired in any sample, written 16 neurons to a half-word by the wave that
// steps them (no atomics).
//
// State is sample-minor, [n][B]: a wave's 64 lanes touch one target's samples
// in one 256-byte run. Spike and event counts are not accumulated per step:
// the host derives them from the per-neuron spike counts (events = spikes of
// neuron i times its out-degree, as cpu_backend.cpp:168-186 counts them).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "forward_plan.hpp"
#include "wave.hpp"

namespace snc_amd {

constexpr int kPullWaveTargets = 16;   // targets one wave steps
constexpr int kPullMaxBlock = 256;     // targets per block (a multiple of 16, at least 64)
constexpr int kPullMinBlock = 64;
constexpr int kPullMaxThreads = kPullMaxBlock / kPullWaveTargets * kWave;
constexpr int kPullMaxSamples = 64;    // bits of the widest F word
constexpr int kPullHubChunk = 512;     // hub column entries per chunk wave
constexpr int kPullHaloCap = 9216;     // fired words a block stages in LDS
constexpr size_t kPullMaxLds = 150 * 1024;
constexpr int kPullMaxCap = 1024;      // largest light in-degree cap

// A light entry's halo index is its LDS slot; kHaloDirect marks an entry past
// its block's halo capacity, which reads F directly ((pre << 8) | delay from
// the side array). kHaloNone marks a lane past a hub chunk's end.
constexpr uint16_t kHaloDirect = 0xFFFFu;
constexpr uint32_t kHaloNone = 0xFFFFFFFFu;
constexpr int kPullStreamPad = 128;    // entries past the light stream's end (over-reads)

// Per-graph, per-cap arrays (device).
struct PullGraph {
  int cap = -1;
  int blk = 0;                    // targets per block
  int n_blocks = 0;
  int n_hub = 0, n_chunks = 0;
  int halo_words = 0;             // largest staged halo (LDS words)
  bool direct = false;            // some light entry reads F directly
  int64_t direct_entries = 0;     // light entries past a block's halo capacity
  float* lw = nullptr;            // light entries' values, target by target in column order
  uint16_t* lh = nullptr;         // their halo indices (kHaloDirect: see ldir)
  uint32_t* ldir = nullptr;       // (pre << 8) | delay of direct entries (else unused)
  int32_t* loff = nullptr;        // [n] first light entry of a target
  int32_t* lcnt = nullptr;        // [n] light entries of a target, -1 for hubs
  int32_t* hub_of = nullptr;      // [n] hub index of a target, -1 for light targets
  int64_t* halo_ptr = nullptr;    // [n_blocks + 1]
  uint32_t* halo = nullptr;       // (pre << 8) | delay, per block by (delay, pre)
  int32_t* hub_ids = nullptr;     // [n_hub]
  int32_t* chunk_hub = nullptr;   // [n_chunks] hub index
  int64_t hub_entries = 0;        // entries of all hub columns
  uint32_t* hub_pk = nullptr;     // hub columns: (pre << 8) | delay, in column order
  float* hub_w = nullptr;         // the value of each hub entry
  int32_t* chunk_pos = nullptr;   // [n_chunks] first entry in hub_pk / hub_w
  int32_t* chunk_len = nullptr;   // [n_chunks] entries
  int32_t* hub_chunk_ptr = nullptr;  // [n_hub + 1]
  int32_t* blk_order = nullptr;   // [n_blocks] target block of each launch slot
  std::vector<int32_t> hub_host;
  int64_t dev_bytes = 0;
  void release();
  ~PullGraph() { release(); }
};

struct PullArgs {
  int n, L, B, T, words, n_blocks;
  int blk;    // targets per block (16 per wave)
  int pool;   // injection entries a block stages per step
  int halo;   // LDS halo words of this launch
  int direct; // some light entry reads F directly
  LifK lif;
  // graph
  const float* lw;
  const uint16_t* lh;
  const uint32_t* ldir;
  const int32_t* loff;
  const int32_t* lcnt;
  const int32_t* hub_of;
  const int64_t* halo_ptr;
  const uint32_t* halo_pk;
  const int32_t* hub_ids;
  const uint32_t* hub_pk;
  const float* hub_w;
  const int32_t* chunk_hub;
  const int32_t* chunk_pos;
  const int32_t* chunk_len;
  const int32_t* hub_chunk_ptr;
  const int32_t* blk_order;
  int n_hub, n_chunks;
  // injections: per (sample, step, block) lists and a dense hub table
  const int32_t* inj_blk_ptr;  // [B][T][n_blocks + 1]
  const int32_t* blk_nid;
  const float* blk_amp;
  const float* inj_hub;        // [B][T][n_hub]
  int64_t inj_stride, hinj_stride;
  // state
  void* F;            // [L][n] words
  uint16_t* fsum;     // [L][sum_words] fired summary: bit i of word q = neuron 16 q + i
  int sum_words;
  float* hub_list;    // [hub entries][lstride] arriving values per chunk and sample
  uint16_t* hub_cnt;  // [n_chunks][lstride] how many of them
  int lstride;        // samples per list row: B rounded up to a multiple of 8
  float* v;           // [n][B]
  uint8_t* refr;      // [n][B] refractory steps left (the refractory period is at most 255)
  int32_t* nspk;      // [n][B]
  uint32_t* raster;   // [B][T][words] or null
  int64_t raster_stride;
};

// Builds the halo-indexed columns of `g` at light cap `cap` for blocks of
// `blk` targets (nothing if pg already holds them).
int pull_prepare(const DevGraph& g, int cap, int blk, PullGraph& pg);
// LDS bytes of a step launch.
size_t pull_lds_bytes(int B, int halo, int pool, int blk);
// Allows the step kernel of B samples the LDS it needs.
int pull_set_lds(int B, int halo, int pool, int blk);
// Whether blk is a valid block size.
bool pull_block_ok(int blk);
// One step: the hub-chunk launch (if the graph has hubs), then the step launch.
int pull_launch_step(const PullArgs& a, int t, hipStream_t st);
// Bytes of one F word for B samples.
inline int pull_word_bytes(int B) { return B > 32 ? 8 : 4; }

}  // namespace snc_amd
