// Forward plan: device buffers and kernel argument block of one (graph,
// sample, variant) forward configuration.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device_graph.hpp"
#include "snc_hip.h"

namespace snc_amd {

struct LifK {
  float threshold, decay, reset;
  int refractory;
  int subtract;
};

// Inline injection table of the fused step kernel: per (sample, step, block)
// kInjInline entries {block-local neuron, amplitude bits}, unused entries
// {-1, 0}. A (step, block) with more injections marks its last entry
// {kInjSpill, 0} and the kernel reads the per-(step, block) list instead.
constexpr int kInjInline = 16;
constexpr int kInjSpill = -2;

// Everything a step kernel touches; passed by value.
// Ints between two hip-sort list counters: each counter has a 128-byte
// cache line of its own, so the expansion's appends to different segments
// do not queue on shared lines.
constexpr int kSortCntStride = 32;
// Largest hip-sort queue storage a plan allocates (L x P records of 8 bytes).
constexpr long long kSortQueueBytesMax = 64ll << 30;

struct StepArgs {
  int n, L, T, words;
  LifK lif;
  // graph
  const int32_t* row_ptr;
  const int32_t* post;
  const float* wsig;
  const uint8_t* delay;
  const int32_t* cpos;
  const int32_t* ccol;
  const float* cval;
  const uint8_t* cdel;
  const int32_t* heavy_ids;
  // inputs (per-step injection CSR, pre-folded per neuron)
  const int32_t* inj_ptr;
  const int32_t* inj_nid;
  const float* inj_amp;
  // state
  float* v;
  int32_t* refr;
  float* ext;
  float* ring;
  uint8_t* dirty;
  uint32_t* bits;
  int32_t* fired;
  int32_t* fcnt;  // [2], indexed by step parity
  int32_t* nspk;  // [n] per-neuron spike counts
  unsigned long long* blk_spikes;
  unsigned long long* blk_events;
  uint32_t* raster;  // [T][words] or null
  uint32_t* fbits;   // [words] fired bitmap of the current step (hip-sort) or null
  // hip-sort: target segments and their arrival queues, per ring slot
  const int4* sseg;            // [n_seg] {first target, end target, first position, end position}
  const int32_t* sseg_e;       // [S] segment of each synapse's target
  uint2* slist;                // [L][P] queue records {column position, value bits}: segment s's at sseg[s].z
  int64_t slist_stride;        // P
  int32_t* scnt;               // [L][n_seg][kSortCntStride]: next free queue position per slot and segment
  const int32_t* sinj_ptr;     // [T][n_seg + 1] each step's injections, per segment
  int n_seg;
  int sort_slot0;              // first stats slot of the segment workgroups
  // fused step kernel
  const int32_t* inj_blk_ptr;  // [T][n_reg_blocks + 1] per-(step, block) injection ranges
  const int32_t* blk_nid;      // injections in (step, block) order, heavy targets excluded
  const float* blk_amp;
  const int2* inj_tab;         // [T][n_reg_blocks][kInjInline] inline injection table
  const float* inj_heavy;      // [T][n_heavy] injections of heavy targets (hip-bucket)
  const uint8_t* is_heavy;     // [n]
  int n_reg_blocks;
  int blk;                     // neurons per block of the fused step kernel
  int n_heavy;
  int64_t bits_half;           // persistent: words of one bitmap half (step parity)
  int64_t dirty_half;          // persistent: bytes of one dirty-flag half
  int64_t bits_slot;           // fused hip-bucket: words of one arrival bitmap (per ring slot)
  // batched plans: samples are grid.y; per-sample strides of the state
  int B;
  int64_t ring_stride;    // L * n
  int64_t bits_stride;    // both bitmap halves of one sample
  int64_t dirty_stride;   // both dirty halves of one sample
  int64_t blk_stride;     // stats slots of one sample
  int64_t inj_stride;     // T * (n_reg_blocks + 1)
  int64_t tab_stride;     // T * n_reg_blocks * kInjInline
  int64_t hinj_stride;    // T * n_heavy
  int64_t raster_stride;  // T * words
};

// The kernel arguments of sample `smp` of a batched plan (uniform offsets).
__host__ __device__ inline StepArgs sample_args(const StepArgs& a0, int smp) {
  StepArgs a = a0;
  if (smp) {
    a.v += (size_t)smp * a.n;
    a.refr += (size_t)smp * a.n;
    a.nspk += (size_t)smp * a.n;
    if (a.ring) a.ring += (size_t)smp * a.ring_stride;
    if (a.bits) a.bits += (size_t)smp * a.bits_stride;
    if (a.dirty) a.dirty += (size_t)smp * a.dirty_stride;
    a.blk_spikes += (size_t)smp * a.blk_stride;
    a.blk_events += (size_t)smp * a.blk_stride;
    a.inj_blk_ptr += (size_t)smp * a.inj_stride;
    if (a.inj_tab) a.inj_tab += (size_t)smp * a.tab_stride;
    if (a.inj_heavy) a.inj_heavy += (size_t)smp * a.hinj_stride;
    if (a.raster) a.raster += (size_t)smp * a.raster_stride;
  }
  return a;
}

struct FwdPlanCore {
  hipStream_t stream = nullptr;
  int variant = SNC_VARIANT_ATOMIC;
  int lif_grid = 1, expand_grid = 1;
  int n_heavy = 0;
  int64_t max_inj_per_step = 0;
};

int fwd_launch_fill(hipStream_t st, void* p, uint32_t value, size_t words);
int fwd_launch_first_injection(const FwdPlanCore& pc, const StepArgs& a);
// Fills a.inj_tab from the per-(sample, step, block) lists.
int fwd_launch_inj_table(hipStream_t st, const StepArgs& a);

// fused per-step kernel (step.hip): one launch per step, hub blocks included
int fused_launch_step(const FwdPlanCore& pc, const StepArgs& a, int t);
bool fused_block_ok(int block);
int fused_light_max_in_degree();  // largest in-degree the step kernel folds per lane

// hip-sort (sort.hip). Segments: runs of targets starting on multiples of 32,
// at most 1024 targets and 65536 column positions each (a longer single
// group of 32 gets its own segment), from the padded column starts ccol [n + 1].
int sort_build_segments(const std::vector<int32_t>& ccol, int n, std::vector<int4>& seg,
                        std::vector<int32_t>& seg_of_word);
// [T][n_seg + 1]: where each step's (neuron-sorted) injections of each segment
// start in inj_nid, from the per-step ranges inj_ptr [T + 1].
std::vector<int32_t> sort_injection_ranges(const std::vector<int4>& seg,
                                           const std::vector<int32_t>& inj_ptr,
                                           const std::vector<int32_t>& inj_nid, int T);
int sort_launch_step(const FwdPlanCore& pc, const StepArgs& a, int t);
int sort_prepare();  // the step kernel's LDS request

}  // namespace snc_amd
