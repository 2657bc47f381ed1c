// Persistent (cooperative) forward kernel: the whole step loop in one launch.
#pragma once

#include <cstdint>
#include <vector>

#include "device_graph.hpp"
#include "forward_plan.hpp"

namespace snc_amd {

struct PersistConfig {
  int blocks = 0;
  int threads = 0;
  int npb = 0;       // neurons per block (multiple of 64)
  int per_lane = 0;  // neurons per lane (registers)
  const int32_t* heavy_blk_ptr = nullptr;  // device [blocks+1] into heavy_ids
  int64_t bits_words = 0;                  // words of one bitmap half
  int64_t dirty_stride = 0;                // bytes of one dirty-flag half
};

struct PersistArgs {
  int npb, per_lane;
  const int32_t* heavy_blk_ptr;
  int64_t bits_words, dirty_stride;
};

// Whether the device can run the persistent kernel at all.
bool persist_supported();
// Decide whether a plan runs as one persistent launch (when `want`, or when
// SNC_HIP_PERSIST=1 asks for it); fills the launch shape.
bool persist_choose(const DevGraph& g, int variant, bool want, PersistConfig& cfg);
std::vector<int32_t> persist_heavy_ptr(const DevGraph& g, const std::vector<int32_t>& heavy_sorted,
                                       const PersistConfig& cfg);
int persist_launch(const DevGraph& g, const FwdPlanCore& pc, const StepArgs& a,
                   const PersistConfig& cfg);

}  // namespace snc_amd
