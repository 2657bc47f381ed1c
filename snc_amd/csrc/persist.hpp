// Persistent (cooperative) forward kernel: the whole step loop in one launch.
#pragma once

#include "device_graph.hpp"
#include "forward_plan.hpp"

namespace snc_amd {

struct PersistConfig {
  int blocks = 0;
  int threads = 0;
  bool lds_ring = false;
};

// Decide whether a plan runs as one persistent launch; fills cfg.
bool persist_choose(const DevGraph& g, int variant, PersistConfig& cfg);
int persist_launch(const DevGraph& g, const FwdPlanCore& pc, const StepArgs& a,
                   const PersistConfig& cfg);

}  // namespace snc_amd
