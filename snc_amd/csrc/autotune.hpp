// Per-graph choice of the delivery strategy and its launch shape.
//
// The search space is {strategy} x {step-kernel block} x {hub threshold} x
// {persistent}; the candidates are derived from the graph's in-degree
// statistics, timed on a trial window of the real input, and the fastest is
// cached with the graph.
#pragma once

#include <cstdint>
#include <vector>

#include "device_graph.hpp"
#include "snc_hip.h"

namespace snc_amd {

// Default configuration of an explicitly named strategy on graph g.
snc_plan_config default_plan_config(const DevGraph& g, int variant);

// hip-pull's default light in-degree cap: the 99th percentile in-degree,
// rounded up to a multiple of 32, within [32, 1024]. Targets above it are
// stepped as hubs.
int pull_default_cap(const DevGraph& g);

// Candidates for the search. exact_only drops hip-atomic; allow_sort says
// whether hip-sort may run (single-sample plans only); num_samples > 1 adds
// hip-pull (plans of up to 64 samples). Persistent launches are never
// candidates.
std::vector<snc_plan_config> autotune_candidates(const DevGraph& g, bool exact_only,
                                                 bool allow_sort, int num_samples);

// Steps of the trial window for a sample of num_steps steps.
int autotune_trial_steps(int num_steps);

}  // namespace snc_amd
