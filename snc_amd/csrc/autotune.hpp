// Per-graph choice of the delivery strategy and its launch shape.
#pragma once

#include <cstdint>

#include "device_graph.hpp"

namespace snc_amd {

int autotune_variant(const DevGraph& g, int64_t n_events, int num_steps);

}  // namespace snc_amd
