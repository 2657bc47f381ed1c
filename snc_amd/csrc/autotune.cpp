// Per-graph delivery strategy choice (first version: fixed preference).
#include "autotune.hpp"

#include "snc_hip.h"

namespace snc_amd {

int autotune_variant(const DevGraph&, int64_t, int) { return SNC_VARIANT_BUCKET; }

}  // namespace snc_amd
