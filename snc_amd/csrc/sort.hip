// hip-sort strategy (placeholder until implemented).
#include "common.hpp"
#include "forward_plan.hpp"

namespace snc_amd {

size_t sort_hist_words(int n) { return (size_t)n + 1; }

int sort_launch_step(const FwdPlanCore&, const StepArgs&, int) {
  return fail(SNC_E_UNSUPPORTED, "hip-sort not built");
}

}  // namespace snc_amd
