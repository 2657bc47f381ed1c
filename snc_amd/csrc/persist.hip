// Persistent cooperative forward kernel (placeholder until implemented).
#include "persist.hpp"

#include "common.hpp"

namespace snc_amd {

bool persist_choose(const DevGraph&, int, PersistConfig&) { return false; }

int persist_launch(const DevGraph&, const FwdPlanCore&, const StepArgs&, const PersistConfig&) {
  return fail(SNC_E_UNSUPPORTED, "persistent forward not built");
}

}  // namespace snc_amd
