// Autotuner candidate generation from in-degree statistics.
//
// What the statistics decide:
//   * hub thresholds of hip-bucket: the per-lane fold of the step kernel
//     handles columns of up to 512 positions; lower thresholds move more
//     targets to the block-per-hub kernel. A threshold is tried only when it
//     changes the hub set (the number of targets above it differs from the
//     next larger threshold) and the hub count stays within kMaxHubs (one
//     block per hub per step).
//   * hip-atomic (unless only bit-exact strategies are wanted) at every block
//     size: its cost is the expansion's atomics, whose contention on hubs is
//     measured by the trial rather than modelled.
//   * hip-sort is tried once (its cost follows the events, not the shape),
//     for single-sample plans.
//   * no persistent (cooperative whole-run) variant: its grid-wide barrier
//     per step cost more than a graph-launched step on every measured graph
//     (delivery stress: 29 vs 7.3 us per step with float atomics, 55 vs 12.1
//     with buckets), so it runs only when a plan asks for it explicitly.
//   * hip-pull (bit-exact) for batched plans of up to 64 samples, at the
//     99th-percentile in-degree cap with 256- and 128-target blocks (the
//     larger won with 64 samples on the delivery-stress graph, the smaller
//     with 8), and at twice the cap.
#include "autotune.hpp"
#include "pull.hpp"

#include <algorithm>

namespace snc_amd {

namespace {
constexpr int kLightLimit = 512;
constexpr int kMaxHubs = 16384;
constexpr int kThresholds[] = {64, 128, 256, 512};
}  // namespace

int pull_default_cap(const DevGraph& g) {
  if (g.in_degree.empty()) return 32;
  std::vector<int32_t> d(g.in_degree);
  const size_t q = std::min(d.size() - 1, d.size() * 99 / 100);
  std::nth_element(d.begin(), d.begin() + (ptrdiff_t)q, d.end());
  const int cap = (d[q] + 31) / 32 * 32;
  return std::min(1024, std::max(32, cap));
}

snc_plan_config default_plan_config(const DevGraph& g, int variant) {
  snc_plan_config c{};
  c.variant = variant;
  c.block_neurons = 256;
  c.heavy_threshold = variant == SNC_VARIANT_PULL ? pull_default_cap(g) : kLightLimit;
  c.persistent = 0;
  return c;
}

std::vector<snc_plan_config> autotune_candidates(const DevGraph& g, bool exact_only,
                                                 bool allow_sort, int num_samples) {
  std::vector<snc_plan_config> out;
  auto add = [&](int variant, int block, int thr, int persistent) {
    if ((int)out.size() >= SNC_TUNE_MAX_CANDIDATES) return;
    snc_plan_config c{variant, block, thr, persistent};
    out.push_back(c);
  };
  auto hubs_above = [&](int thr) {
    int64_t k = 0;
    for (int d : g.in_degree) k += d > thr;
    return k;
  };

  // hip-bucket: hub thresholds that change the hub set, largest first.
  std::vector<int> thr;
  int64_t prev = -1;
  for (int q = (int)(sizeof(kThresholds) / sizeof(int)) - 1; q >= 0; --q) {
    const int64_t h = hubs_above(kThresholds[q]);
    if (h > kMaxHubs) break;
    if (h != prev) thr.push_back(kThresholds[q]);
    prev = h;
  }
  if (thr.empty() || thr.front() != kLightLimit) thr.insert(thr.begin(), kLightLimit);
  for (int t : thr)
    for (int blk : {128, 256, 512}) add(SNC_VARIANT_BUCKET, blk, t, 0);

  if (!exact_only)
    for (int blk : {128, 256, 512}) add(SNC_VARIANT_ATOMIC, blk, kLightLimit, 0);
  if (allow_sort) add(SNC_VARIANT_SORT, 256, kLightLimit, 0);
  if (num_samples > 1 && num_samples <= kPullMaxSamples) {
    const int cap = pull_default_cap(g);
    add(SNC_VARIANT_PULL, 256, cap, 0);
    add(SNC_VARIANT_PULL, 128, cap, 0);
    if (std::min(1024, 2 * cap) != cap) add(SNC_VARIANT_PULL, 256, std::min(1024, 2 * cap), 0);
  }
  return out;
}

int autotune_trial_steps(int num_steps) { return std::max(1, std::min(num_steps, 128)); }

}  // namespace snc_amd
