// Checks the reference-side binding of libsnc_hip (the patch and the three
// new files of the integration guide) once compiled into the reference's own
// runtime: the enum numbering against SNC_VARIANT_*, the event layout, the
// backend names, and the fallback contract of forward() and of the trainer
// (with no device: the CPU result, reported as Backend::Cpu; with a device:
// the HIP result of the bit-exact strategy, equal to the CPU's).
//
// Built and run by tests/test_integration_build.py; prints "binding ok".

#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "snc/dataset.hpp"
#include "snc/encoders.hpp"
#include "snc/hip_trainer.hpp"
#include "snc/runtime.hpp"
#include "snc/snn_graph.hpp"
#include "snc/trainer.hpp"
#include "snc_hip.h"

using snc::Backend;

static_assert(static_cast<int>(Backend::HipAtomic) == SNC_VARIANT_ATOMIC, "hip-atomic number");
static_assert(static_cast<int>(Backend::HipBucket) == SNC_VARIANT_BUCKET, "hip-bucket number");
static_assert(static_cast<int>(Backend::HipSort) == SNC_VARIANT_SORT, "hip-sort number");
static_assert(static_cast<int>(Backend::HipAuto) == SNC_VARIANT_AUTO, "hip-auto number");
static_assert(static_cast<int>(Backend::HipAutoExact) == SNC_VARIANT_AUTO_EXACT,
              "hip-auto-exact number");
static_assert(static_cast<int>(Backend::HipPull) == SNC_VARIANT_PULL, "hip-pull number");
static_assert(static_cast<int>(Backend::CudaSort) < SNC_VARIANT_ATOMIC,
              "the HIP values follow the reference's own");
static_assert(sizeof(snc::InputEvent) == sizeof(snc_event), "event size");
static_assert(offsetof(snc::InputEvent, sample_id) == offsetof(snc_event, sample_id) &&
                  offsetof(snc::InputEvent, channel_id) == offsetof(snc_event, channel_id) &&
                  offsetof(snc::InputEvent, time) == offsetof(snc_event, time) &&
                  offsetof(snc::InputEvent, amplitude) == offsetof(snc_event, amplitude),
              "event field offsets");

static int failures = 0;
#define CHECK(cond)                                               \
  do {                                                            \
    if (!(cond)) {                                                \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                 \
    }                                                             \
  } while (0)

int main() {
  // names: every HIP backend parses, prints back and is HIP, not CUDA
  const char* names[] = {"hip-atomic", "hip-bucket", "hip-sort",
                         "hip-auto",   "hip-auto-exact", "hip-pull"};
  for (int k = 0; k < 6; ++k) {
    Backend b{};
    CHECK(snc::parse_backend(names[k], b));
    CHECK(static_cast<int>(b) == SNC_VARIANT_ATOMIC + k);
    CHECK(std::string(snc::backend_name(b)) == names[k]);
    CHECK(snc::backend_is_hip(b));
    CHECK(!snc::backend_is_cuda(b));
  }
  for (Backend b : {Backend::Cpu, Backend::OpenMP, Backend::CudaAtomic, Backend::CudaBucket,
                    Backend::CudaSort})
    CHECK(!snc::backend_is_hip(b));
  Backend none{};
  CHECK(!snc::parse_backend("hip", none));

  // forward: a small recurrent-free sparse graph driven by direct currents
  const snc::SNNGraph g = snc::make_random_sparse({12, 40, 5}, 300, 0.45f, 2, 0.2f, 11);
  std::string err;
  CHECK(g.validate(err));
  std::vector<float> feat(12);
  for (int c = 0; c < 12; ++c) feat[c] = 0.1f + 0.07f * c;
  snc::EncoderParams ep;
  ep.gain = 1.3f;
  const std::vector<snc::InputEvent> ev =
      snc::encode(snc::Encoder::Direct, feat.data(), 12, 40, ep, 5);
  snc::LIFParams lif;
  const snc::ForwardResult cpu = snc::forward(g, ev, 40, lif, Backend::Cpu);
  CHECK(cpu.stats.spikes > 0);
  const bool dev = snc::hip::available();
  CHECK(dev == (snc_hip_available() == 1));
  for (Backend b : {Backend::HipBucket, Backend::HipSort, Backend::HipAutoExact}) {
    const snc::ForwardResult r = snc::forward(g, ev, 40, lif, b);
    if (!dev)
      CHECK(r.used_backend == Backend::Cpu);
    else if (b == Backend::HipAutoExact)
      CHECK(r.used_backend == Backend::HipBucket || r.used_backend == Backend::HipSort);
    else
      CHECK(r.used_backend == b);
    CHECK(r.logits == cpu.logits);
    CHECK(r.stats.spikes == cpu.stats.spikes);
    CHECK(r.stats.synaptic_events == cpu.stats.synaptic_events);
  }

  // trainer: without a device hiptrain declines and the caller keeps the CPU
  // Trainer, as the reference does with cudatrain
  snc::TrainConfig cfg;
  cfg.num_steps = 20;
  cfg.encoder = snc::Encoder::Direct;
  const snc::Dataset data = snc::make_synthetic(24, 12, 5, 0.2f, 3);
  CHECK(snc::hiptrain::available(cfg) == dev);
  snc::HipTrainSession* s = snc::hiptrain::create(g, cfg, 8);
  CHECK((s != nullptr) == dev);
  if (s) {
    const snc::EpochStats st = snc::hiptrain::train_epoch(s, data, data, 0);
    CHECK(st.spikes > 0);
    CHECK(snc::hiptrain::get_weights(s).size() == (size_t)g.num_synapses());
    snc::hiptrain::destroy(s);
  } else {
    CHECK(snc::hiptrain::get_weights(nullptr).empty());
    CHECK(snc::hiptrain::evaluate(nullptr, data) == 0.0);
    snc::hiptrain::destroy(nullptr);
    snc::Trainer t(g, cfg);
    const snc::EpochStats st = t.train_epoch(data, data, 0);
    CHECK(st.spikes > 0);
  }

  if (failures) return 1;
  std::printf("binding ok (device %s)\n", dev ? "present" : "absent");
  return 0;
}
