// snc_hip_export: write a generated graph as an SNCG file, with the options of
// the reference's snc_export (src/bench/snc_export.cpp:45-70), so the same
// command line gives the same bytes:
//
//   snc_hip_export --structure static-snc --layers 784,256,10
//                  --synapse-budget 40000 --seed 1 --out graph.bin
//
// Structures: static-snc (alias dynamic-snc), random-sparse, dense; all with
// weight 0.5 as in the exporter. --delay-max > 1 spreads delays by
// --delay-mode random | distance. Host-only: no GPU call is made.

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cli_common.hpp"

using namespace snc_cli;

namespace {

std::vector<int32_t> parse_ints(const std::string& spec) {
  std::vector<int32_t> v;
  size_t i = 0;
  while (i < spec.size()) {
    size_t j = spec.find(',', i);
    if (j == std::string::npos) j = spec.size();
    if (j > i) v.push_back(std::atoi(spec.substr(i, j - i).c_str()));
    i = j + 1;
  }
  return v;
}

}  // namespace

int main(int argc, char** argv) {
  std::string structure = "static-snc", layers_spec = "784,256,10", out = "graph.bin";
  std::string delay_mode = "random";
  int budget = 40000, delay = 1, delay_max = 1;
  float inhib = 0.2f;
  uint64_t seed = 1;
  for (int i = 1; i < argc; ++i) {
    if (std::strcmp(argv[i], "--help") == 0) {
      std::printf(
          "usage: snc_hip_export [--structure static-snc|dynamic-snc|random-sparse|dense]\n"
          "       [--layers 784,256,10] [--synapse-budget 40000] [--delay 1] [--delay-max 1]\n"
          "       [--delay-mode random|distance] [--inhib 0.2] [--seed 1] [--out graph.bin]\n");
      return 0;
    }
    if (take(i, argc, argv, "--structure", structure) ||
        take(i, argc, argv, "--layers", layers_spec) || take(i, argc, argv, "--out", out) ||
        take(i, argc, argv, "--synapse-budget", budget) ||
        take(i, argc, argv, "--delay", delay) || take(i, argc, argv, "--delay-max", delay_max) ||
        take(i, argc, argv, "--delay-mode", delay_mode) ||
        take(i, argc, argv, "--inhib", inhib) || take(i, argc, argv, "--seed", seed))
      continue;
    std::fprintf(stderr, "unknown arg: %s\n", argv[i]);
    return 2;
  }
  const std::vector<int32_t> layers = parse_ints(layers_spec);
  const int32_t nl = (int32_t)layers.size();
  snc_graph_buf* b = nullptr;
  int rc;
  if (structure == "dense")
    rc = snc_gen_dense(layers.data(), nl, 0.5f, delay, &b);
  else if (structure == "random-sparse")
    rc = snc_gen_random_sparse(layers.data(), nl, budget, 0.5f, delay, inhib, seed, &b);
  else if (structure == "static-snc" || structure == "dynamic-snc")
    rc = snc_gen_static_snc(layers.data(), nl, budget, 0.5f, delay, inhib, seed, &b);
  else {
    std::fprintf(stderr, "unknown structure: %s\n", structure.c_str());
    return 2;
  }
  check(rc, "graph generation");
  Graph g(b);
  if (snc_graph_validate(&g.view) != 1) {
    std::fprintf(stderr, "invalid graph: %s\n", snc_hip_last_error());
    return 1;
  }
  const int mode = delay_mode == "distance" ? SNC_DELAYS_DISTANCE : SNC_DELAYS_RANDOM;
  check(snc_graph_spread_delays(g.buf, mode, delay_max, layers.data(), nl, seed), "delays");
  check(snc_graph_write_sncg(&g.view, out.c_str()), "SNCG write");
  int32_t n_in = 0, n_out = 0;
  check(snc_graph_io_channels(&g.view, &n_in, &n_out), "io channels");
  std::printf("wrote %s: %s  N=%d S=%d in=%d out=%d\n", out.c_str(), structure.c_str(),
              g.view.num_neurons, g.view.num_synapses, n_in, n_out);
  return 0;
}
