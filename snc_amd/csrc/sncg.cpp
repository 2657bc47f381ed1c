// SNCG flat graph files and the exporter's delay spreads (include/snc_gen.h).
//
// The file layout is the reference's interchange format between its C++
// generators and its PyTorch bridge, so it is kept byte for byte:
//   src/bench/snc_export.cpp:10-12, 107-123   writer
//   python/snc/graph.py:11-24                 reader
// little-endian int32: magic 'SNCG' (0x534E4347), N, S, num_input_channels,
// num_output_channels, pre[S], post[S], delays[S], role[N], channel[N].
// Weights and signs are not part of the format.
//
// The channel counts follow SNNGraph::rebuild_io_index
// (src/graph/snn_graph.cpp:13-28): one more than the largest channel of an
// INPUT (resp. OUTPUT) neuron, 0 when there is none.
//
// snc_graph_spread_delays restates snc_export's heterogeneous delays
// (src/bench/snc_export.cpp:81-101): "random" draws std::mt19937_64 seeded
// with seed ^ 0xDE1A0 through std::uniform_int_distribution<int>(1, dmax)
// (libstdc++, the reference's own draws), "distance" maps a synapse's span in
// the 1-D layer embedding of a 2-layer [A, B] graph onto [1, dmax].

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "common.hpp"
#include "snc_gen.h"

namespace snc_amd {
namespace {

constexpr int32_t kSncgMagic = 0x534E4347;

void io_channels(const snc_graph_view& g, int32_t* n_in, int32_t* n_out) {
  int32_t max_in = -1, max_out = -1;
  for (int i = 0; i < g.num_neurons; ++i) {
    if (g.role[i] == 1) max_in = std::max(max_in, g.channel[i]);
    else if (g.role[i] == 2) max_out = std::max(max_out, g.channel[i]);
  }
  *n_in = max_in + 1;
  *n_out = max_out + 1;
}

struct File {
  std::FILE* f = nullptr;
  explicit File(std::FILE* p) : f(p) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

bool put_i32(std::FILE* f, const int32_t* p, size_t count) {
  // Little-endian on disk; x86-64 and the GPU hosts are little-endian.
  return std::fwrite(p, sizeof(int32_t), count, f) == count;
}

bool get_i32(std::FILE* f, int32_t* p, size_t count) {
  return std::fread(p, sizeof(int32_t), count, f) == count;
}

}  // namespace
}  // namespace snc_amd

using namespace snc_amd;

extern "C" {

int snc_graph_io_channels(const snc_graph_view* g, int32_t* n_in, int32_t* n_out) {
  if (!g || !n_in || !n_out || (g->num_neurons > 0 && (!g->role || !g->channel)))
    return fail(SNC_E_INVALID, "null argument");
  io_channels(*g, n_in, n_out);
  return SNC_OK;
}

int snc_graph_write_sncg(const snc_graph_view* g, const char* path) {
  if (!g || !path) return fail(SNC_E_INVALID, "null argument");
  const std::string err = validate_view(*g);
  if (!err.empty()) return fail(SNC_E_INVALID, "graph validation failed: " + err);
  const int32_t n = g->num_neurons, m = g->num_synapses;
  int32_t head[5] = {kSncgMagic, n, m, 0, 0};
  io_channels(*g, &head[3], &head[4]);
  std::vector<int32_t> pre((size_t)m), role((size_t)n);
  for (int i = 0; i < n; ++i) {
    for (int e = g->row_ptr[i]; e < g->row_ptr[i + 1]; ++e) pre[e] = i;
    role[i] = (int32_t)g->role[i];
  }
  File f(std::fopen(path, "wb"));
  if (!f.f) return fail(SNC_E_INVALID, std::string("cannot open ") + path);
  const bool ok = put_i32(f.f, head, 5) && put_i32(f.f, pre.data(), (size_t)m) &&
                  put_i32(f.f, g->post_ids, (size_t)m) && put_i32(f.f, g->delays, (size_t)m) &&
                  put_i32(f.f, role.data(), (size_t)n) && put_i32(f.f, g->channel, (size_t)n);
  if (!ok) return fail(SNC_E_INVALID, std::string("write failed: ") + path);
  return SNC_OK;
}

int snc_graph_read_sncg(const char* path, float weight, snc_graph_buf** out, int32_t* n_in,
                        int32_t* n_out) {
  if (!path || !out) return fail(SNC_E_INVALID, "null argument");
  *out = nullptr;
  File f(std::fopen(path, "rb"));
  if (!f.f) return fail(SNC_E_INVALID, std::string("cannot open ") + path);
  int32_t head[5];
  if (!get_i32(f.f, head, 5)) return fail(SNC_E_INVALID, "SNCG header truncated");
  if (head[0] != kSncgMagic) return fail(SNC_E_INVALID, "not an SNCG file (bad magic)");
  const int32_t n = head[1], m = head[2];
  if (n < 0 || m < 0) return fail(SNC_E_INVALID, "SNCG sizes negative");
  std::fseek(f.f, 0, SEEK_END);
  const long long bytes = std::ftell(f.f);
  if (bytes != 4LL * (5 + 3LL * m + 2LL * n))
    return fail(SNC_E_INVALID, "SNCG file size does not match N and S");
  std::fseek(f.f, 20, SEEK_SET);
  std::vector<int32_t> pre((size_t)m), post((size_t)m), delay((size_t)m), role((size_t)n),
      channel((size_t)n);
  if (!get_i32(f.f, pre.data(), (size_t)m) || !get_i32(f.f, post.data(), (size_t)m) ||
      !get_i32(f.f, delay.data(), (size_t)m) || !get_i32(f.f, role.data(), (size_t)n) ||
      !get_i32(f.f, channel.data(), (size_t)n))
    return fail(SNC_E_INVALID, "SNCG body truncated");
  for (int32_t k = 0; k < m; ++k) {
    if (pre[k] < 0 || pre[k] >= n) return fail(SNC_E_INVALID, "SNCG pre id out of range");
    if (post[k] < 0 || post[k] >= n) return fail(SNC_E_INVALID, "post id out of range");
    if (delay[k] < 1) return fail(SNC_E_INVALID, "delay < 1");
  }
  for (int32_t i = 0; i < n; ++i)
    if (role[i] < 0 || role[i] > 2) return fail(SNC_E_INVALID, "SNCG role out of range");
  // Rows in file order within each pre (a stable counting sort; files written
  // by the exporter are already row-major).
  auto* b = new snc_graph_buf();
  GraphArrays& g = b->g;
  g.n = n;
  g.row_ptr.assign((size_t)n + 1, 0);
  for (int32_t k = 0; k < m; ++k) g.row_ptr[pre[k] + 1]++;
  for (int32_t i = 0; i < n; ++i) g.row_ptr[i + 1] += g.row_ptr[i];
  g.post.resize((size_t)m);
  g.delay.resize((size_t)m);
  g.weight.assign((size_t)m, weight);
  std::vector<int32_t> fill(g.row_ptr.begin(), g.row_ptr.end() - 1);
  for (int32_t k = 0; k < m; ++k) {
    const int32_t e = fill[pre[k]]++;
    g.post[e] = post[k];
    g.delay[e] = delay[k];
  }
  g.sign.assign((size_t)n, 1);
  g.role.resize((size_t)n);
  for (int32_t i = 0; i < n; ++i) g.role[i] = (uint8_t)role[i];
  g.channel = std::move(channel);
  if (n_in) *n_in = head[3];
  if (n_out) *n_out = head[4];
  *out = b;
  return SNC_OK;
}

int snc_graph_spread_delays(snc_graph_buf* b, int32_t mode, int32_t delay_max,
                            const int32_t* layers, int32_t n_layers, uint64_t seed) {
  if (!b) return fail(SNC_E_INVALID, "null graph");
  if (mode != SNC_DELAYS_RANDOM && mode != SNC_DELAYS_DISTANCE)
    return fail(SNC_E_INVALID, "unknown delay mode");
  if (delay_max <= 1) return SNC_OK;  // the exporter leaves delays alone
  GraphArrays& g = b->g;
  const int32_t m = (int32_t)g.post.size();
  if (mode == SNC_DELAYS_DISTANCE && layers && n_layers >= 2) {
    const int32_t A = layers[0], B = layers[1];
    const double denom = std::max(1, std::max(A, B) - 1);
    for (int32_t i = 0; i < g.n; ++i)
      for (int32_t e = g.row_ptr[i]; e < g.row_ptr[i + 1]; ++e) {
        const double d = std::abs(i - (g.post[e] - A)) / denom;
        const int del = 1 + (int)std::lround(d * (delay_max - 1));
        g.delay[e] = std::max(1, std::min(delay_max, del));
      }
    return SNC_OK;
  }
  std::mt19937_64 rng(seed ^ 0xDE1A0ULL);
  std::uniform_int_distribution<int> dd(1, delay_max);
  for (int32_t e = 0; e < m; ++e) g.delay[e] = dd(rng);
  return SNC_OK;
}

}  // extern "C"
