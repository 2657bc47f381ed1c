// Host-side helpers shared by the library's translation units.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "snc_hip.h"

namespace snc_amd {

// Thread-local last error, reported through snc_hip_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// SNNGraph-shaped owned arrays (snn_graph.hpp:43-77).
struct GraphArrays {
  int32_t n = 0;
  std::vector<int32_t> row_ptr, post, delay, channel;
  std::vector<float> weight;
  std::vector<int8_t> sign;
  std::vector<uint8_t> role;

  snc_graph_view view() const {
    snc_graph_view v;
    v.num_neurons = n;
    v.num_synapses = (int32_t)post.size();
    v.row_ptr = row_ptr.data();
    v.post_ids = post.data();
    v.weights = weight.data();
    v.delays = delay.data();
    v.sign = sign.data();
    v.role = role.data();
    v.channel = channel.data();
    return v;
  }
};

// One synapse before CSR packing.
struct EdgeRec {
  int32_t pre, post;
  float w;
  int32_t d;
};

// Pack edges into CSR over outgoing synapses; rows sorted by pre, and within
// a row by (post, insertion order) when sort_rows is set, else insertion order.
void pack_csr(int32_t n, std::vector<EdgeRec>& edges, bool sort_rows, GraphArrays& g);

// SNNGraph::validate (snn_graph.cpp:30-56). Empty string on success.
std::string validate_view(const snc_graph_view& g);

}  // namespace snc_amd

// Opaque handle of include/snc_gen.h: owned SNNGraph arrays.
struct snc_graph_buf {
  snc_amd::GraphArrays g;
};
