// Shared pieces of the command-line drivers: argument parsing and owning
// wrappers over the C ABI handles. The drivers are plain C++ (g++) programs
// that reach the GPU only through libsnc_hip's C ABI.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "snc_gen.h"
#include "snc_hip.h"
#include "snc_struct.h"

namespace snc_cli {

// Exit with the library's message when a C ABI call fails.
inline void check(int rc, const char* what) {
  if (rc != SNC_OK) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, snc_hip_last_error());
    std::exit(1);
  }
}

inline bool take(int& i, int argc, char** argv, const char* name, std::string& out) {
  if (std::strcmp(argv[i], name) != 0) return false;
  if (i + 1 >= argc) {
    std::fprintf(stderr, "%s needs a value\n", name);
    std::exit(2);
  }
  out = argv[++i];
  return true;
}
inline bool take(int& i, int argc, char** argv, const char* name, int& out) {
  std::string s;
  if (!take(i, argc, argv, name, s)) return false;
  out = std::atoi(s.c_str());
  return true;
}
inline bool take(int& i, int argc, char** argv, const char* name, float& out) {
  std::string s;
  if (!take(i, argc, argv, name, s)) return false;
  out = (float)std::atof(s.c_str());
  return true;
}
inline bool take(int& i, int argc, char** argv, const char* name, uint64_t& out) {
  std::string s;
  if (!take(i, argc, argv, name, s)) return false;
  out = (uint64_t)std::strtoull(s.c_str(), nullptr, 10);
  return true;
}

// hip-* backend names (the reference's Backend numbering, extended).
inline int parse_variant(const std::string& s) {
  if (s == "hip-atomic") return SNC_VARIANT_ATOMIC;
  if (s == "hip-bucket") return SNC_VARIANT_BUCKET;
  if (s == "hip-sort") return SNC_VARIANT_SORT;
  if (s == "hip-auto") return SNC_VARIANT_AUTO;
  if (s == "hip-auto-exact") return SNC_VARIANT_AUTO_EXACT;
  if (s == "hip-pull") return SNC_VARIANT_PULL;
  return -1;
}
inline const char* variant_name(int v) {
  switch (v) {
    case SNC_VARIANT_ATOMIC: return "hip-atomic";
    case SNC_VARIANT_BUCKET: return "hip-bucket";
    case SNC_VARIANT_SORT: return "hip-sort";
    case SNC_VARIANT_AUTO: return "hip-auto";
    case SNC_VARIANT_AUTO_EXACT: return "hip-auto-exact";
    case SNC_VARIANT_PULL: return "hip-pull";
    default: return "?";
  }
}

// An owned SNNGraph-shaped buffer and its view.
struct Graph {
  snc_graph_buf* buf = nullptr;
  snc_graph_view view{};
  Graph() = default;
  explicit Graph(snc_graph_buf* b) : buf(b) { check(snc_graph_buf_view(b, &view), "graph view"); }
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;
  Graph(Graph&& o) noexcept : buf(o.buf), view(o.view) { o.buf = nullptr; }
  ~Graph() { snc_graph_buf_free(buf); }
  int num_output_channels() const {
    int c = 0;
    for (int i = 0; i < view.num_neurons; ++i)
      if (view.role[i] == SNC_ROLE_OUTPUT) c = std::max(c, view.channel[i] + 1);
    return c;
  }
};

// An owned event list.
struct Events {
  snc_event_buf* buf = nullptr;
  const snc_event* data = nullptr;
  int64_t n = 0;
  Events() = default;
  explicit Events(snc_event_buf* b) : buf(b) { check(snc_event_buf_data(b, &data, &n), "events"); }
  Events(const Events&) = delete;
  Events& operator=(const Events&) = delete;
  Events(Events&& o) noexcept : buf(o.buf), data(o.data), n(o.n) { o.buf = nullptr; }
  Events& operator=(Events&& o) noexcept {
    if (this != &o) {
      snc_event_buf_free(buf);
      buf = o.buf;
      data = o.data;
      n = o.n;
      o.buf = nullptr;
    }
    return *this;
  }
  ~Events() { snc_event_buf_free(buf); }
};

// A graph resident in HBM.
struct DeviceGraph {
  snc_hip_graph* h = nullptr;
  explicit DeviceGraph(const snc_graph_view& v) { check(snc_hip_graph_upload(&v, &h), "graph upload"); }
  DeviceGraph(const DeviceGraph&) = delete;
  DeviceGraph& operator=(const DeviceGraph&) = delete;
  ~DeviceGraph() { snc_hip_graph_free(h); }
};

}  // namespace snc_cli
