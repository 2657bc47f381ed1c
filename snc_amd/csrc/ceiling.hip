// Scattered-access ceilings of the device, for the bench's roofline.
//
// The fused step kernel's delivery is one scattered device-scope atomic per
// synaptic event (a float add into the ring, or an OR into the arrival
// bitmap). Its ceiling is the rate at which the device retires such accesses
// at random addresses, not HBM bandwidth: an access moves a 32-64-byte sector
// for 4 useful bytes. k_scatter_probe measures that rate in the step kernel's
// own launch shape: `blocks` blocks of 512 threads, `ops` accesses per launch
// at hashed addresses over a buffer of `target_bytes`, launched `reps` times
// from one captured HIP graph like the step loop.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "device_graph.hpp"
#include "hip_err.hpp"

namespace snc_amd {
namespace {

constexpr int kProbeThreads = 512;

__device__ __forceinline__ uint32_t probe_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

template <int KIND>
__global__ __launch_bounds__(kProbeThreads) void k_scatter_probe(float* __restrict__ buf,
                                                                 uint32_t words, int64_t ops,
                                                                 int64_t per_block, uint32_t rep,
                                                                 float* __restrict__ sink) {
  const int64_t b0 = (int64_t)blockIdx.x * per_block;
  const int64_t b1 = min(ops, b0 + per_block);
  float acc = 0.0f;
  for (int64_t k = b0 + threadIdx.x; k < b1; k += kProbeThreads) {
    const uint32_t i = probe_hash((uint32_t)k * 0x9E3779B9u + rep * 0x85EBCA6Bu) % words;
    if (KIND == SNC_SCATTER_ATOMIC_ADD) atomicAdd(buf + i, 1.0f);
    if (KIND == SNC_SCATTER_ATOMIC_OR) atomicOr(reinterpret_cast<uint32_t*>(buf) + i, 1u);
    if (KIND == SNC_SCATTER_STORE) buf[i] = 1.0f;
    if (KIND == SNC_SCATTER_LOAD) acc += buf[i];
  }
  if (KIND == SNC_SCATTER_LOAD && acc == -1.0f) sink[0] = acc;  // keeps the loads
}

template <int KIND>
void launch_probe(float* buf, uint32_t words, int64_t ops, int blocks, uint32_t rep,
                  float* sink, hipStream_t st) {
  const int64_t per = (ops + blocks - 1) / blocks;
  hipLaunchKernelGGL(k_scatter_probe<KIND>, dim3(blocks), dim3(kProbeThreads), 0, st, buf, words,
                     ops, per, rep, sink);
}

// The fused step kernel's expansion alone: block b's waves take the fired
// rows of its 512 neurons (rows[bp[b] .. bp[b+1])), one wave per row, lanes
// striding the row four edges at a time, float atomics into ring slot
// (tm + d) % L -- the same code path and placement as the step kernel.
__global__ __launch_bounds__(512) void k_expand_probe(const int32_t* __restrict__ row_ptr,
                                                      const int32_t* __restrict__ post,
                                                      const float* __restrict__ wsig,
                                                      const uint8_t* __restrict__ delay,
                                                      const int32_t* __restrict__ bp,
                                                      const int32_t* __restrict__ rows, int n,
                                                      int L, int tm, float* __restrict__ ring) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = bp[blockIdx.x] + wave; k < bp[blockIdx.x + 1]; k += 8) {
    const int i = rows[k];
    const int rb = row_ptr[i], re = row_ptr[i + 1];
    for (int g = rb; g < re; g += 4 * 64) {
      int d[4], tgt[4];
      float w[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = min(g + lane + u * 64, re - 1);
        d[u] = delay[e];
        tgt[u] = post[e];
        w[u] = wsig[e];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (g + lane + u * 64 >= re) continue;
        int s = tm + d[u];
        if (s >= L) s -= L;
        atomicAdd(ring + (size_t)s * n + tgt[u], w[u]);
      }
    }
  }
}

}  // namespace
}  // namespace snc_amd

using namespace snc_amd;

extern "C" int snc_hip_measure_expansion(const snc_hip_graph* G, const int32_t* fired,
                                         int32_t n_fired, int32_t reps, double* us_per_launch) {
  if (!G || !us_per_launch || n_fired < 0 || (n_fired > 0 && !fired) || reps < 1)
    return fail(SNC_E_INVALID, "bad expansion probe arguments");
  const DevGraph& g = G->g;
  const int n = g.n, R = std::max(1, (n + 511) / 512);
  std::vector<int32_t> bp((size_t)R + 1, 0), rows;
  for (int k = 0; k < n_fired; ++k) {
    if (fired[k] < 0 || fired[k] >= n) return fail(SNC_E_INVALID, "fired neuron out of range");
    bp[fired[k] / 512 + 1]++;
  }
  for (int b = 0; b < R; ++b) bp[b + 1] += bp[b];
  rows.resize((size_t)n_fired);
  {
    std::vector<int32_t> at(bp.begin(), bp.end() - 1);
    for (int k = 0; k < n_fired; ++k) rows[at[fired[k] / 512]++] = fired[k];
  }
  hip_clear();
  int32_t *d_bp = nullptr, *d_rows = nullptr;
  float* ring = nullptr;
  hipStream_t st = nullptr;
  hipGraphExec_t ge = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = SNC_OK;
  hipError_t e = hipMalloc((void**)&d_bp, bp.size() * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&d_rows, std::max<size_t>(rows.size(), 1) * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&ring, (size_t)g.L * n * 4);
  if (e == hipSuccess) e = hipMemcpy(d_bp, bp.data(), bp.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && !rows.empty())
    e = hipMemcpy(d_rows, rows.data(), rows.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(ring, 0, (size_t)g.L * n * 4);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) rc = hip_fail(e, "expansion probe setup");
  if (!rc) {
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_expand_probe, dim3(R), dim3(512), 0, st, g.row_ptr, g.post, g.wsig,
                         g.delay, d_bp, d_rows, n, g.L, r % g.L, ring);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(st, &graph);
    if (e == hipSuccess && graph) {
      e = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
    }
    if (e != hipSuccess) rc = hip_fail(e, "expansion probe capture");
  }
  if (!rc) {
    float best = 1e30f;
    for (int it = 0; it < 4 && !rc; ++it) {
      e = hipEventRecord(e0, st);
      if (e == hipSuccess) e = hipGraphLaunch(ge, st);
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) rc = hip_fail(e, "expansion probe run");
      else if (it > 0) best = std::min(best, ms);
    }
    if (!rc) *us_per_launch = (double)best * 1e3 / reps;
  }
  if (ge) (void)hipGraphExecDestroy(ge);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (d_bp) (void)hipFree(d_bp);
  if (d_rows) (void)hipFree(d_rows);
  if (ring) (void)hipFree(ring);
  return rc;
}

extern "C" int snc_hip_measure_scatter(int32_t kind, int64_t target_bytes, int64_t ops,
                                       int32_t blocks, int32_t reps, double* us_per_launch) {
  if (!us_per_launch || target_bytes < 4 || target_bytes > (int64_t)1 << 34 || ops < 0 ||
      blocks < 1 || reps < 1 || kind < SNC_SCATTER_ATOMIC_ADD || kind > SNC_SCATTER_LOAD)
    return fail(SNC_E_INVALID, "bad scatter probe arguments");
  hip_clear();
  const uint32_t words = (uint32_t)std::min<int64_t>(target_bytes / 4, 0xFFFFFFFFll);
  float* buf = nullptr;
  float* sink = nullptr;
  hipStream_t st = nullptr;
  hipGraphExec_t ge = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = SNC_OK;
  if (hipMalloc((void**)&buf, (size_t)words * 4) != hipSuccess ||
      hipMalloc((void**)&sink, 64) != hipSuccess) {
    hip_clear();
    rc = fail(SNC_E_NOMEM, "scatter probe allocation failed");
  }
  hipError_t e = hipSuccess;
  if (!rc) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (!rc && e == hipSuccess) e = hipMemsetAsync(buf, 0, (size_t)words * 4, st);
  if (!rc && e == hipSuccess) e = hipEventCreate(&e0);
  if (!rc && e == hipSuccess) e = hipEventCreate(&e1);
  if (!rc && e == hipSuccess) e = hipStreamSynchronize(st);
  if (!rc && e == hipSuccess) e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (!rc && e != hipSuccess) rc = hip_fail(e, "scatter probe setup");
  if (!rc) {
    for (int r = 0; r < reps; ++r) {
      switch (kind) {
        case SNC_SCATTER_ATOMIC_ADD:
          launch_probe<SNC_SCATTER_ATOMIC_ADD>(buf, words, ops, blocks, r, sink, st);
          break;
        case SNC_SCATTER_ATOMIC_OR:
          launch_probe<SNC_SCATTER_ATOMIC_OR>(buf, words, ops, blocks, r, sink, st);
          break;
        case SNC_SCATTER_STORE:
          launch_probe<SNC_SCATTER_STORE>(buf, words, ops, blocks, r, sink, st);
          break;
        default:
          launch_probe<SNC_SCATTER_LOAD>(buf, words, ops, blocks, r, sink, st);
          break;
      }
    }
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(st, &graph);
    if (e == hipSuccess && graph) {
      e = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
    }
    if (e != hipSuccess) rc = hip_fail(e, "scatter probe capture");
  }
  if (!rc) {
    float best = 1e30f;
    for (int it = 0; it < 4 && !rc; ++it) {  // the first replay warms up
      e = hipEventRecord(e0, st);
      if (e == hipSuccess) e = hipGraphLaunch(ge, st);
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) rc = hip_fail(e, "scatter probe run");
      else if (it > 0) best = std::min(best, ms);
    }
    if (!rc) *us_per_launch = (double)best * 1e3 / reps;
  }
  if (ge) (void)hipGraphExecDestroy(ge);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (buf) (void)hipFree(buf);
  if (sink) (void)hipFree(sink);
  return rc;
}
