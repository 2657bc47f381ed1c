// Shared forward kernels: fills, injection staging, and the integrate/fire
// kernel of the hip-sort strategy (sort.hip). hip-atomic and hip-bucket run
// the fused step kernels of step.hip; the persistent form is persist.hip.
//
// Semantics follow the reference CPU path step for step
// (src/runtime/cpu_backend.cpp:95-183):
//   1. inject the step's external drive,
//   2. integrate/fire every neuron, reading and clearing ring slot t % L,
//   3. decode output spikes,
//   4. expand fired rows into slots (t + d) % L.
// The float expression of step 2 is evaluated as ((v * decay) + ext) + slot
// with explicit round-to-nearest intrinsics, never contracted to an FMA.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "hip_err.hpp"
#include "device_graph.hpp"
#include "fold.hpp"
#include "forward_plan.hpp"
#include "wave.hpp"

namespace snc_amd {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t value, size_t count) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride)
    p[k] = value;
}

// Write the pre-folded injections of one step into ext (one entry per neuron).
__global__ void k_scatter_inj(const int32_t* __restrict__ inj_ptr,
                              const int32_t* __restrict__ inj_nid,
                              const float* __restrict__ inj_amp, int t, float* __restrict__ ext) {
  const int b = inj_ptr[t], e = inj_ptr[t + 1];
  for (int k = b + (int)(blockIdx.x * blockDim.x + threadIdx.x); k < e;
       k += (int)(gridDim.x * blockDim.x))
    ext[inj_nid[k]] = inj_amp[k];
}

// One thread per (sample, step, block): copies a list of at most kInjInline
// injections into the inline table (block-local neuron, amplitude bits), or
// marks the cell as spilled. Cells are kInjInline int2 = 128 bytes, written
// as eight 16-byte stores.
__global__ void k_inj_table(const int32_t* __restrict__ blk_ptr, const int32_t* __restrict__ nid,
                            const float* __restrict__ amp, int2* __restrict__ tab, int T, int R,
                            int blk, int64_t cells) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  const int64_t per_sample = (int64_t)T * R;
  const int64_t smp = i / per_sample, rem = i - smp * per_sample;
  const int64_t t = rem / R, b = rem - t * R;
  const int32_t* bp = blk_ptr + smp * (int64_t)T * (R + 1) + t * (R + 1) + b;
  const int k0 = bp[0], cnt = bp[1] - k0;
  const int base = (int)b * blk;
  int4* out = reinterpret_cast<int4*>(tab + i * kInjInline);
#pragma unroll
  for (int q = 0; q < kInjInline; q += 2) {
    int4 w = make_int4(-1, 0, -1, 0);
    if (cnt <= kInjInline) {
      if (q < cnt) w.x = nid[k0 + q] - base, w.y = __float_as_int(amp[k0 + q]);
      if (q + 1 < cnt) w.z = nid[k0 + q + 1] - base, w.w = __float_as_int(amp[k0 + q + 1]);
    } else if (q + 2 == kInjInline) {
      w.z = kInjSpill;
    }
    out[q / 2] = w;
  }
}

// Appends the wave's fired neurons to the fired list, records the raster
// words and returns the wave's spike count (valid in every lane).
__device__ __forceinline__ int wave_emit_fired(bool f, int i, int32_t* __restrict__ fired,
                                               int32_t* __restrict__ fcnt,
                                               uint32_t* __restrict__ raster_row, int words) {
  const uint64_t m = ballot(f);
  if (m == 0) return 0;
  const int cnt = __popcll(m);
  const int lane = lane_id();
  int base = 0;
  if (lane == 0) base = atomicAdd(fcnt, cnt);
  base = __shfl(base, 0, kWave);
  if (f) fired[base + rank_below(m)] = i;
  if (raster_row && lane == 0) {
    const int w0 = (i - lane) >> 5;  // the wave covers 64 consecutive ids
    if ((uint32_t)m) raster_row[w0] = (uint32_t)m;
    if ((uint32_t)(m >> 32) && w0 + 1 < words) raster_row[w0 + 1] = (uint32_t)(m >> 32);
  }
  return cnt;
}

// Block reduction of per-wave counts into the block's own stats slot.
__device__ __forceinline__ void block_accumulate(int wave_count, unsigned long long* slot) {
  __shared__ int s_sum;
  if (threadIdx.x == 0) s_sum = 0;
  __syncthreads();
  if (lane_id() == 0 && wave_count) atomicAdd(&s_sum, wave_count);
  __syncthreads();
  if (threadIdx.x == 0 && s_sum) *slot += (unsigned long long)s_sum;
}

// Integrate/fire for every neuron (step t, slot tm = t % L), hip-sort: also
// writes the step's fired bitmap, which k_sort_compact turns into the
// ordered fired list.
__global__ __launch_bounds__(256) void k_lif(StepArgs a, int t, int tm) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  bool f = false;
  if (i < a.n) {
    float* slotp = a.ring + (size_t)tm * a.n + i;
    float v = a.v[i];
    int r = a.refr[i];
    f = lif_step(v, r, a.ext[i], *slotp, a.lif);
    a.v[i] = v;
    a.refr[i] = r;
    a.ext[i] = 0.0f;
    *slotp = 0.0f;
    if (f) a.nspk[i] += 1;
  }
  if (a.fbits) {  // every wave rewrites its own words, zeros included
    const uint64_t m = ballot(f);
    const int first = i - lane_id();
    if (lane_id() == 0 && first < a.n) {
      a.fbits[first >> 5] = (uint32_t)m;
      if ((first >> 5) + 1 < a.words) a.fbits[(first >> 5) + 1] = (uint32_t)(m >> 32);
    }
  }
  const int cnt = wave_emit_fired(f, i, a.fired, a.fcnt + (t & 1),
                                  a.raster ? a.raster + (size_t)t * a.words : nullptr, a.words);
  block_accumulate(cnt, a.blk_spikes + blockIdx.x);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

namespace {

int grid_for(int64_t work, int block, int cap) {
  int64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

}  // namespace

int fwd_launch_fill(hipStream_t st, void* p, uint32_t value, size_t words) {
  if (!words) return SNC_OK;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for((int64_t)words, 256, 2048)), dim3(256), 0, st,
                     (uint32_t*)p, value, words);
  return hip_last("fill launch failed");
}

int fwd_launch_lif(const FwdPlanCore& pc, const StepArgs& a, int t) {
  hipLaunchKernelGGL(k_lif, dim3(pc.lif_grid), dim3(256), 0, pc.stream, a, t, t % a.L);
  return hip_last("LIF launch failed");
}

int fwd_launch_first_injection(const FwdPlanCore& pc, const StepArgs& a) {
  if (a.T <= 0) return SNC_OK;
  hipLaunchKernelGGL(k_scatter_inj, dim3(grid_for(pc.max_inj_per_step, 256, 1024)), dim3(256), 0,
                     pc.stream, a.inj_ptr, a.inj_nid, a.inj_amp, 0, a.ext);
  return hip_last("injection launch failed");
}

int fwd_launch_inj_table(hipStream_t st, const StepArgs& a) {
  const int64_t cells = (int64_t)a.B * a.T * a.n_reg_blocks;
  if (cells <= 0) return SNC_OK;
  hipLaunchKernelGGL(k_inj_table, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st,
                     a.inj_blk_ptr, a.blk_nid, a.blk_amp, const_cast<int2*>(a.inj_tab), a.T,
                     a.n_reg_blocks, a.blk, cells);
  return hip_last("injection table launch failed");
}

}  // namespace snc_amd
