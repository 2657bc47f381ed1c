// Shared forward kernels: fills and injection staging. hip-atomic and
// hip-bucket run the fused step kernels of step.hip, hip-sort the kernels of
// sort.hip; the persistent form is persist.hip.
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
