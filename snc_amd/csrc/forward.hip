// HIP forward backend for gfx950: LIF membrane update, conduction-delay ring
// and spike delivery (hip-atomic, hip-bucket, hip-sort).
//
// Semantics follow the reference CPU path step for step
// (src/runtime/cpu_backend.cpp:95-183):
//   1. inject the step's external drive,
//   2. integrate/fire every neuron, reading and clearing ring slot t % L,
//   3. decode output spikes,
//   4. expand fired rows into slots (t + d) % L.
// The float expression of step 2 is evaluated as ((v * decay) + ext) + slot
// with explicit round-to-nearest intrinsics, never contracted to an FMA.
//
// Delivery strategies:
//   hip-atomic  one wave per fired row streams post/weight/delay with
//               coalesced loads and adds into the ring with hardware float
//               atomics. Addition order is whatever the L2 sees: voltages
//               match the CPU to ~1e-6 relative, not bit for bit.
//   hip-bucket  every target owns a bucket -- its column of incoming
//               synapses, ordered (pre, CSR position) exactly as the CPU adds
//               them -- and a presence bitmap over that column. Expansion
//               only sets bits (order-free atomicOr); at the start of the next
//               step the owning lane walks its bits in ascending order and
//               folds the values into its ring slots. The per-(target, slot)
//               sum is therefore the CPU's sum, bit for bit. Targets with a
//               large in-degree are folded by a whole wave (k_fold_heavy).
//   hip-sort    see sort.hip.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
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

// Integrate/fire for every neuron (step t, slot tm = t % L). With FOLD the
// lane first folds its target's pending bucket (hip-bucket).
template <bool FOLD>
__global__ __launch_bounds__(256) void k_lif(StepArgs a, int t, int tm) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  bool f = false;
  if (i < a.n) {
    if (FOLD && t > 0 && a.dirty[i]) {
      fold_light(i, a.n, a.L, tm == 0 ? a.L - 1 : tm - 1, a.ccol, a.cval, a.cdel, a.bits, a.ring);
      a.dirty[i] = 0;
    }
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

// Wave-per-heavy-target ordered fold (hip-bucket), one wave per block.
__global__ __launch_bounds__(64) void k_fold_heavy(StepArgs a, int tprev) {
  __shared__ float s_val[kHeavyChunk];
  __shared__ uint8_t s_slot[kHeavyChunk];
  fold_heavy_wave<kHeavyChunk>(a.heavy_ids[blockIdx.x], tprev, a.L, a.n, a.ccol, a.cval, a.cdel,
                               a.bits, a.dirty, a.ring, s_val, s_slot);
}

// Expansion of step t's fired rows, one wave per row, lanes striding the row
// so post/weight/delay loads coalesce. BUCKET sets presence bits instead of
// adding. Also stages step t+1's injections and clears its fired counter.
template <bool BUCKET>
__global__ __launch_bounds__(256) void k_expand(StepArgs a, int t, int tm) {
  const int gtid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int gsize = (int)(gridDim.x * blockDim.x);
  if (t + 1 < a.T) {
    const int b = a.inj_ptr[t + 1], e = a.inj_ptr[t + 2];
    for (int k = b + gtid; k < e; k += gsize) a.ext[a.inj_nid[k]] = a.inj_amp[k];
    if (gtid == 0) a.fcnt[(t + 1) & 1] = 0;
  }
  const int nf = a.fcnt[t & 1];
  const int lane = lane_id();
  const int wave = gtid >> 6, nwaves = gsize >> 6;
  int ev = 0;
  for (int k = wave; k < nf; k += nwaves) {
    const int i = a.fired[k];
    const int rb = a.row_ptr[i], re = a.row_ptr[i + 1];
    for (int e = rb + lane; e < re; e += kWave) {
      if (BUCKET) {
        const int p = a.cpos[e];
        atomicOr(a.bits + (p >> 5), 1u << (p & 31));
        a.dirty[a.post[e]] = 1;
      } else {
        int s = tm + (int)a.delay[e];
        if (s >= a.L) s -= a.L;
        atomicAdd(a.ring + (size_t)s * a.n + a.post[e], a.wsig[e]);
      }
    }
    ev += re - rb;
  }
  block_accumulate(lane == 0 ? ev : 0, a.blk_events + blockIdx.x);
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
  return hipGetLastError() == hipSuccess ? SNC_OK : fail(SNC_E_HIP, "fill launch failed");
}

int fwd_launch_heavy(const FwdPlanCore& pc, const StepArgs& a, int t) {
  const int tm = t % a.L;
  if (pc.variant == SNC_VARIANT_BUCKET && t > 0 && pc.n_heavy > 0)
    hipLaunchKernelGGL(k_fold_heavy, dim3(pc.n_heavy), dim3(64), 0, pc.stream, a,
                       tm == 0 ? a.L - 1 : tm - 1);
  return hipGetLastError() == hipSuccess ? SNC_OK : fail(SNC_E_HIP, "heavy-fold launch failed");
}

int fwd_launch_lif(const FwdPlanCore& pc, const StepArgs& a, int t) {
  const int tm = t % a.L;
  if (pc.variant == SNC_VARIANT_BUCKET)
    hipLaunchKernelGGL(k_lif<true>, dim3(pc.lif_grid), dim3(256), 0, pc.stream, a, t, tm);
  else
    hipLaunchKernelGGL(k_lif<false>, dim3(pc.lif_grid), dim3(256), 0, pc.stream, a, t, tm);
  return hipGetLastError() == hipSuccess ? SNC_OK : fail(SNC_E_HIP, "LIF launch failed");
}

int fwd_launch_expand(const FwdPlanCore& pc, const StepArgs& a, int t) {
  const int tm = t % a.L;
  if (pc.variant == SNC_VARIANT_BUCKET)
    hipLaunchKernelGGL(k_expand<true>, dim3(pc.expand_grid), dim3(256), 0, pc.stream, a, t, tm);
  else
    hipLaunchKernelGGL(k_expand<false>, dim3(pc.expand_grid), dim3(256), 0, pc.stream, a, t, tm);
  return hipGetLastError() == hipSuccess ? SNC_OK : fail(SNC_E_HIP, "expand launch failed");
}

// Launch one step of the per-step chain on the plan's stream.
int fwd_launch_step(const FwdPlanCore& pc, const StepArgs& a, int t) {
  int rc = fwd_launch_heavy(pc, a, t);
  if (!rc) rc = fwd_launch_lif(pc, a, t);
  if (!rc) rc = fwd_launch_expand(pc, a, t);
  return rc;
}

int fwd_launch_first_injection(const FwdPlanCore& pc, const StepArgs& a) {
  if (a.T <= 0) return SNC_OK;
  hipLaunchKernelGGL(k_scatter_inj, dim3(grid_for(pc.max_inj_per_step, 256, 1024)), dim3(256), 0,
                     pc.stream, a.inj_ptr, a.inj_nid, a.inj_amp, 0, a.ext);
  return hipGetLastError() == hipSuccess ? SNC_OK : fail(SNC_E_HIP, "injection launch failed");
}

}  // namespace snc_amd
