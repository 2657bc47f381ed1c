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

__device__ __forceinline__ bool lif_step(float& v, int& refr, float ext, float slot,
                                         const LifK& p) {
  if (refr > 0) {
    --refr;
    if (!p.subtract) v = p.reset;
    return false;
  }
  const float x = __fadd_rn(__fadd_rn(__fmul_rn(v, p.decay), ext), slot);
  if (x >= p.threshold) {
    v = p.subtract ? __fsub_rn(x, p.threshold) : p.reset;
    refr = p.refractory;
    return true;
  }
  v = x;
  return false;
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

// Ordered fold of one light target's pending deposits (made at step t-1) into
// its ring slots: walk the presence bits of its column in ascending order and
// add each value to slot (t-1 + delay) % L. A running accumulator is kept per
// slot run; switching slots writes it back, so every (target, slot) sum is
// the sequential sum in column order.
__device__ __forceinline__ void fold_light(int i, int n, int L, int tprev,
                                           const int32_t* __restrict__ ccol,
                                           const float* __restrict__ cval,
                                           const uint8_t* __restrict__ cdel,
                                           uint32_t* __restrict__ bits, float* __restrict__ ring) {
  const int w0 = ccol[i] >> 5, w1 = ccol[i + 1] >> 5;
  int cur = -1;
  float acc = 0.0f;
  for (int w = w0; w < w1; ++w) {
    uint32_t b = bits[w];
    if (!b) continue;
    bits[w] = 0u;
    while (b) {
      const int p = (w << 5) | __builtin_ctz(b);
      b &= b - 1u;
      int s = tprev + (int)cdel[p];
      if (s >= L) s -= L;
      if (s != cur) {
        if (cur >= 0) ring[(size_t)cur * n + i] = acc;
        cur = s;
        acc = ring[(size_t)s * n + i];
      }
      acc = __fadd_rn(acc, cval[p]);
    }
  }
  if (cur >= 0) ring[(size_t)cur * n + i] = acc;
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
  const int cnt = wave_emit_fired(f, i, a.fired, a.fcnt + (t & 1),
                                  a.raster ? a.raster + (size_t)t * a.words : nullptr, a.words);
  block_accumulate(cnt, a.blk_spikes + blockIdx.x);
}

// Wave-per-heavy-target ordered fold (hip-bucket). The wave gathers the set
// positions of the target's column in ascending order into LDS, then lane s
// folds, in that order, every value bound for ring slot s (lanes cover slots
// s, s+64, ... for rings longer than 64).
constexpr int kHeavyChunk = 2048;
constexpr int kMaxSlotsPerLane = 4;  // rings up to 256 slots

__global__ __launch_bounds__(64) void k_fold_heavy(StepArgs a, int tprev) {
  __shared__ float s_val[kHeavyChunk];
  __shared__ uint8_t s_slot[kHeavyChunk];
  const int h = a.heavy_ids[blockIdx.x];
  if (!a.dirty[h]) return;
  const int lane = lane_id();
  const int L = a.L, n = a.n;
  float acc[kMaxSlotsPerLane];
#pragma unroll
  for (int k = 0; k < kMaxSlotsPerLane; ++k) {
    const int s = lane + k * kWave;
    acc[k] = s < L ? a.ring[(size_t)s * n + h] : 0.0f;
  }
  int fill = 0;
  auto drain = [&]() {
    for (int q = 0; q < fill; ++q) {
      const int s = s_slot[q];
      const float x = s_val[q];
#pragma unroll
      for (int k = 0; k < kMaxSlotsPerLane; ++k)
        if (s == lane + k * kWave) acc[k] = __fadd_rn(acc[k], x);
    }
    fill = 0;
  };
  const int w0 = a.ccol[h] >> 5, w1 = a.ccol[h + 1] >> 5;
  for (int wb = w0; wb < w1; wb += kWave) {
    const int w = wb + lane;
    uint32_t b = 0;
    if (w < w1) {
      b = a.bits[w];
      if (b) a.bits[w] = 0u;
    }
    const int c = __popc(b);
    const int incl = wave_inclusive_scan(c);
    const int total = __shfl(incl, kWave - 1, kWave);
    if (total == 0) continue;
    if (fill + total > kHeavyChunk) {
      __syncthreads();
      drain();
      __syncthreads();
    }
    int o = fill + incl - c;
    while (b) {
      const int p = (w << 5) | __builtin_ctz(b);
      b &= b - 1u;
      int s = tprev + (int)a.cdel[p];
      if (s >= L) s -= L;
      s_slot[o] = (uint8_t)s;
      s_val[o] = a.cval[p];
      ++o;
    }
    fill += total;
  }
  __syncthreads();
  drain();
#pragma unroll
  for (int k = 0; k < kMaxSlotsPerLane; ++k) {
    const int s = lane + k * kWave;
    if (s < L) a.ring[(size_t)s * n + h] = acc[k];
  }
  if (lane == 0) a.dirty[h] = 0;
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
