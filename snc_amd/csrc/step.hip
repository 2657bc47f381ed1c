// Fused per-step kernels of hip-atomic and hip-bucket: one launch per
// timestep does fold + integrate/fire + expansion for every neuron.
//
// A block owns BLOCK consecutive neurons (128, 256 or 512: a launch-shape
// choice of the autotuner), one per thread. Within step t it
//   1. stages its own injections of step t in LDS (per-(step, block) ranges),
//   2. [hip-bucket] folds the deposits made at step t-1 into its targets:
//      the target's column words are loaded together, set bits are walked in
//      ascending column position; values bound for slot t % L accumulate in a
//      register on top of the slot value, others go to their ring slot,
//   3. integrates/fires (ext, slot t % L read and cleared),
//   4. expands its fired rows, one wave per row: float atomics into slots
//      (t + d) % L (hip-atomic) or presence bits of step parity t & 1
//      (hip-bucket).
// No two blocks of one launch touch the same ring slot, bitmap half or dirty
// half: the ring has max_delay + 1 slots, so slot t % L never receives a
// deposit of step t, and the bucket state is double-buffered by step parity.
//
// Heavy targets (in-degree above the light limit) are stepped by
// k_hub_step, one block per hub, which may run concurrently with k_step: the
// block folds the hub's column in ordered rounds (block scan, LDS list, one
// lane per ring slot), integrates the hub and expands its row.

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "fold.hpp"
#include "forward_plan.hpp"
#include "wave.hpp"

namespace snc_amd {

constexpr int kStepThreads = 256;
constexpr int kLightMaxWords = 16;  // light targets: in-degree <= 512
constexpr int kHubRoundWords = 256; // column words per hub-fold round (one per thread)
constexpr int kHubCap = kHubRoundWords * 32;

__device__ __forceinline__ void raster_or(uint32_t* row, int first, uint64_t m, int words) {
  const int w0 = first >> 5;
  if ((uint32_t)m) atomicOr(row + w0, (uint32_t)m);
  if ((uint32_t)(m >> 32) && w0 + 1 < words) atomicOr(row + w0 + 1, (uint32_t)(m >> 32));
}

// Expansion of one fired row by one wave.
template <bool BUCKET>
__device__ __forceinline__ int expand_row(const StepArgs& a, int i, int tm, uint32_t* bits_now,
                                          uint8_t* dirty_now, int lane) {
  const int rb = a.row_ptr[i], re = a.row_ptr[i + 1];
  for (int e = rb + lane; e < re; e += kWave) {
    if (BUCKET) {
      const int p = a.cpos[e];
      atomicOr(bits_now + (p >> 5), 1u << (p & 31));
      dirty_now[a.post[e]] = 1;
    } else {
      int s = tm + (int)a.delay[e];
      if (s >= a.L) s -= a.L;
      atomicAdd(a.ring + (size_t)s * a.n + a.post[e], a.wsig[e]);
    }
  }
  return re - rb;
}

template <bool BUCKET, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_step(StepArgs a, int t, int tm) {
  __shared__ float s_ext[BLOCK];
  __shared__ int32_t s_fired[BLOCK];
  __shared__ int s_nf;
  __shared__ unsigned long long s_ev;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int b = blockIdx.x;
  const int base = b * BLOCK;
  const int j = base + tid;
  const int n = a.n, L = a.L;
  s_ext[tid] = 0.0f;
  if (tid == 0) {
    s_nf = 0;
    s_ev = 0;
  }
  __syncthreads();
  const int32_t* bp = a.inj_blk_ptr + (size_t)t * (a.n_reg_blocks + 1) + b;
  for (int k = bp[0] + tid; k < bp[1]; k += BLOCK) s_ext[a.blk_nid[k] - base] = a.blk_amp[k];
  __syncthreads();

  bool f = false;
  if (j < n && !(BUCKET && a.is_heavy[j])) {
    float v = a.v[j];
    int r = a.refr[j];
    float* slotp = a.ring + (size_t)tm * n + j;
    float slot = *slotp;
    if (BUCKET && t > 0) {
      uint8_t* dprev = a.dirty + (size_t)((t - 1) & 1) * a.dirty_half;
      if (dprev[j]) {
        dprev[j] = 0;
        uint32_t* bprev = a.bits + (size_t)((t - 1) & 1) * a.bits_half;
        const int tprev = tm == 0 ? L - 1 : tm - 1;
        const int w0 = a.ccol[j] >> 5, nw = (a.ccol[j + 1] >> 5) - w0;
        uint32_t wv[kLightMaxWords];
#pragma unroll
        for (int k = 0; k < kLightMaxWords; ++k) wv[k] = k < nw ? bprev[w0 + k] : 0u;
#pragma unroll
        for (int k = 0; k < kLightMaxWords; ++k)
          if (wv[k]) bprev[w0 + k] = 0u;
        int cur = -1;  // run of a slot other than tm, kept in a register
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < kLightMaxWords; ++k) {
          uint32_t bw = wv[k];
          while (bw) {
            const int p = ((w0 + k) << 5) | __builtin_ctz(bw);
            bw &= bw - 1u;
            int s = tprev + (int)a.cdel[p];
            if (s >= L) s -= L;
            const float x = a.cval[p];
            if (s == tm) {
              slot = __fadd_rn(slot, x);
            } else {
              if (s != cur) {
                if (cur >= 0) a.ring[(size_t)cur * n + j] = acc;
                cur = s;
                acc = a.ring[(size_t)s * n + j];
              }
              acc = __fadd_rn(acc, x);
            }
          }
        }
        if (cur >= 0) a.ring[(size_t)cur * n + j] = acc;
      }
    }
    f = lif_step(v, r, s_ext[tid], slot, a.lif);
    a.v[j] = v;
    a.refr[j] = r;
    if (__float_as_uint(slot) != 0u) *slotp = 0.0f;
    if (f) a.nspk[j] += 1;
  }
  const uint64_t m = ballot(f);
  if (m) {
    int pos = 0;
    if (lane == 0) pos = atomicAdd(&s_nf, __popcll(m));
    pos = __shfl(pos, 0, kWave);
    if (f) s_fired[pos + rank_below(m)] = j;
    if (a.raster && lane == 0) raster_or(a.raster + (size_t)t * a.words, j - lane, m, a.words);
  }
  __syncthreads();
  const int nf = s_nf;
  uint32_t* bits_now = BUCKET ? a.bits + (size_t)(t & 1) * a.bits_half : nullptr;
  uint8_t* dirty_now = BUCKET ? a.dirty + (size_t)(t & 1) * a.dirty_half : nullptr;
  unsigned long long ev = 0;
  for (int k = wave; k < nf; k += BLOCK / kWave)
    ev += (unsigned long long)expand_row<BUCKET>(a, s_fired[k], tm, bits_now, dirty_now, lane);
  if (lane == 0 && ev) atomicAdd(&s_ev, ev);
  __syncthreads();
  if (tid == 0) {
    if (nf) a.blk_spikes[b] += (unsigned long long)nf;
    if (s_ev) a.blk_events[b] += s_ev;
  }
}

// One block per heavy target (hip-bucket).
__global__ __launch_bounds__(kStepThreads) void k_hub_step(StepArgs a, int t, int tm) {
  __shared__ float s_val[kHubCap];
  __shared__ uint8_t s_slot[kHubCap];
  __shared__ int s_wsum[kStepThreads / kWave + 1];
  __shared__ int s_fire;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int hidx = blockIdx.x;
  const int h = a.heavy_ids[hidx];
  const int n = a.n, L = a.L;
  const int slot_base = a.n_reg_blocks;  // stats slots after the regular blocks
  if (t > 0) {
    uint8_t* dprev = a.dirty + (size_t)((t - 1) & 1) * a.dirty_half;
    if (dprev[h]) {  // same value for every thread of the block
      uint32_t* bprev = a.bits + (size_t)((t - 1) & 1) * a.bits_half;
      const int tprev = tm == 0 ? L - 1 : tm - 1;
      float acc[kMaxSlotsPerLane];
      if (wave == 0)
#pragma unroll
        for (int k = 0; k < kMaxSlotsPerLane; ++k) {
          const int s = lane + k * kWave;
          acc[k] = s < L ? a.ring[(size_t)s * n + h] : 0.0f;
        }
      const int w0 = a.ccol[h] >> 5, w1 = a.ccol[h + 1] >> 5;
      for (int wb = w0; wb < w1; wb += kHubRoundWords) {  // uniform trip count
        const int w = wb + tid;
        uint32_t bw = 0;
        if (w < w1) {
          bw = bprev[w];
          if (bw) bprev[w] = 0u;
        }
        // block-wide exclusive scan of the set-bit counts (word order = column order)
        const int c = __popc(bw);
        const int incl = wave_inclusive_scan(c);
        if (lane == kWave - 1) s_wsum[wave] = incl;
        __syncthreads();
        if (tid == 0) {
          int run = 0;
          for (int q = 0; q < kStepThreads / kWave; ++q) {
            const int x = s_wsum[q];
            s_wsum[q] = run;
            run += x;
          }
          s_wsum[kStepThreads / kWave] = run;
        }
        __syncthreads();
        int o = s_wsum[wave] + incl - c;
        const int total = s_wsum[kStepThreads / kWave];
        while (bw) {
          const int p = (w << 5) | __builtin_ctz(bw);
          bw &= bw - 1u;
          int s = tprev + (int)a.cdel[p];
          if (s >= L) s -= L;
          s_slot[o] = (uint8_t)s;
          s_val[o] = a.cval[p];
          ++o;
        }
        __syncthreads();
        if (wave == 0)
          for (int q = 0; q < total; ++q) {
            const int s = s_slot[q];
            const float x = s_val[q];
#pragma unroll
            for (int k = 0; k < kMaxSlotsPerLane; ++k)
              if (s == lane + k * kWave) acc[k] = __fadd_rn(acc[k], x);
          }
        __syncthreads();
      }
      if (wave == 0)
#pragma unroll
        for (int k = 0; k < kMaxSlotsPerLane; ++k) {
          const int s = lane + k * kWave;
          if (s < L) a.ring[(size_t)s * n + h] = acc[k];
        }
      if (tid == 0) dprev[h] = 0;
      __syncthreads();
    }
  }
  if (tid == 0) {
    float v = a.v[h];
    int r = a.refr[h];
    float* slotp = a.ring + (size_t)tm * n + h;
    const float slot = *slotp;
    const bool f = lif_step(v, r, a.inj_heavy[(size_t)t * a.n_heavy + hidx], slot, a.lif);
    a.v[h] = v;
    a.refr[h] = r;
    if (__float_as_uint(slot) != 0u) *slotp = 0.0f;
    s_fire = f ? 1 : 0;
    if (f) {
      a.nspk[h] += 1;
      a.blk_spikes[slot_base + hidx] += 1ull;
      if (a.raster) atomicOr(a.raster + (size_t)t * a.words + (h >> 5), 1u << (h & 31));
    }
  }
  __syncthreads();
  if (s_fire && wave == 0) {
    const int deg = expand_row<true>(a, h, tm, a.bits + (size_t)(t & 1) * a.bits_half,
                                     a.dirty + (size_t)(t & 1) * a.dirty_half, lane);
    if (lane == 0) a.blk_events[slot_base + hidx] += (unsigned long long)deg;
  }
}

int fused_launch_hub(const FwdPlanCore& pc, const StepArgs& a, int t, hipStream_t st) {
  if (pc.variant == SNC_VARIANT_BUCKET && a.n_heavy > 0)
    hipLaunchKernelGGL(k_hub_step, dim3(a.n_heavy), dim3(kStepThreads), 0, st, a, t, t % a.L);
  return hipGetLastError() == hipSuccess ? SNC_OK : fail(SNC_E_HIP, "hub step launch failed");
}

template <int BLOCK>
static void launch_step_block(const FwdPlanCore& pc, const StepArgs& a, int t, int tm) {
  if (pc.variant == SNC_VARIANT_BUCKET)
    hipLaunchKernelGGL((k_step<true, BLOCK>), dim3(a.n_reg_blocks), dim3(BLOCK), 0, pc.stream, a,
                       t, tm);
  else
    hipLaunchKernelGGL((k_step<false, BLOCK>), dim3(a.n_reg_blocks), dim3(BLOCK), 0, pc.stream,
                       a, t, tm);
}

int fused_launch_regular(const FwdPlanCore& pc, const StepArgs& a, int t) {
  const int tm = t % a.L;
  switch (a.blk) {
    case 128: launch_step_block<128>(pc, a, t, tm); break;
    case 256: launch_step_block<256>(pc, a, t, tm); break;
    case 512: launch_step_block<512>(pc, a, t, tm); break;
    default: return fail(SNC_E_INVALID, "step block must be 128, 256 or 512 neurons");
  }
  return hipGetLastError() == hipSuccess ? SNC_OK : fail(SNC_E_HIP, "step launch failed");
}

int fused_launch_step(const FwdPlanCore& pc, const StepArgs& a, int t, hipStream_t hub_stream) {
  int rc = fused_launch_hub(pc, a, t, hub_stream ? hub_stream : pc.stream);
  if (!rc) rc = fused_launch_regular(pc, a, t);
  return rc;
}

bool fused_block_ok(int block) { return block == 128 || block == 256 || block == 512; }

int fused_light_max_in_degree() { return kLightMaxWords * 32; }

}  // namespace snc_amd
