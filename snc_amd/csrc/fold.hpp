// Device building blocks shared by the per-step and persistent kernels:
// the LIF update and the ordered folds of the hip-bucket strategy.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "forward_plan.hpp"
#include "wave.hpp"

namespace snc_amd {

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

// Block reduction of per-wave counts (lane 0's value of each wave) into the
// block's own stats slot. Every thread of the block must call it.
__device__ __forceinline__ void block_accumulate(int wave_count, unsigned long long* slot) {
  __shared__ int s_sum;
  if (threadIdx.x == 0) s_sum = 0;
  __syncthreads();
  if (lane_id() == 0 && wave_count) atomicAdd(&s_sum, wave_count);
  __syncthreads();
  if (threadIdx.x == 0 && s_sum) *slot += (unsigned long long)s_sum;
}

// One LIF step (cpu_backend.cpp:129-143 with the reset-mode extension).
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

// Wave-level ordered fold of a heavy target's pending deposits. The wave
// gathers the set positions of the target's column in ascending order into
// LDS (CHUNK entries at a time), then lane s folds, in that order, every value
// bound for ring slot s (lanes cover s, s+64, ... for rings above 64 slots).
// Every (target, slot) sum is the sequential sum in column order.
constexpr int kMaxSlotsPerLane = 4;  // rings up to 256 slots
constexpr int kHeavyChunk = 2048;    // LDS entries of the one-wave heavy-fold kernel

template <int CHUNK>
__device__ void fold_heavy_wave(int h, int tprev, int L, int n, const int32_t* __restrict__ ccol,
                                const float* __restrict__ cval, const uint8_t* __restrict__ cdel,
                                uint32_t* __restrict__ bits, uint8_t* __restrict__ dirty,
                                float* __restrict__ ring, float* s_val, uint8_t* s_slot) {
  if (!dirty[h]) return;
  const int lane = lane_id();
  float acc[kMaxSlotsPerLane];
#pragma unroll
  for (int k = 0; k < kMaxSlotsPerLane; ++k) {
    const int s = lane + k * kWave;
    acc[k] = s < L ? ring[(size_t)s * n + h] : 0.0f;
  }
  int fill = 0;
  auto drain = [&]() {
    wave_lds_sync();
    for (int q = 0; q < fill; ++q) {
      const int s = s_slot[q];
      const float x = s_val[q];
#pragma unroll
      for (int k = 0; k < kMaxSlotsPerLane; ++k)
        if (s == lane + k * kWave) acc[k] = __fadd_rn(acc[k], x);
    }
    wave_lds_sync();
    fill = 0;
  };
  const int w0 = ccol[h] >> 5, w1 = ccol[h + 1] >> 5;
  for (int wb = w0; wb < w1; wb += kWave) {
    const int w = wb + lane;
    uint32_t b = 0;
    if (w < w1) {
      b = bits[w];
      if (b) bits[w] = 0u;
    }
    const int c = __popc(b);
    const int incl = wave_inclusive_scan(c);
    int total = __shfl(incl, kWave - 1, kWave);
    if (total == 0) continue;
    if (fill + total > CHUNK) drain();
    // A single 64-word slab can exceed a small CHUNK: spill it in pieces.
    int o = fill + incl - c;
    while (total > 0) {
      int wrote = 0;
      while (b && o < CHUNK) {
        const int p = (w << 5) | __builtin_ctz(b);
        b &= b - 1u;
        int s = tprev + (int)cdel[p];
        if (s >= L) s -= L;
        s_slot[o] = (uint8_t)s;
        s_val[o] = cval[p];
        ++o;
        ++wrote;
      }
      const int wrote_all = wave_sum(wrote);
      total -= wrote_all;
      fill += wrote_all;
      if (total > 0) {
        // Lanes whose bits did not fit resume after a drain, keeping order:
        // the lowest lane with bits left continues first.
        drain();
        const int left = __popc(b);
        const int li = wave_inclusive_scan(left);
        o = li - left;
      }
    }
  }
  drain();
#pragma unroll
  for (int k = 0; k < kMaxSlotsPerLane; ++k) {
    const int s = lane + k * kWave;
    if (s < L) ring[(size_t)s * n + h] = acc[k];
  }
  if (lane == 0) dirty[h] = 0;
}

}  // namespace snc_amd
