// Persistent forward kernel: the whole step loop of one sample in a single
// cooperative launch, one grid barrier per timestep.
//
// Each block owns a contiguous, 64-aligned range of neurons for the whole
// run; each lane keeps the membrane and refractory counter of its neurons in
// registers. Per step the block
//   (a) [hip-bucket] folds the deposits made at step t-1 into its targets'
//       ring slots, heavy targets by one wave each, light ones by their lane,
//   (b) integrates/fires its neurons (ext and ring slot t % L read and
//       cleared), compacting the fired ids into LDS,
//   (c) expands its own fired rows, one wave per row,
//   (d) stages step t+1's injections,
// then all blocks meet at the grid barrier. With ring_len = max_delay + 1 no
// deposit of step t lands in slot t % L, so (b) of one block and (c) of
// another never touch the same slot within a step. ext, the presence bitmap
// and the dirty flags are double-buffered by step parity: step t writes
// parity (t+1) & 1 (ext) or t & 1 (bitmap, dirty) while readers of the same
// step use the other half.

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "common.hpp"
#include "hip_err.hpp"
#include "fold.hpp"
#include "persist.hpp"

namespace cg = cooperative_groups;

namespace snc_amd {

constexpr int kPersistThreads = 1024;
constexpr int kPersistWaves = kPersistThreads / kWave;
constexpr int kMaxPerLane = 4;        // neurons per lane kept in registers
constexpr int kPersistHeavyChunk = 256;

template <bool BUCKET>
__global__ __launch_bounds__(kPersistThreads) void k_persist(StepArgs a, PersistArgs q) {
  __shared__ int32_t s_fired[kMaxPerLane * kPersistThreads];
  __shared__ int s_nf;
  __shared__ float s_hval[kPersistWaves][kPersistHeavyChunk];
  __shared__ uint8_t s_hslot[kPersistWaves][kPersistHeavyChunk];
  __shared__ unsigned long long s_ev;
  cg::grid_group grid = cg::this_grid();

  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int base = blockIdx.x * q.npb;
  const int n = a.n, L = a.L;
  float v[kMaxPerLane];
  int r[kMaxPerLane];
#pragma unroll
  for (int k = 0; k < kMaxPerLane; ++k) {
    v[k] = 0.0f;
    r[k] = 0;
  }
  if (tid == 0) {
    s_nf = 0;
    s_ev = 0;
  }
  unsigned long long my_spikes = 0;
  const int h0 = q.heavy_blk_ptr[blockIdx.x], h1 = q.heavy_blk_ptr[blockIdx.x + 1];
  __syncthreads();

  int tm = 0;
  for (int t = 0; t < a.T; ++t) {
    const int par = t & 1;
    const int tprev = tm == 0 ? L - 1 : tm - 1;
    float* ext_now = a.ext + (size_t)par * n;
    if (BUCKET && t > 0) {
      uint32_t* bits_prev = a.bits + (size_t)(par ^ 1) * q.bits_words;
      uint8_t* dirty_prev = a.dirty + (size_t)(par ^ 1) * q.dirty_stride;
      for (int hi = h0 + wave; hi < h1; hi += kPersistWaves)
        fold_heavy_wave<kPersistHeavyChunk>(a.heavy_ids[hi], tprev, L, n, a.ccol, a.cval, a.cdel,
                                            bits_prev, dirty_prev, a.ring, s_hval[wave],
                                            s_hslot[wave]);
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kMaxPerLane; ++k) {
        const int j = base + k * kPersistThreads + tid;
        if (k < q.per_lane && j < n && j < base + q.npb && dirty_prev[j]) {
          fold_light(j, n, L, tprev, a.ccol, a.cval, a.cdel, bits_prev, a.ring);
          dirty_prev[j] = 0;
        }
      }
    }

    // (b) integrate / fire, compaction into LDS, raster words.
#pragma unroll
    for (int k = 0; k < kMaxPerLane; ++k) {
      if (k >= q.per_lane) break;  // uniform across the block
      const int j = base + k * kPersistThreads + tid;
      const bool mine = j < n && j < base + q.npb;
      bool f = false;
      if (mine) {
        const float x = ext_now[j];
        if (__float_as_uint(x) != 0u) ext_now[j] = 0.0f;
        float* sp = a.ring + (size_t)tm * n + j;
        const float s = *sp;
        if (__float_as_uint(s) != 0u) *sp = 0.0f;
        f = lif_step(v[k], r[k], x, s, a.lif);
        if (f) a.nspk[j] += 1;
      }
      const uint64_t m = ballot(f);
      if (m) {
        const int cnt = __popcll(m);
        int pos = 0;
        if (lane == 0) pos = atomicAdd(&s_nf, cnt);
        pos = __shfl(pos, 0, kWave);
        if (f) s_fired[pos + rank_below(m)] = j;
        if (lane == 0) {
          my_spikes += (unsigned long long)cnt;
          if (a.raster) {
            const int w0 = (j - lane) >> 5;
            uint32_t* row = a.raster + (size_t)t * a.words;
            if ((uint32_t)m) row[w0] = (uint32_t)m;
            if ((uint32_t)(m >> 32) && w0 + 1 < a.words) row[w0 + 1] = (uint32_t)(m >> 32);
          }
        }
      }
    }
    __syncthreads();

    // (c) expansion of this block's fired rows.
    const int nf = s_nf;
    unsigned long long ev = 0;
    uint32_t* bits_now = BUCKET ? a.bits + (size_t)par * q.bits_words : nullptr;
    uint8_t* dirty_now = BUCKET ? a.dirty + (size_t)par * q.dirty_stride : nullptr;
    for (int k = wave; k < nf; k += kPersistWaves) {
      const int i = s_fired[k];
      const int rb = a.row_ptr[i], re = a.row_ptr[i + 1];
      for (int e = rb + lane; e < re; e += kWave) {
        if (BUCKET) {
          const int p = a.cpos[e];
          atomicOr(bits_now + (p >> 5), 1u << (p & 31));
          dirty_now[a.post[e]] = 1;
        } else {
          int s = tm + (int)a.delay[e];
          if (s >= L) s -= L;
          atomicAdd(a.ring + (size_t)s * n + a.post[e], a.wsig[e]);
        }
      }
      ev += (unsigned long long)(re - rb);
    }
    if (lane == 0 && ev) atomicAdd(&s_ev, ev);

    // (d) injections of step t+1 into the other ext half.
    if (t + 1 < a.T) {
      float* ext_next = a.ext + (size_t)(par ^ 1) * n;
      const int b0 = a.inj_ptr[t + 1], b1 = a.inj_ptr[t + 2];
      const int g = (int)(blockIdx.x * kPersistThreads + tid), gs = (int)(gridDim.x * kPersistThreads);
      for (int k = b0 + g; k < b1; k += gs) ext_next[a.inj_nid[k]] = a.inj_amp[k];
    }
    __syncthreads();
    if (tid == 0) s_nf = 0;
    tm = tm + 1 == L ? 0 : tm + 1;
    grid.sync();
  }

  // Final state for capture.
#pragma unroll
  for (int k = 0; k < kMaxPerLane; ++k) {
    const int j = base + k * kPersistThreads + tid;
    if (k < q.per_lane && j < n && j < base + q.npb) {
      a.v[j] = v[k];
      a.refr[j] = r[k];
    }
  }
  if (lane == 0 && my_spikes) atomicAdd(a.blk_spikes + blockIdx.x, my_spikes);
  __syncthreads();
  if (tid == 0) a.blk_events[blockIdx.x] += s_ev;
}

bool persist_supported() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return false;
  return prop.cooperativeLaunch != 0;
}

bool persist_choose(const DevGraph& g, int variant, bool want, PersistConfig& cfg) {
  if (variant != SNC_VARIANT_ATOMIC && variant != SNC_VARIANT_BUCKET) return false;
  // Opt-in (the autotuner's candidate, or SNC_HIP_PERSIST=1): on MI355X the
  // graph-launched fused step kernels measured faster than one cooperative
  // launch with a grid barrier per step.
  const char* e = std::getenv("SNC_HIP_PERSIST");
  if (!want && !(e && e[0] == '1')) return false;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return false;
  if (!prop.cooperativeLaunch) return false;
  int per_cu = 0;
  const void* fn = variant == SNC_VARIANT_BUCKET ? (const void*)k_persist<true>
                                                 : (const void*)k_persist<false>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kPersistThreads, 0) != hipSuccess ||
      per_cu < 1)
    return false;
  const int max_blocks = per_cu * prop.multiProcessorCount;
  // Enough blocks that each lane keeps at most kMaxPerLane neurons; no more
  // blocks than needed (fewer blocks make the barrier cheaper).
  int blocks = std::min(max_blocks, std::max(1, (g.n + kPersistThreads - 1) / kPersistThreads));
  blocks = std::max(blocks, std::min(max_blocks, prop.multiProcessorCount));
  int64_t npb = ((int64_t)g.n + blocks - 1) / blocks;
  npb = (npb + 63) / 64 * 64;
  const int per_lane = (int)((npb + kPersistThreads - 1) / kPersistThreads);
  if (per_lane > kMaxPerLane || per_lane < 1) return false;
  blocks = (int)((g.n + npb - 1) / npb);
  cfg.blocks = std::max(1, blocks);
  cfg.threads = kPersistThreads;
  cfg.npb = (int)npb;
  cfg.per_lane = per_lane;
  return true;
}

int persist_launch(const DevGraph& g, const FwdPlanCore& pc, const StepArgs& a,
                   const PersistConfig& cfg) {
  (void)g;
  PersistArgs q{};
  q.npb = cfg.npb;
  q.per_lane = cfg.per_lane;
  q.heavy_blk_ptr = cfg.heavy_blk_ptr;
  q.bits_words = cfg.bits_words;
  q.dirty_stride = cfg.dirty_stride;
  StepArgs aa = a;
  void* args[] = {(void*)&aa, (void*)&q};
  const void* fn = pc.variant == SNC_VARIANT_BUCKET ? (const void*)k_persist<true>
                                                    : (const void*)k_persist<false>;
  const hipError_t e =
      hipLaunchCooperativeKernel(fn, dim3(cfg.blocks), dim3(cfg.threads), args, 0, pc.stream);
  if (e != hipSuccess) return hip_fail(e, "cooperative launch of the persistent kernel");
  return SNC_OK;
}

// Heavy targets grouped by owning block (heavy_ids is ascending).
std::vector<int32_t> persist_heavy_ptr(const DevGraph& g, const std::vector<int32_t>& heavy_sorted,
                                       const PersistConfig& cfg) {
  (void)g;
  std::vector<int32_t> ptr((size_t)cfg.blocks + 1, 0);
  size_t k = 0;
  for (int b = 0; b < cfg.blocks; ++b) {
    ptr[b] = (int32_t)k;
    const int64_t hi = (int64_t)(b + 1) * cfg.npb;
    while (k < heavy_sorted.size() && heavy_sorted[k] < hi) ++k;
  }
  ptr[cfg.blocks] = (int32_t)k;
  return ptr;
}

}  // namespace snc_amd
