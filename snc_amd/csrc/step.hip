// Fused per-step kernels of hip-atomic and hip-bucket: one launch per
// timestep does delivery-in + integrate/fire + expansion for every neuron.
//
// A block owns BLOCK consecutive neurons (128, 256 or 512: a launch-shape
// choice of the autotuner), one per thread; grid.y is the sample of a batched
// plan. Within step t a block
//   1. issues every load that depends only on the neuron (state, row
//      pointers, column bounds) before its first barrier,
//   2. stages its own injections of step t in LDS (per-(step, block) ranges),
//   3. gathers the input arriving at step t:
//        hip-atomic  the ring slot t % L (read and cleared);
//        hip-bucket  the target's column bits in the arrival bitmap of step t,
//                    walked in column order (delay descending, pre, CSR
//                    position), their values summed from 0 -- the CPU's
//                    ring-slot sum, bit for bit -- then cleared,
//   4. integrates/fires,
//   5. expands its fired rows, one wave per row, lanes striding the row:
//        hip-atomic  float atomics into ring slots (t + d) % L;
//        hip-bucket  presence bits in the arrival bitmap of step t + d.
// Arrival state of step t is only read (and cleared) at step t and only
// written at steps t - L + 1 .. t - 1 (delays are 1 .. L - 1), so within one
// launch readers and writers never touch the same bitmap or ring slot.
//
// Heavy targets (in-degree above the hub threshold) are stepped by
// k_hub_step, one block per hub, concurrently with k_step: the block loads
// the hub's arrival words in one round (up to 2048 words), orders the set
// entries by a block scan into an LDS list, gathers their values in one
// round of independent loads, and one lane sums them in column order.

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "hip_err.hpp"
#include "fold.hpp"
#include "forward_plan.hpp"
#include "wave.hpp"

namespace snc_amd {

constexpr int kStepThreads = 256;
constexpr int kLightMaxWords = 16;     // light targets: in-degree <= 512
constexpr int kHubWordsPerThread = 8;  // column words a hub-block thread loads per round
constexpr int kHubSuperWords = kStepThreads * kHubWordsPerThread;
constexpr int kHubCap = 8192;          // LDS entries of one gather pass

__device__ __forceinline__ void raster_or(uint32_t* row, int first, uint64_t m, int words) {
  const int w0 = first >> 5;
  if ((uint32_t)m) atomicOr(row + w0, (uint32_t)m);
  if ((uint32_t)(m >> 32) && w0 + 1 < words) atomicOr(row + w0 + 1, (uint32_t)(m >> 32));
}

// Expansion of CSR row [rb, re) by one wave.
template <bool BUCKET>
__device__ __forceinline__ int expand_row(const StepArgs& a, int rb, int re, int tm, int lane) {
  for (int e = rb + lane; e < re; e += kWave) {
    int s = tm + (int)a.delay[e];
    if (s >= a.L) s -= a.L;
    if (BUCKET) {
      const int p = a.cpos[e];
      atomicOr(a.bits + (size_t)s * a.bits_slot + (p >> 5), 1u << (p & 31));
    } else {
      atomicAdd(a.ring + (size_t)s * a.n + a.post[e], a.wsig[e]);
    }
  }
  return re - rb;
}

template <bool BUCKET, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_step(StepArgs a0, int t, int tm) {
  const StepArgs a = sample_args(a0, blockIdx.y);
  __shared__ float s_ext[BLOCK];
  __shared__ int2 s_rows[BLOCK];  // CSR bounds of the block's fired rows
  __shared__ int s_nf;
  __shared__ unsigned long long s_ev;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int b = blockIdx.x;
  const int base = b * BLOCK;
  const int j = base + tid;
  const int n = a.n;

  bool mine = j < n;
  float v = 0.0f, slot = 0.0f;
  int r = 0, rb = 0, re = 0, c0 = 0, c1 = 0;
  float* slotp = BUCKET ? nullptr : a.ring + (size_t)tm * n + j;
  if (mine) {
    if (BUCKET) mine = a.is_heavy[j] == 0;
    v = a.v[j];
    r = a.refr[j];
    rb = a.row_ptr[j];
    re = a.row_ptr[j + 1];
    if (BUCKET) {
      c0 = a.ccol[j];
      c1 = a.ccol[j + 1];
    } else {
      slot = *slotp;
    }
  }
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
  if (mine) {
    if (BUCKET) {  // neighbouring lanes read neighbouring columns: the slot's
                   // bitmap streams through once per step
      uint32_t* bnow = a.bits + (size_t)tm * a.bits_slot;
      const int w0 = c0 >> 5, nw = (c1 >> 5) - w0;
      uint32_t wv[kLightMaxWords];
#pragma unroll
      for (int k = 0; k < kLightMaxWords; ++k) wv[k] = k < nw ? bnow[w0 + k] : 0u;
#pragma unroll
      for (int k = 0; k < kLightMaxWords; ++k)
        if (wv[k]) bnow[w0 + k] = 0u;
      float acc = 0.0f;  // the cleared slot the CPU adds onto
#pragma unroll
      for (int k = 0; k < kLightMaxWords; ++k) {
        uint32_t bw = wv[k];
        while (bw) {
          const int p = ((w0 + k) << 5) | __builtin_ctz(bw);
          bw &= bw - 1u;
          acc = __fadd_rn(acc, a.cval[p]);
        }
      }
      slot = acc;
    }
    f = lif_step(v, r, s_ext[tid], slot, a.lif);
    a.v[j] = v;
    a.refr[j] = r;
    if (!BUCKET && __float_as_uint(slot) != 0u) *slotp = 0.0f;
    if (f) a.nspk[j] += 1;
  }
  const uint64_t m = ballot(f);
  if (m) {
    int pos = 0;
    if (lane == 0) pos = atomicAdd(&s_nf, __popcll(m));
    pos = __shfl(pos, 0, kWave);
    if (f) s_rows[pos + rank_below(m)] = make_int2(rb, re);
    if (a.raster && lane == 0) raster_or(a.raster + (size_t)t * a.words, j - lane, m, a.words);
  }
  __syncthreads();
  const int nf = s_nf;
  unsigned long long ev = 0;
  for (int k = wave; k < nf; k += BLOCK / kWave) {
    const int2 row = s_rows[k];
    ev += (unsigned long long)expand_row<BUCKET>(a, row.x, row.y, tm, lane);
  }
  if (lane == 0 && ev) atomicAdd(&s_ev, ev);
  __syncthreads();
  if (tid == 0) {
    if (nf) a.blk_spikes[b] += (unsigned long long)nf;
    if (s_ev) a.blk_events[b] += s_ev;
  }
}

// One block per heavy target (hip-bucket).
__global__ __launch_bounds__(kStepThreads) void k_hub_step(StepArgs a0, int t, int tm) {
  const StepArgs a = sample_args(a0, blockIdx.y);
  // s_val holds an entry's column position until the gather replaces it by
  // the entry's value (same thread, same index).
  __shared__ __align__(16) float s_val[kHubCap];
  __shared__ int s_wsum[kStepThreads / kWave + 1];
  __shared__ int s_fire;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int hidx = blockIdx.x;
  const int h = a.heavy_ids[hidx];
  const int slot_base = a.n_reg_blocks;  // stats slots after the regular blocks
  // neuron-only loads first (the same addresses in every lane)
  const int c0 = a.ccol[h], c1 = a.ccol[h + 1];
  float v = a.v[h];
  int r = a.refr[h];
  const float x_in = a.inj_heavy[(size_t)t * a.n_heavy + hidx];
  float in = 0.0f;  // sum of the arriving values, kept by thread 0
  {
    uint32_t* bnow = a.bits + (size_t)tm * a.bits_slot;
    const int w0 = c0 >> 5, w1 = c1 >> 5;
    for (int sb = w0; sb < w1; sb += kHubSuperWords) {  // uniform trip count
      const int nw = min(kHubSuperWords, w1 - sb);
      const int per = (nw + kStepThreads - 1) / kStepThreads;
      const int my0 = sb + tid * per;
      const int my1 = min(sb + nw, my0 + per);
      uint32_t wv[kHubWordsPerThread];
      int c = 0;
#pragma unroll
      for (int k = 0; k < kHubWordsPerThread; ++k) {
        wv[k] = my0 + k < my1 ? bnow[my0 + k] : 0u;
        c += __popc(wv[k]);
      }
#pragma unroll
      for (int k = 0; k < kHubWordsPerThread; ++k)
        if (wv[k]) bnow[my0 + k] = 0u;
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
      const int off = s_wsum[wave] + incl - c;
      const int total = s_wsum[kStepThreads / kWave];
      for (int p0 = 0; p0 < total; p0 += kHubCap) {  // uniform trip count
        const int cnt = min(kHubCap, total - p0);
        // (1) column positions of this pass's entries, in column order
        if (c && off < p0 + kHubCap && off + c > p0) {
          int o = off;
#pragma unroll
          for (int k = 0; k < kHubWordsPerThread; ++k) {
            uint32_t bw = wv[k];
            while (bw) {
              if (o >= p0 && o < p0 + kHubCap)
                s_val[o - p0] = __int_as_float(((my0 + k) << 5) | __builtin_ctz(bw));
              bw &= bw - 1u;
              ++o;
            }
          }
        }
        __syncthreads();
        // (2) one round of independent value loads
        for (int q = tid; q < cnt; q += kStepThreads) s_val[q] = a.cval[__float_as_int(s_val[q])];
        __syncthreads();
        // (3) the ordered sum, four entries per LDS read
        if (tid == 0) {
          int q = 0;
          for (; q + 4 <= cnt; q += 4) {
            const float4 x = *(const float4*)&s_val[q];
            in = __fadd_rn(__fadd_rn(__fadd_rn(__fadd_rn(in, x.x), x.y), x.z), x.w);
          }
          for (; q < cnt; ++q) in = __fadd_rn(in, s_val[q]);
        }
        __syncthreads();
      }
      __syncthreads();  // s_wsum is reused by the next super-round
    }
  }
  if (tid == 0) {
    const bool f = lif_step(v, r, x_in, in, a.lif);
    a.v[h] = v;
    a.refr[h] = r;
    s_fire = f ? 1 : 0;
    if (f) {
      a.nspk[h] += 1;
      a.blk_spikes[slot_base + hidx] += 1ull;
      if (a.raster) atomicOr(a.raster + (size_t)t * a.words + (h >> 5), 1u << (h & 31));
    }
  }
  __syncthreads();
  if (s_fire && wave == 0) {
    const int deg = expand_row<true>(a, a.row_ptr[h], a.row_ptr[h + 1], tm, lane);
    if (lane == 0) a.blk_events[slot_base + hidx] += (unsigned long long)deg;
  }
}

int fused_launch_hub(const FwdPlanCore& pc, const StepArgs& a, int t, hipStream_t st) {
  if (pc.variant == SNC_VARIANT_BUCKET && a.n_heavy > 0)
    hipLaunchKernelGGL(k_hub_step, dim3(a.n_heavy, a.B), dim3(kStepThreads), 0, st, a, t,
                       t % a.L);
  return hip_last("hub step launch failed");
}

template <int BLOCK>
static void launch_step_block(const FwdPlanCore& pc, const StepArgs& a, int t, int tm) {
  if (pc.variant == SNC_VARIANT_BUCKET)
    hipLaunchKernelGGL((k_step<true, BLOCK>), dim3(a.n_reg_blocks, a.B), dim3(BLOCK), 0,
                       pc.stream, a, t, tm);
  else
    hipLaunchKernelGGL((k_step<false, BLOCK>), dim3(a.n_reg_blocks, a.B), dim3(BLOCK), 0,
                       pc.stream, a, t, tm);
}

int fused_launch_regular(const FwdPlanCore& pc, const StepArgs& a, int t) {
  const int tm = t % a.L;
  switch (a.blk) {
    case 128: launch_step_block<128>(pc, a, t, tm); break;
    case 256: launch_step_block<256>(pc, a, t, tm); break;
    case 512: launch_step_block<512>(pc, a, t, tm); break;
    default: return fail(SNC_E_INVALID, "step block must be 128, 256 or 512 neurons");
  }
  return hip_last("step launch failed");
}

int fused_launch_step(const FwdPlanCore& pc, const StepArgs& a, int t, hipStream_t hub_stream) {
  int rc = fused_launch_hub(pc, a, t, hub_stream ? hub_stream : pc.stream);
  if (!rc) rc = fused_launch_regular(pc, a, t);
  return rc;
}

bool fused_block_ok(int block) { return block == 128 || block == 256 || block == 512; }

int fused_light_max_in_degree() { return kLightMaxWords * 32; }

}  // namespace snc_amd
