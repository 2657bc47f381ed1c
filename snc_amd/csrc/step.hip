// Fused per-step kernels of hip-atomic and hip-bucket: one launch per
// timestep does delivery-in + integrate/fire + expansion for every neuron.
//
// A block owns BLOCK consecutive neurons (128, 256 or 512: a launch-shape
// choice of the autotuner), one per thread; grid.y is the sample of a batched
// plan. Within step t a block
//   1. issues every load that depends only on the neuron, the step and the
//      block (state, row pointers, column bounds, the inline injection
//      entries of (step, block)) in one round,
//   2. gathers the input arriving at step t, before any barrier:
//        hip-atomic  the ring slot t % L (read and cleared);
//        hip-bucket  the target's column bits in the arrival bitmap of step t,
//                    walked in column order (delay descending, pre, CSR
//                    position), their values summed from 0 -- the CPU's
//                    ring-slot sum, bit for bit -- then cleared,
//   3. stages its injections of step t in LDS,
//   4. integrates/fires; spike counts and block statistics are no-return
//      atomics, so no load-modify-store round trip sits on the step's path,
//   5. expands its fired rows, one wave per row, lanes striding the row:
//        hip-atomic  float atomics into ring slots (t + d) % L;
//        hip-bucket  presence bits in the arrival bitmap of step t + d.
// Arrival state of step t is only read (and cleared) at step t and only
// written at steps t - L + 1 .. t - 1 (delays are 1 .. L - 1), so within one
// launch readers and writers never touch the same bitmap or ring slot.
//
// Heavy targets (in-degree above the hub threshold) of hip-bucket are the
// launch's leading blocks, one per hub, so one launch per step covers every
// neuron: a hub block loads the hub's arrival words in rounds of 8 words per
// thread, orders the set entries by a block scan into an LDS list, gathers
// their values in one round of independent loads, and one lane sums them in
// column order.

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "hip_err.hpp"
#include "fold.hpp"
#include "forward_plan.hpp"
#include "wave.hpp"

namespace snc_amd {

constexpr int kLightMaxWords = 16;     // light targets: in-degree <= 512
constexpr int kHubWordsPerThread = 8;  // column words a hub-block thread loads per round
constexpr int kHubCap = 2048;          // LDS entries of one gather pass
constexpr int kGather = 8;             // value loads a light target keeps in flight

// LDS of one block: the regular step's arrays, or (hip-bucket) a hub block's
// gather list -- a block is one or the other, so they share the bytes.
constexpr size_t step_lds_bytes(int block) { return (size_t)block * 12 + 16; }
constexpr size_t hub_lds_bytes(int block) { return (size_t)kHubCap * 4 + ((size_t)block / 64 + 2) * 4; }
constexpr size_t lds_bytes(bool bucket, int block) {
  return bucket && hub_lds_bytes(block) > step_lds_bytes(block) ? hub_lds_bytes(block)
                                                                 : step_lds_bytes(block);
}

__device__ __forceinline__ void raster_or(uint32_t* row, int first, uint64_t m, int words) {
  const int w0 = first >> 5;
  if ((uint32_t)m) atomicOr(row + w0, (uint32_t)m);
  if ((uint32_t)(m >> 32) && w0 + 1 < words) atomicOr(row + w0 + 1, (uint32_t)(m >> 32));
}

// Expansion of CSR row [rb, re) by one wave. Each lane loads U edges before
// its first deposit, so a row of up to 64 * U edges costs one round of
// loads; indexes past the row's end are clamped to its last edge (in bounds,
// the deposit is skipped). Measured on the delivery-stress graph (100-edge
// rows): hip-atomic is 1.4 % faster per step with two edges per lane than
// with four (eight is slower still); hip-bucket is 1.1 % faster with four.
template <bool BUCKET>
constexpr int expand_unroll() { return BUCKET ? 4 : 2; }

template <bool BUCKET>
__device__ __forceinline__ void expand_row(const StepArgs& a, int rb, int re, int tm, int lane) {
  constexpr int kExpandUnroll = expand_unroll<BUCKET>();
  for (int g = rb; g < re; g += kExpandUnroll * kWave) {  // wave-uniform trip count
    const int e0 = g + lane;
    int d[kExpandUnroll], tgt[kExpandUnroll];
    float w[kExpandUnroll];
#pragma unroll
    for (int u = 0; u < kExpandUnroll; ++u) {
      const int e = min(e0 + u * kWave, re - 1);
      d[u] = a.delay[e];
      tgt[u] = BUCKET ? a.cpos[e] : a.post[e];
      w[u] = BUCKET ? 0.0f : a.wsig[e];
    }
#pragma unroll
    for (int u = 0; u < kExpandUnroll; ++u) {
      if (e0 + u * kWave >= re) continue;
      int s = tm + d[u];
      if (s >= a.L) s -= a.L;
      if (BUCKET)
        atomicOr(a.bits + (size_t)s * a.bits_slot + (tgt[u] >> 5), 1u << (tgt[u] & 31));
      else
        atomicAdd(a.ring + (size_t)s * a.n + tgt[u], w[u]);
    }
  }
}

template <int NT>
__device__ void hub_block(const StepArgs& a, int t, int tm, int hidx, unsigned char* lds);

template <bool BUCKET, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_step(StepArgs a0, int t, int tm) {
  const StepArgs a = sample_args(a0, blockIdx.y);
  __shared__ __align__(16) unsigned char s_lds[lds_bytes(BUCKET, BLOCK)];
  if (BUCKET && (int)blockIdx.x < a.n_heavy) {  // hub blocks first: they run longest
    hub_block<BLOCK>(a, t, tm, blockIdx.x, s_lds);
    return;
  }
  int2* s_rows = reinterpret_cast<int2*>(s_lds);  // CSR bounds of the block's fired rows
  unsigned long long& s_ev = *reinterpret_cast<unsigned long long*>(s_lds + BLOCK * 8);
  float* s_ext = reinterpret_cast<float*>(s_lds + BLOCK * 8 + 8);
  int& s_nf = *reinterpret_cast<int*>(s_lds + BLOCK * 12 + 8);
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int b = (int)blockIdx.x - (BUCKET ? a.n_heavy : 0);
  const int base = b * BLOCK;
  const int j = base + tid;
  const int n = a.n;

  // First batch of loads: everything that depends only on the neuron, the step and the
  // block -- state, row and column bounds, the inline injection entries.
  const int2* tab = a.inj_tab + ((size_t)t * a.n_reg_blocks + b) * kInjInline;
  const bool spill = tab[kInjInline - 1].x == kInjSpill;
  int2 ent = make_int2(-1, 0);
  if (tid < kInjInline) ent = tab[tid];
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
  if (BUCKET && mine) {
    // Second and third batches of loads (registers only, before any
    // barrier): the target's column words of this step's arrival bitmap,
    // then the values of the set bits.
    // Neighbouring lanes read neighbouring columns, so the bitmap streams
    // through once per step.
    uint32_t* bnow = a.bits + (size_t)tm * a.bits_slot;
    const int w0 = c0 >> 5, nw = (c1 >> 5) - w0;
    uint32_t wv[kLightMaxWords];
#pragma unroll
    for (int k = 0; k < kLightMaxWords; ++k) wv[k] = k < nw ? bnow[w0 + k] : 0u;
#pragma unroll
    for (int k = 0; k < kLightMaxWords; ++k)
      if (wv[k]) bnow[w0 + k] = 0u;
    // The set bits' values, kGather loads in flight per round (a load per
    // arrival in turn would cost the block its busiest target's arrival
    // count in round trips), added in column order onto the cleared slot.
    float acc = 0.0f;
    int pb[kGather];
    int nb = 0;
    auto flush = [&]() {
      float x[kGather];
#pragma unroll
      for (int u = 0; u < kGather; ++u) x[u] = u < nb ? a.cval[pb[u]] : 0.0f;
#pragma unroll
      for (int u = 0; u < kGather; ++u)
        if (u < nb) acc = __fadd_rn(acc, x[u]);
      nb = 0;
    };
#pragma unroll
    for (int k = 0; k < kLightMaxWords; ++k) {
      uint32_t bw = wv[k];
      while (bw) {
        const int p = ((w0 + k) << 5) | __builtin_ctz(bw);
        bw &= bw - 1u;
#pragma unroll
        for (int u = 0; u < kGather; ++u)
          if (nb == u) pb[u] = p;
        if (++nb == kGather) flush();
      }
    }
    flush();
    slot = acc;
  }
  __syncthreads();
  if (spill) {  // more than kInjInline injections: the (step, block) list
    const int32_t* bp = a.inj_blk_ptr + (size_t)t * (a.n_reg_blocks + 1) + b;
    for (int k = bp[0] + tid; k < bp[1]; k += BLOCK) s_ext[a.blk_nid[k] - base] = a.blk_amp[k];
  } else if (ent.x >= 0) {
    s_ext[ent.x] = __int_as_float(ent.y);
  }
  __syncthreads();

  bool f = false;
  if (mine) {
    f = lif_step(v, r, s_ext[tid], slot, a.lif);
    a.v[j] = v;
    a.refr[j] = r;
    if (!BUCKET && __float_as_uint(slot) != 0u) *slotp = 0.0f;
    if (f) atomicAdd(a.nspk + j, 1);
  }
  const uint64_t m = ballot(f);
  if (m) {
    int pos = 0;
    const unsigned long long deg = wave_sum_ll(f ? (long long)(re - rb) : 0ll);
    if (lane == 0) {
      pos = atomicAdd(&s_nf, __popcll(m));
      atomicAdd(&s_ev, deg);
    }
    pos = __shfl(pos, 0, kWave);
    if (f) s_rows[pos + rank_below(m)] = make_int2(rb, re);
    if (a.raster && lane == 0) raster_or(a.raster + (size_t)t * a.words, j - lane, m, a.words);
  }
  __syncthreads();
  const int nf = s_nf;
  if (tid == 0 && nf) {
    atomicAdd(a.blk_spikes + b, (unsigned long long)nf);
    atomicAdd(a.blk_events + b, s_ev);
  }
  for (int k = wave; k < nf; k += BLOCK / kWave) {
    const int2 row = s_rows[k];
    expand_row<BUCKET>(a, row.x, row.y, tm, lane);
  }
}

// One block of NT threads per heavy target (hip-bucket).
template <int NT>
__device__ void hub_block(const StepArgs& a, int t, int tm, int hidx, unsigned char* lds) {
  constexpr int kHubSuperWords = NT * kHubWordsPerThread;
  // s_val holds an entry's column position until the gather replaces it by
  // the entry's value (same thread, same index).
  float* s_val = reinterpret_cast<float*>(lds);
  int* s_wsum = reinterpret_cast<int*>(lds + kHubCap * 4);  // [NT / kWave + 1]
  int& s_fire = s_wsum[NT / kWave + 1];
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int h = a.heavy_ids[hidx];
  const int slot_base = a.n_reg_blocks;  // stats slots after the regular blocks
  // neuron-only loads first (the same addresses in every lane)
  const int c0 = a.ccol[h], c1 = a.ccol[h + 1];
  const int hrb = a.row_ptr[h], hre = a.row_ptr[h + 1];
  float v = a.v[h];
  int r = a.refr[h];
  const float x_in = a.inj_heavy[(size_t)t * a.n_heavy + hidx];
  float in = 0.0f;  // sum of the arriving values, kept by thread 0
  {
    uint32_t* bnow = a.bits + (size_t)tm * a.bits_slot;
    const int w0 = c0 >> 5, w1 = c1 >> 5;
    for (int sb = w0; sb < w1; sb += kHubSuperWords) {  // uniform trip count
      const int nw = min(kHubSuperWords, w1 - sb);
      const int per = (nw + NT - 1) / NT;
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
        for (int q = 0; q < NT / kWave; ++q) {
          const int x = s_wsum[q];
          s_wsum[q] = run;
          run += x;
        }
        s_wsum[NT / kWave] = run;
      }
      __syncthreads();
      const int off = s_wsum[wave] + incl - c;
      const int total = s_wsum[NT / kWave];
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
        for (int q = tid; q < cnt; q += NT) s_val[q] = a.cval[__float_as_int(s_val[q])];
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
      atomicAdd(a.nspk + h, 1);
      atomicAdd(a.blk_spikes + slot_base + hidx, 1ull);
      atomicAdd(a.blk_events + slot_base + hidx, (unsigned long long)(hre - hrb));
      if (a.raster) atomicOr(a.raster + (size_t)t * a.words + (h >> 5), 1u << (h & 31));
    }
  }
  __syncthreads();
  if (s_fire && wave == 0) expand_row<true>(a, hrb, hre, tm, lane);
}

template <int BLOCK>
static void launch_step_block(const FwdPlanCore& pc, const StepArgs& a, int t, int tm) {
  if (pc.variant == SNC_VARIANT_BUCKET)
    hipLaunchKernelGGL((k_step<true, BLOCK>), dim3(a.n_heavy + a.n_reg_blocks, a.B), dim3(BLOCK),
                       0, pc.stream, a, t, tm);
  else
    hipLaunchKernelGGL((k_step<false, BLOCK>), dim3(a.n_reg_blocks, a.B), dim3(BLOCK), 0,
                       pc.stream, a, t, tm);
}

int fused_launch_step(const FwdPlanCore& pc, const StepArgs& a, int t) {
  const int tm = t % a.L;
  switch (a.blk) {
    case 128: launch_step_block<128>(pc, a, t, tm); break;
    case 256: launch_step_block<256>(pc, a, t, tm); break;
    case 512: launch_step_block<512>(pc, a, t, tm); break;
    default: return fail(SNC_E_INVALID, "step block must be 128, 256 or 512 neurons");
  }
  return hip_last("step launch failed");
}

bool fused_block_ok(int block) { return block == 128 || block == 256 || block == 512; }

int fused_light_max_in_degree() { return kLightMaxWords * 32; }

}  // namespace snc_amd
