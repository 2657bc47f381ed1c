// hip-sort: deterministic delivery by sorting each step's arrivals.
//
// Sort key. Every synapse has a column position cpos[e]: its rank in the
// order (post, delay descending, pre, CSR position) of the padded incoming
// columns. The deposits that arrive at one target at one step come from
// steps t - d in order of their origin (longest delay first) and, within an
// origin step, in (pre, CSR position) order -- the order the CPU path adds
// them onto the ring slot it reads at step t (cpu_backend.cpp:173-181). That
// is column order, so sorting a step's arrivals by column position and
// adding each target's run from zero is the CPU's slot sum bit for bit. This
// replaces the reference's Thrust sort_by_key + reduce_by_key over all S
// synapses (cuda_backend.cu:138-169, 317-328), whose segmented reduction has
// no fixed association order.
//
// Targets are cut into segments: runs of at most 1024 targets, starting on a
// multiple of 32, whose columns span at most kWin positions. Each (ring slot,
// segment) pair has an arrival queue of records {column position, value}: a
// deposit made at step t with delay d is appended to the queue of slot
// (t + d) % L of its target's segment. One launch per step, one workgroup per
// segment (k_sort_step):
//   1. take the segment's queue of slot t % L (and reset it: the next
//      deposits into that slot are made at step t + 1 or later),
//   2. sort it by counting: presence bits of the column positions in LDS,
//      an exclusive popcount prefix over the words (every record's rank),
//      the values staged at their ranks,
//   3. each target adds its run of staged values in order, from zero, and
//      integrates/fires for step t (its injections come from the step's
//      per-segment range),
//   4. expand the segment's fired rows: one atomic reservation and one
//      8-byte record store per deposit (queue order is arbitrary; step 2
//      sorts it).
// A segment longer than kWin positions (a hub's column) is sorted one window
// of kWin positions at a time, and a window with more records than the
// staging holds is added in rank order, kStage values at a time; both keep
// every target's sum in column order. There is no ring: a queue holds the
// deposits themselves, so no target's partial sum is read back or written.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "fold.hpp"
#include "forward_plan.hpp"
#include "hip_err.hpp"
#include "wave.hpp"

namespace snc_amd {

constexpr int kSortThreads = 1024;  // = most targets of a segment
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kWin = 65536;         // column positions per window
constexpr int kWinWords = kWin / 32;
constexpr int kStage = 4096;        // staged values per round
constexpr int kCache = 4096;        // records kept in LDS
constexpr int kHead = 512;          // records loaded before the queue's count is known

// Dynamic LDS of k_sort_step: presence bits, word prefixes (+ wave sums),
// cached records, staged values, per-target staging (injections, fired
// rows).
constexpr int kPrefInts = kWinWords + 1 + kSortWaves + 15;
constexpr size_t kSortLds = sizeof(uint32_t) * kWinWords + sizeof(int) * kPrefInts +
                            sizeof(uint2) * kCache + sizeof(float) * kStage +
                            sizeof(float) * kSortThreads + sizeof(int2) * kSortThreads;
static_assert(kPrefInts % 4 == 0, "the record cache must stay 16-byte aligned");
static_assert(kSortLds <= 80 * 1024, "two hip-sort workgroups must fit in the CU's 160 KB of LDS");

// ---------------------------------------------------------------------------
// host: segments and per-segment injection ranges
// ---------------------------------------------------------------------------

int sort_build_segments(const std::vector<int32_t>& ccol, int n, std::vector<int4>& seg,
                        std::vector<int32_t>& seg_of_word) {
  seg.clear();
  const int64_t P = ccol[n];
  seg_of_word.assign((size_t)(P / 32), 0);
  int j0 = 0;
  while (j0 < n) {
    // grow by groups of 32 targets while the window and target limits hold;
    // a first group that alone exceeds the window becomes its own segment
    int j1 = std::min(n, j0 + 32);
    while (j1 < n && j1 - j0 < kSortThreads) {
      const int jn = std::min(n, j1 + 32);
      if ((int64_t)ccol[jn] - ccol[j0] > kWin || jn - j0 > kSortThreads) break;
      j1 = jn;
    }
    const int s = (int)seg.size();
    seg.push_back(make_int4(j0, j1, ccol[j0], ccol[j1]));
    for (int64_t w = ccol[j0] / 32; w < ccol[j1] / 32; ++w) seg_of_word[(size_t)w] = s;
    j0 = j1;
  }
  return SNC_OK;
}

std::vector<int32_t> sort_injection_ranges(const std::vector<int4>& seg,
                                           const std::vector<int32_t>& inj_ptr,
                                           const std::vector<int32_t>& inj_nid, int T) {
  const size_t S1 = seg.size() + 1;
  std::vector<int32_t> r((size_t)std::max(1, T) * S1, 0);
  for (int t = 0; t < T; ++t) {
    const auto b = inj_nid.begin() + inj_ptr[t], e = inj_nid.begin() + inj_ptr[t + 1];
    for (size_t s = 0; s < seg.size(); ++s)
      r[(size_t)t * S1 + s] = (int32_t)(std::lower_bound(b, e, seg[s].x) - inj_nid.begin());
    r[(size_t)t * S1 + seg.size()] = inj_ptr[t + 1];
  }
  return r;
}

// ---------------------------------------------------------------------------
// device
// ---------------------------------------------------------------------------

// One step t of one segment. Eight waves per SIMD (at most 64 VGPRs and 96
// SGPRs; the kernel arguments spill to VGPR lanes): two workgroups fit on a
// CU, LDS included, so graphs of up to 512 segments run every segment at
// once on the 256 CUs.
__global__ __launch_bounds__(kSortThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void
k_sort_step(StepArgs a, int t) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t* s_bits = (uint32_t*)smem;                 // [kWinWords]
  int* s_pref = (int*)(s_bits + kWinWords);           // [kWinWords + 1]
  int* s_wave = s_pref + kWinWords + 1;               // [kSortWaves]
  uint2* s_rec = (uint2*)(s_pref + kPrefInts);        // [kCache]
  float* s_val = (float*)(s_rec + kCache);            // [kStage]
  float* s_ext = s_val + kStage;                      // [kSortThreads] this step's injections
  int2* s_rows = (int2*)(s_ext + kSortThreads);       // [kSortThreads] fired rows
  __shared__ int s_count;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int sg = blockIdx.x;
  const int4 g = a.sseg[sg];  // targets [x, y), column positions [z, w)
  const int j = g.x + tid;    // this thread's target
  const bool own = j < g.y;
  const int tm = t % a.L;
  // this step's injection range of the segment, with the first loads
  const int* ir = a.sinj_ptr + (size_t)t * (a.n_seg + 1) + sg;
  const int inj0 = ir[0], inj1 = ir[1];
  // the own target's state, loaded before the sort
  float v = 0.0f, in = 0.0f;
  int refr = 0, cj0 = 0, cj1 = 0, rb0 = 0, re0 = 0;
  if (own) {
    cj0 = a.ccol[j];
    cj1 = a.ccol[j + 1];
    v = a.v[j];
    refr = a.refr[j];
    rb0 = a.row_ptr[j];  // the own row, expanded below if the target fires
    re0 = a.row_ptr[j + 1];
  }
  // The queue of slot tm; its first kHead records are loaded before the
  // count is known (within the queue's capacity, so in bounds).
  const uint2* q = a.slist + (size_t)tm * a.slist_stride + g.z;
  const int cap = g.w - g.z;
  uint2 head = make_uint2(0u, 0u);
  if (tid < min(cap, kHead)) head = q[tid];
  // the first injection entry of this thread, loaded in the same round
  int inj_k = inj0 + tid, inj_n = -1;
  float inj_a = 0.0f;
  if (inj_k < inj1) {
    inj_n = a.inj_nid[inj_k];
    inj_a = a.inj_amp[inj_k];
  }
  s_ext[tid] = 0.0f;
  if (tid == 0) {
    // the counter holds the next free queue position, absolute (from the
    // segment's first column position), so producers need no segment table
    int32_t* c = a.scnt + ((size_t)tm * a.n_seg + sg) * kSortCntStride;
    s_count = *c - g.z;
    *c = g.z;  // deposits into slot tm resume at step t + 1
  }
  __syncthreads();
  // this step's injections into the segment's targets
  if (inj_n >= 0) s_ext[inj_n - g.x] = inj_a;
  for (int k = inj_k + kSortThreads; k < inj1; k += kSortThreads) s_ext[a.inj_nid[k] - g.x] = a.inj_amp[k];
  const int count = s_count;
  const bool cached = count <= kCache;
  if (cached) {
    if (tid < min(count, kHead)) s_rec[tid] = head;
    for (int k = tid < kHead ? tid + kSortThreads : tid; k < count; k += kSortThreads) s_rec[k] = q[k];
  }
  // record k is only ever read by thread k % kSortThreads, its writer
  auto rec = [&](int k) { return cached ? s_rec[k] : q[k]; };
  for (int wp = g.z; count > 0 && wp < g.w; wp += kWin) {
    const int wend = min(wp + kWin, g.w);
    const int nw = (wend - wp) >> 5;
    for (int w = tid; w < nw; w += kSortThreads) s_bits[w] = 0u;
    __syncthreads();
    for (int k = tid; k < count; k += kSortThreads) {
      const int p = (int)rec(k).x;
      if (p >= wp && p < wend) atomicOr(&s_bits[(p - wp) >> 5], 1u << (p & 31));
    }
    __syncthreads();
    // exclusive popcount prefix of the window's words: the sorted rank
    constexpr int kPer = kWinWords / kSortThreads;
    int c = 0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int w = tid * kPer + u;
      c += w < nw ? __popc(s_bits[w]) : 0;
    }
    const int incl = wave_inclusive_scan(c);
    if (lane == kWave - 1) s_wave[wave] = incl;
    __syncthreads();
    int run = incl - c;
    for (int k = 0; k < wave; ++k) run += s_wave[k];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int w = tid * kPer + u;
      if (w < nw) {
        s_pref[w] = run;
        run += __popc(s_bits[w]);
      }
    }
    if (tid == kSortThreads - 1) s_pref[nw] = run;  // the last thread's running total
    __syncthreads();
    const int total = s_pref[nw];
    // this thread's target: its arrivals are the ranks of its column's words
    int tr0 = 0, tr1 = 0;
    if (own) {
      const int a0 = max(cj0, wp), a1 = min(cj1, wend);
      if (a0 < a1) {
        tr0 = s_pref[(a0 - wp) >> 5];
        tr1 = s_pref[(a1 - wp) >> 5];
      }
    }
    for (int r0 = 0; r0 < total; r0 += kStage) {
      const int r1 = min(total, r0 + kStage);
      for (int k = tid; k < count; k += kSortThreads) {
        const uint2 e = rec(k);
        const int p = (int)e.x;
        if (p < wp || p >= wend) continue;
        const int w = (p - wp) >> 5;
        const int rk = s_pref[w] + __popc(s_bits[w] & ((1u << (p & 31)) - 1u));
        if (rk >= r0 && rk < r1) s_val[rk - r0] = __uint_as_float(e.y);
      }
      __syncthreads();
      // the own run, in column order, four staged values per round
      int r = max(tr0, r0) - r0;
      const int re = min(tr1, r1) - r0;
      for (; r + 4 <= re; r += 4) {
        const float x0 = s_val[r], x1 = s_val[r + 1], x2 = s_val[r + 2], x3 = s_val[r + 3];
        in = __fadd_rn(__fadd_rn(__fadd_rn(__fadd_rn(in, x0), x1), x2), x3);
      }
      for (; r < re; ++r) in = __fadd_rn(in, s_val[r]);
      __syncthreads();
    }
  }
  // integrate/fire of step t (the barrier orders the injections' LDS writes
  // before their reads when no sort ran)
  __syncthreads();
  bool f = false;
  if (own) {
    f = lif_step(v, refr, s_ext[tid], in, a.lif);
    a.v[j] = v;
    a.refr[j] = refr;
    if (f) a.nspk[j] += 1;
  }
  // the fired list in id order: wave counts, their prefix, each lane's rank
  const uint64_t m = ballot(f);
  if (lane == 0) s_wave[wave] = __popcll(m);
  __syncthreads();
  int off = 0, nf = 0;
  for (int k = 0; k < kSortWaves; ++k) {
    off += k < wave ? s_wave[k] : 0;
    nf += s_wave[k];
  }
  if (f) s_rows[off + rank_below(m)] = make_int2(rb0, re0);
  // The segment starts on a multiple of 32 targets and only own targets
  // fire, so every nonzero raster word of a wave is the segment's own.
  if (a.raster && lane == 0 && m) {
    uint32_t* row = a.raster + (size_t)t * a.words + ((g.x + wave * kWave) >> 5);
    if ((uint32_t)m) row[0] = (uint32_t)m;
    if ((uint32_t)(m >> 32)) row[1] = (uint32_t)(m >> 32);
  }
  __syncthreads();
  // Expansion of the fired rows, flattened over the workgroup: deposit k of
  // the rows' concatenation goes to thread k % kSortThreads, kExpand per
  // thread per round, so a segment with many fired rows still pays one round
  // of loads, one of reservations and one of stores. The exclusive prefix of
  // the row lengths (in the sort's presence bits, free now) maps k to its
  // row. Deposits of the last step would never be read and are skipped.
  constexpr int kExpand = 4;
  int* s_pre = (int*)s_bits;
  const int my_len = tid < nf ? s_rows[tid].y - s_rows[tid].x : 0;
  const int my_inc = wave_inclusive_scan(my_len);
  if (lane == kWave - 1) s_wave[wave] = my_inc;  // every read of the fired counts came before
  __syncthreads();
  int below = 0, events = 0;
  for (int k = 0; k < kSortWaves; ++k) {
    below += k < wave ? s_wave[k] : 0;
    events += s_wave[k];
  }
  if (tid < nf) s_pre[tid] = below + my_inc - my_len;
  __syncthreads();
  if (t + 1 < a.T) {
    for (int k0 = 0; k0 < events; k0 += kExpand * kSortThreads) {
      int p[kExpand], sgu[kExpand], sl[kExpand], idx[kExpand];
      float w[kExpand];
      bool valid[kExpand];
#pragma unroll
      for (int u = 0; u < kExpand; ++u) {
        const int k = k0 + u * kSortThreads + tid;
        valid[u] = k < events;
        int e = 0;
        if (valid[u]) {
          int lo = 0, hi = nf - 1;  // the last row starting at or before k
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_pre[mid] <= k) lo = mid;
            else hi = mid - 1;
          }
          e = s_rows[lo].x + (k - s_pre[lo]);
          p[u] = a.cpos[e];
          w[u] = a.wsig[e];
          sgu[u] = a.sseg_e[e];
          sl[u] = tm + (int)a.delay[e];
        }
      }
#pragma unroll
      for (int u = 0; u < kExpand; ++u)
        if (valid[u]) {
          if (sl[u] >= a.L) sl[u] -= a.L;
          idx[u] = atomicAdd(a.scnt + ((size_t)sl[u] * a.n_seg + sgu[u]) * kSortCntStride, 1);
        }
#pragma unroll
      for (int u = 0; u < kExpand; ++u)
        if (valid[u])
          a.slist[(size_t)sl[u] * a.slist_stride + idx[u]] =
              make_uint2((uint32_t)p[u], __float_as_uint(w[u]));
    }
  }
  if (tid == 0) {
    if (nf) a.blk_spikes[a.sort_slot0 + sg] += (unsigned long long)nf;
    if (events) a.blk_events[a.sort_slot0 + sg] += (unsigned long long)events;
  }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------

int sort_prepare() {
  const hipError_t e = hipFuncSetAttribute((const void*)k_sort_step,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)kSortLds);
  return e == hipSuccess ? SNC_OK : hip_fail(e, "LDS request of the hip-sort step");
}

int sort_launch_step(const FwdPlanCore& pc, const StepArgs& a, int t) {
  hipLaunchKernelGGL(k_sort_step, dim3(a.n_seg), dim3(kSortThreads), kSortLds, pc.stream, a, t);
  return hip_last("sort step launch failed");
}

}  // namespace snc_amd
