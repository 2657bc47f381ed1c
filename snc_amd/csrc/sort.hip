// hip-sort: deterministic delivery by sorting each step's events.
//
// Sort key. Every synapse has a column position cpos[e]: its rank in the
// order (post, delay descending, pre, CSR position) of the padded incoming
// columns. For one step, the deposits onto one (target, ring slot) share a
// delay, so sorting the step's events by column position puts them in
// (pre, CSR position) order -- the order the CPU path adds them
// (cpu_backend.cpp:173-181) -- and the sorted runs of a target are its ring
// slots' runs. Adding each run, in that order, onto the slot's value is the
// CPU's sum bit for bit. This replaces the reference's Thrust sort_by_key +
// reduce_by_key over all S synapses (cuda_backend.cu:138-169, 317-328), whose
// segmented reduction has no fixed association order.
//
// Targets are cut into segments: runs of at most 1024 targets, starting on a
// multiple of 32, whose columns span at most kWin positions. Each segment has
// an event list per step parity. One launch per step, one workgroup per
// segment (k_sort_step):
//   1. sort the segment's events of step t by counting: presence bits in
//      LDS, an exclusive popcount prefix over the words (every event's sorted
//      rank), the events' values and slots staged at their ranks;
//   2. fold each target's sorted runs onto the ring: one thread per target,
//      or, for a target with many events this step, one wave whose lanes take
//      one run (one ring slot) each;
//   3. integrate/fire the segment's targets for step t + 1 (their injections
//      come from the step's per-segment range);
//   4. expand the segment's fired rows: every event of step t + 1 (its
//      column position) is appended to its target segment's list of the
//      other parity, with one atomic per distinct segment per wave (the list
//      order is arbitrary; step 1 sorts it).
// A launch before step 0 does 3 and 4 for step 0. A segment longer than kWin
// positions (a hub's column) is sorted one window of kWin positions at a
// time, and a window with more events than the staging holds is folded in
// rank order, kStage events at a time; both keep every (target, slot) sum in
// sorted order.

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
constexpr int kStage = 4096;        // staged events per fold round
constexpr int kCache = 4096;        // event positions kept in LDS
constexpr int kLightRun = 16;       // a target with more events this round folds by wave

// Dynamic LDS of k_sort_step: presence bits, word prefixes (+ wave sums),
// cached positions, staged events, per-target staging (next-step drive,
// wave-fold list, fired list).
constexpr int kPrefInts = kWinWords + 1 + kSortWaves + 15;
constexpr size_t kSortLds = sizeof(uint32_t) * kWinWords + sizeof(int) * kPrefInts +
                            sizeof(uint32_t) * kCache + sizeof(uint2) * kStage +
                            sizeof(float) * kSortThreads + 2 * sizeof(int32_t) * kSortThreads;
static_assert(kPrefInts % 4 == 0, "the position cache must stay 16-byte aligned");
static_assert(kSortLds <= 160 * 1024, "hip-sort workgroup LDS above the CU's 160 KB");

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

// Ordered fold of target j's staged events [r0, r1) (column order) onto its
// ring slots: consecutive events of one slot are one run, added in order onto
// the slot's value; the slot values of up to kBatch runs load together. Slot
// s1 (the next step's, which the same thread integrates right after) is not
// read or written here: its value lives in the register slot1.
__device__ __forceinline__ void fold_sorted(const uint2* s_ev, int r0, int r1, float* ring,
                                            int n, int j, int s1, float& slot1) {
  constexpr int kBatch = 8;
  int cur = -1;
  float acc = 0.0f;
  auto put = [&](int sl, float v) {
    if (sl == s1) slot1 = v;
    else ring[(size_t)sl * n + j] = v;
  };
  for (int r = r0; r < r1; r += kBatch) {
    int sl[kBatch];
    float val[kBatch], old[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      sl[u] = -1;
      val[u] = 0.0f;
      if (r + u < r1) {
        val[u] = __uint_as_float(s_ev[r + u].x);
        sl[u] = (int)s_ev[r + u].y;
      }
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
      old[u] = sl[u] >= 0 && sl[u] != cur && sl[u] != s1 ? ring[(size_t)sl[u] * n + j] : 0.0f;
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (sl[u] < 0) break;
      if (sl[u] != cur) {
        if (cur >= 0) put(cur, acc);
        cur = sl[u];
        acc = cur == s1 ? slot1 : old[u];
      }
      acc = __fadd_rn(acc, val[u]);
    }
  }
  if (cur >= 0) put(cur, acc);
}

// Wave fold of target j's staged events [r0, r1). Within one step a target's
// events of one ring slot are one contiguous run (column order is delay
// descending), so the wave marks where each run starts, 64 positions at a
// time, and the lane holding a start adds that run, in order, onto its
// slot's value: the runs (one per slot) proceed in parallel.
__device__ __forceinline__ void fold_wave(const uint2* s_ev, int r0, int r1, float* ring, int n,
                                          int j) {
  const int lane = lane_id();
  for (int base = r0; base < r1; base += kWave) {
    const int r = base + lane;
    if (r >= r1) continue;
    const uint32_t sl = s_ev[r].y;
    if (r > r0 && s_ev[r - 1].y == sl) continue;  // not the start of a run
    float* dst = ring + (size_t)sl * n + j;
    float acc = *dst;
    int q = r;
    for (; q + 4 <= r1 && s_ev[q + 3].y == sl; q += 4) {
      acc = __fadd_rn(acc, __uint_as_float(s_ev[q].x));
      acc = __fadd_rn(acc, __uint_as_float(s_ev[q + 1].x));
      acc = __fadd_rn(acc, __uint_as_float(s_ev[q + 2].x));
      acc = __fadd_rn(acc, __uint_as_float(s_ev[q + 3].x));
    }
    for (; q < r1 && s_ev[q].y == sl; ++q) acc = __fadd_rn(acc, __uint_as_float(s_ev[q].x));
    *dst = acc;
  }
}

// One step of one segment; t = -1 is the launch before step 0 (no fold).
// Eight waves per SIMD (at most 64 VGPRs and 96 SGPRs; the kernel arguments
// spill to VGPR lanes): two workgroups fit on a CU, LDS included, so graphs of
// up to 512 segments run every segment at once on the 256 CUs.
__global__ __launch_bounds__(kSortThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void
k_sort_step(StepArgs a, int t) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t* s_bits = (uint32_t*)smem;                        // [kWinWords]
  int* s_pref = (int*)(s_bits + kWinWords);                  // [kWinWords + 1]
  int* s_wave = s_pref + kWinWords + 1;                      // [kSortWaves]
  uint32_t* s_pos = (uint32_t*)(s_pref + kPrefInts);         // [kCache]
  uint2* s_ev = (uint2*)(s_pos + kCache);                    // [kStage]
  float* s_ext = (float*)(s_ev + kStage);                    // [kSortThreads] next-step drive
  int32_t* s_heavy = (int32_t*)(s_ext + kSortThreads);       // [kSortThreads] wave-fold list, then row ends
  int32_t* s_fired = s_heavy + kSortThreads;                 // [kSortThreads] wave-fold ids, then row starts
  __shared__ int s_count, s_nheavy;
  __shared__ unsigned long long s_ev_total;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int sg = blockIdx.x;
  const int4 g = a.sseg[sg];  // targets [x, y), column positions [z, w)
  const int j = g.x + tid;    // this thread's target
  const bool own = j < g.y;
  const int tm = t < 0 ? -1 : t % a.L;  // slot of step t (t = -1: none)
  const bool lif = t + 1 < a.T;
  int s1 = tm + 1;                      // slot of step t + 1
  if (s1 >= a.L) s1 -= a.L;
  const int par = t & 1;                // list parity of step t's events (t >= 0)
  // the next step's state of the own target, loaded before the sort
  float v = 0.0f, slot1 = 0.0f;
  int refr = 0, cj0 = 0, cj1 = 0, rb0 = 0, re0 = 0;
  if (own) {
    cj0 = a.ccol[j];
    cj1 = a.ccol[j + 1];
    if (lif) {
      v = a.v[j];
      refr = a.refr[j];
      rb0 = a.row_ptr[j];  // the own row, expanded below if the target fires
      re0 = a.row_ptr[j + 1];
    }
    slot1 = a.ring[(size_t)s1 * a.n + j];
  }
  s_ext[tid] = 0.0f;
  if (tid == 0) {
    // the counters hold the next free list position (absolute, from the
    // segment's first column position), so appends need no segment table
    s_count = 0;
    if (t >= 0) {
      int32_t* cnt = a.scnt + (size_t)(par * a.n_seg + sg) * kSortCntStride;
      s_count = *cnt - g.z;
      *cnt = g.z;  // step t + 2's expansion appends from the start
    }
    s_nheavy = 0;
    s_ev_total = 0ull;
  }
  __syncthreads();
  if (lif) {  // step t + 1's injections into this segment's targets
    const int* r = a.sinj_ptr + (size_t)(t + 1) * (a.n_seg + 1) + sg;
    for (int k = r[0] + tid; k < r[1]; k += kSortThreads) s_ext[a.inj_nid[k] - g.x] = a.inj_amp[k];
  }
  const int count = s_count;
  const int32_t* lst = a.slist + (size_t)par * a.slist_stride + g.z;
  const bool cached = count <= kCache;
  if (cached)
    for (int k = tid; k < count; k += kSortThreads) s_pos[k] = (uint32_t)lst[k];
  auto pos = [&](int k) { return cached ? (int)s_pos[k] : lst[k]; };
  for (int wp = g.z; count > 0 && wp < g.w; wp += kWin) {
    const int wend = min(wp + kWin, g.w);
    const int nw = (wend - wp) >> 5;
    for (int q = tid; q < nw; q += kSortThreads) s_bits[q] = 0u;
    __syncthreads();
    for (int k = tid; k < count; k += kSortThreads) {
      const int p = pos(k);
      if (p >= wp && p < wend) atomicOr(&s_bits[(p - wp) >> 5], 1u << (p & 31));
    }
    __syncthreads();
    // exclusive popcount prefix of the window's words: the sorted rank
    constexpr int kPer = kWinWords / kSortThreads;
    int c = 0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int q = tid * kPer + u;
      c += q < nw ? __popc(s_bits[q]) : 0;
    }
    const int incl = wave_inclusive_scan(c);
    if (lane == kWave - 1) s_wave[wave] = incl;
    __syncthreads();
    int run = incl - c;
    for (int k = 0; k < wave; ++k) run += s_wave[k];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int q = tid * kPer + u;
      if (q < nw) {
        s_pref[q] = run;
        run += __popc(s_bits[q]);
      }
    }
    if (tid == kSortThreads - 1) s_pref[nw] = run;  // the last thread's running total
    __syncthreads();
    const int total = s_pref[nw];
    // this thread's target: its events are the ranks of its column's words
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
        const int p = pos(k);
        if (p < wp || p >= wend) continue;
        const int q = (p - wp) >> 5;
        const int rk = s_pref[q] + __popc(s_bits[q] & ((1u << (p & 31)) - 1u));
        if (rk < r0 || rk >= r1) continue;
        int s = tm + (int)a.cdel[p];
        if (s >= a.L) s -= a.L;
        s_ev[rk - r0] = make_uint2(__float_as_uint(a.cval[p]), (uint32_t)s);
      }
      __syncthreads();
      const int f0 = max(tr0, r0), f1 = min(tr1, r1);
      const bool heavy = f1 - f0 > kLightRun;
      if (heavy) {
        // the wave fold reads and writes every slot in the ring: the register
        // copy of slot s1 goes back first, and comes back after it
        a.ring[(size_t)s1 * a.n + j] = slot1;
        // the target's ranks in this round (both below kStage) and its id;
        // the fired list's space is free until the integrate/fire
        const int h = atomicAdd(&s_nheavy, 1);
        s_heavy[h] = (f0 - r0) | ((f1 - r0) << 16);
        s_fired[h] = j;
      } else if (f0 < f1) {
        fold_sorted(s_ev, f0 - r0, f1 - r0, a.ring, a.n, j, s1, slot1);
      }
      __syncthreads();
      const int nh = s_nheavy;
      for (int h = wave; h < nh; h += kSortWaves) {
        const int e = s_heavy[h];
        fold_wave(s_ev, e & 0xffff, e >> 16, a.ring, a.n, s_fired[h]);
      }
      __syncthreads();
      if (heavy) slot1 = a.ring[(size_t)s1 * a.n + j];
      if (tid == 0) s_nheavy = 0;
      __syncthreads();
    }
  }
  if (!lif) {
    // last step: the deposits onto slot s1 stay in the ring
    if (own) a.ring[(size_t)s1 * a.n + j] = slot1;
    return;
  }
  // integrate/fire of the segment's targets for step t + 1 (the barrier
  // orders the injections' LDS writes before their reads when no fold ran)
  __syncthreads();
  bool f = false;
  if (own) {
    f = lif_step(v, refr, s_ext[tid], slot1, a.lif);
    a.v[j] = v;
    a.refr[j] = refr;
    a.ring[(size_t)s1 * a.n + j] = 0.0f;
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
  if (f) {
    const int q = off + rank_below(m);
    s_fired[q] = rb0;  // the fired rows, in id order: CSR bounds
    s_heavy[q] = re0;
  }
  // The segment starts on a multiple of 32 targets, so a wave's raster words
  // are its own.
  if (a.raster && lane == 0 && m) {
    uint32_t* row = a.raster + (size_t)(t + 1) * a.words + ((g.x + wave * kWave) >> 5);
    if ((uint32_t)m) row[0] = (uint32_t)m;
    if ((uint32_t)(m >> 32)) row[1] = (uint32_t)(m >> 32);
  }
  __syncthreads();
  // expansion of the fired rows into step t + 1's lists: wave per row, each
  // lane two events in flight
  const int npar = (t + 1) & 1;
  int32_t* nlist = a.slist + (size_t)npar * a.slist_stride;
  int32_t* ncnt = a.scnt + (size_t)npar * a.n_seg * kSortCntStride;
  unsigned long long ev = 0;
  for (int k = wave; k < nf; k += kSortWaves) {
    const int rb = s_fired[k], re = s_heavy[k];
    ev += (unsigned long long)(re - rb);
    for (int e0 = rb; e0 < re; e0 += 2 * kWave) {
      int p[2], sgu[2], leader[2], rank[2], base[2];
      bool valid[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int e = e0 + u * kWave + lane;
        valid[u] = e < re;
        p[u] = valid[u] ? a.cpos[e] : 0;
        sgu[u] = valid[u] ? a.sseg_e[e] : -1;
      }
      // group the wave's events by segment: every lane learns its group's
      // leader and its rank in the group; leaders reserve list space with one
      // atomic each, both halves in one round trip
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        uint64_t rem = ballot(valid[u]);
        int cnt = 0;
        leader[u] = 0;
        rank[u] = 0;
        while (rem) {
          const int l = __builtin_ctzll(rem);
          const int lseg = __shfl(sgu[u], l, kWave);
          const uint64_t gm = ballot(valid[u] && sgu[u] == lseg);
          if ((gm >> lane) & 1ull) {
            leader[u] = l;
            rank[u] = rank_below(gm);
          }
          if (lane == l) cnt = __popcll(gm);
          rem &= ~gm;
        }
        base[u] = valid[u] && lane == leader[u] ? atomicAdd(ncnt + (size_t)sgu[u] * kSortCntStride, cnt) : 0;
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int bs = __shfl(base[u], leader[u], kWave);
        if (valid[u]) nlist[bs + rank[u]] = p[u];
      }
    }
  }
  if (lane == 0 && ev) atomicAdd(&s_ev_total, ev);
  __syncthreads();
  if (tid == 0) {
    if (nf) a.blk_spikes[a.sort_slot0 + sg] += (unsigned long long)nf;
    if (s_ev_total) a.blk_events[a.sort_slot0 + sg] += s_ev_total;
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
  hipStream_t st = pc.stream;
  if (t == 0)  // step 0's integrate/fire and expansion
    hipLaunchKernelGGL(k_sort_step, dim3(a.n_seg), dim3(kSortThreads), kSortLds, st, a, -1);
  hipLaunchKernelGGL(k_sort_step, dim3(a.n_seg), dim3(kSortThreads), kSortLds, st, a, t);
  return hip_last("sort step launch failed");
}

}  // namespace snc_amd
