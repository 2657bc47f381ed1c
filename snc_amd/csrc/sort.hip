// hip-sort: deterministic delivery by an explicit stable sort.
//
// Per step, after the LIF kernel (which also writes the fired bitmap):
//   1. k_sort_compact   one block turns the fired bitmap into the ascending
//                       fired list and the exclusive prefix of their
//                       out-degrees (event offsets).
//   2. k_sort_expand    expands fired rows in that order, so event k carries
//                       CSR index e in ascending order -- the CPU's canonical
//                       delivery order (cpu_backend.cpp:173-181). Key =
//                       post * L + slot.
//   3. LSD radix sort   stable 8-bit passes (histogram, one-block scan,
//                       scatter). Within a block, elements are ranked among
//                       equal digits by wave ballots, so equal keys keep
//                       ascending e.
//   4. k_sort_fold      each run of equal keys is added sequentially, in e
//                       order, onto its ring value: the CPU's sum bit for bit.
// It replaces Thrust sort_by_key + reduce_by_key (cuda_backend.cu:317-328),
// whose segmented reduction has no fixed association order, and expands only
// fired rows instead of all S synapses.

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "hip_err.hpp"
#include "forward_plan.hpp"
#include "wave.hpp"

namespace snc_amd {

constexpr int kSortBlocks = 256;  // blocks of the hist / scatter / fold grids
constexpr int kSortThreads = 256;
constexpr int kScanThreads = 1024;

size_t sort_hist_words(int n) {
  (void)n;
  return (size_t)256 * kSortBlocks + 8;
}

// Exclusive scan of one int per thread over a block of kScanThreads threads.
__device__ int block_exclusive_scan_1024(int x, int* s_wave, int* total) {
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const int incl = wave_inclusive_scan(x);
  if (lane == kWave - 1) s_wave[wave] = incl;
  __syncthreads();
  if (wave == 0) {
    const int nw = kScanThreads / kWave;
    int v = lane < nw ? s_wave[lane] : 0;
    const int vi = wave_inclusive_scan(v);
    if (lane < nw) s_wave[lane] = vi - v;
    if (lane == nw - 1) s_wave[nw] = vi;
  }
  __syncthreads();
  const int r = s_wave[wave] + incl - x;
  *total = s_wave[kScanThreads / kWave];
  __syncthreads();
  return r;
}

// (1) ordered fired list + event offsets; also accounts the step's events.
__global__ __launch_bounds__(kScanThreads) void k_sort_compact(StepArgs a, int t) {
  __shared__ int s_wave[kScanThreads / kWave + 1];
  const uint32_t* fb = a.fbits;  // the fired bitmap of this step
  const int tid = threadIdx.x;
  const int chunk = (a.words + kScanThreads - 1) / kScanThreads;
  const int w0 = min(a.words, tid * chunk), w1 = min(a.words, w0 + chunk);
  int c = 0;
  for (int w = w0; w < w1; ++w) c += __popc(fb[w]);
  int nf = 0;
  int pos = block_exclusive_scan_1024(c, s_wave, &nf);
  for (int w = w0; w < w1; ++w) {
    uint32_t b = fb[w];
    while (b) {
      a.fired[pos++] = (w << 5) | __builtin_ctz(b);
      b &= b - 1u;
    }
  }
  __syncthreads();
  const int fchunk = (nf + kScanThreads - 1) / kScanThreads;
  const int k0 = min(nf, tid * fchunk), k1 = min(nf, k0 + fchunk);
  int deg = 0;
  for (int k = k0; k < k1; ++k) {
    const int i = a.fired[k];
    deg += a.row_ptr[i + 1] - a.row_ptr[i];
  }
  int total = 0;
  int off = block_exclusive_scan_1024(deg, s_wave, &total);
  for (int k = k0; k < k1; ++k) {
    a.ev_ptr[k] = off;
    const int i = a.fired[k];
    off += a.row_ptr[i + 1] - a.row_ptr[i];
  }
  if (tid == 0) {
    a.ev_ptr[nf] = total;
    a.ev_ptr[a.n + 1] = nf;  // fired count of this step
    a.hist[256 * kSortBlocks] = total;
    a.blk_events[0] += (unsigned long long)total;
  }
}

// (2) canonical expansion; also stages step t+1's injections and clears its
// fired counter (the duties of k_expand in the other strategies).
__global__ __launch_bounds__(kSortThreads) void k_sort_expand(StepArgs a, int t, int tm) {
  const int gtid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int gsize = (int)(gridDim.x * blockDim.x);
  if (t + 1 < a.T) {
    const int b = a.inj_ptr[t + 1], e = a.inj_ptr[t + 2];
    for (int k = b + gtid; k < e; k += gsize) a.ext[a.inj_nid[k]] = a.inj_amp[k];
    if (gtid == 0) a.fcnt[(t + 1) & 1] = 0;
  }
  const int nf = a.ev_ptr[a.n + 1];
  const int lane = lane_id(), wave = gtid >> 6, nwaves = gsize >> 6;
  for (int k = wave; k < nf; k += nwaves) {
    const int i = a.fired[k];
    const int rb = a.row_ptr[i], re = a.row_ptr[i + 1];
    const int base = a.ev_ptr[k] - rb;
    for (int e = rb + lane; e < re; e += kWave) {
      int s = tm + (int)a.delay[e];
      if (s >= a.L) s -= a.L;
      a.keys[base + e] = (uint32_t)a.post[e] * (uint32_t)a.L + (uint32_t)s;
      a.vals[base + e] = e;
    }
  }
}

__device__ __forceinline__ void block_range(int count, int& lo, int& hi) {
  const int chunk = (count + kSortBlocks - 1) / kSortBlocks;
  lo = min(count, (int)blockIdx.x * chunk);
  hi = min(count, lo + chunk);
}

// (3a) per-block digit histogram, digit-major: hist[d * kSortBlocks + block].
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                             const int32_t* __restrict__ count_p,
                                                             int shift, int32_t* __restrict__ hist) {
  __shared__ int s_h[256];
  s_h[threadIdx.x] = 0;
  __syncthreads();
  int lo, hi;
  block_range(*count_p, lo, hi);
  for (int k = lo + (int)threadIdx.x; k < hi; k += kSortThreads)
    atomicAdd(&s_h[(keys[k] >> shift) & 255u], 1);
  __syncthreads();
  hist[threadIdx.x * kSortBlocks + blockIdx.x] = s_h[threadIdx.x];
}

// (3b) exclusive scan of the 256 x kSortBlocks histogram in place.
__global__ __launch_bounds__(kScanThreads) void k_radix_scan(int32_t* __restrict__ hist) {
  __shared__ int s_wave[kScanThreads / kWave + 1];
  constexpr int kTotal = 256 * kSortBlocks;
  constexpr int kPer = kTotal / kScanThreads;
  const int base = threadIdx.x * kPer;
  int v[kPer];
  int c = 0;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    v[q] = hist[base + q];
    c += v[q];
  }
  int total = 0;
  int off = block_exclusive_scan_1024(c, s_wave, &total);
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    hist[base + q] = off;
    off += v[q];
  }
}

// (3c) stable scatter. The block walks its range in rounds of kSortThreads
// consecutive elements; an element's destination is the digit's running
// offset + the count of equal digits in earlier waves of the round + its rank
// among equal digits in lower lanes of its wave (eight ballots give the lanes
// sharing all 8 digit bits).
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, int32_t* __restrict__ vout, const int32_t* __restrict__ count_p,
    int shift, const int32_t* __restrict__ hist) {
  constexpr int kWaves = kSortThreads / kWave;
  __shared__ int s_run[256];
  __shared__ int s_wc[kWaves][256];
  const int tid = threadIdx.x, wave = tid >> 6;
  s_run[tid] = hist[tid * kSortBlocks + blockIdx.x];
  for (int w = 0; w < kWaves; ++w) s_wc[w][tid] = 0;
  __syncthreads();
  int lo, hi;
  block_range(*count_p, lo, hi);
  for (int r0 = lo; r0 < hi; r0 += kSortThreads) {
    const int k = r0 + tid;
    const bool valid = k < hi;
    uint32_t key = 0;
    int val = 0;
    uint32_t d = 0;
    if (valid) {
      key = kin[k];
      val = vin[k];
      d = (key >> shift) & 255u;
    }
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = ballot(valid && bit);
      peers &= bit ? m : ~m;
    }
    const int rank = rank_below(peers);
    if (valid && rank == 0) s_wc[wave][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int off = s_run[d] + rank;
      for (int w = 0; w < wave; ++w) off += s_wc[w][d];
      kout[off] = key;
      vout[off] = val;
    }
    __syncthreads();
    int add = 0;
    for (int w = 0; w < kWaves; ++w) {
      add += s_wc[w][tid];
      s_wc[w][tid] = 0;
    }
    s_run[tid] += add;
    __syncthreads();
  }
}

// (4) ordered fold of every run of equal keys onto its ring value.
__global__ __launch_bounds__(kSortThreads) void k_sort_fold(StepArgs a,
                                                            const uint32_t* __restrict__ keys,
                                                            const int32_t* __restrict__ vals) {
  const int count = a.ev_ptr[a.ev_ptr[a.n + 1]];
  const int gsize = (int)(gridDim.x * blockDim.x);
  for (int k = (int)(blockIdx.x * blockDim.x + threadIdx.x); k < count; k += gsize) {
    const uint32_t key = keys[k];
    if (k > 0 && keys[k - 1] == key) continue;
    const uint32_t post = key / (uint32_t)a.L, slot = key - post * (uint32_t)a.L;
    float* dst = a.ring + (size_t)slot * a.n + post;
    float acc = *dst;
    for (int j = k; j < count && keys[j] == key; ++j) acc = __fadd_rn(acc, a.wsig[vals[j]]);
    *dst = acc;
  }
}

int sort_launch_step(const FwdPlanCore& pc, const StepArgs& a, int t) {
  const int tm = t % a.L;
  hipStream_t st = pc.stream;
  int rc = fwd_launch_lif(pc, a, t);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sort_compact, dim3(1), dim3(kScanThreads), 0, st, a, t);
  hipLaunchKernelGGL(k_sort_expand, dim3(kSortBlocks), dim3(kSortThreads), 0, st, a, t, tm);
  // The radix kernels read the step's event count from hist's tail, where
  // k_sort_compact stores it.
  const int32_t* count_p = a.hist + 256 * kSortBlocks;
  const uint32_t* kin = a.keys;
  const int32_t* vin = a.vals;
  uint32_t* kout = a.keys_alt;
  int32_t* vout = a.vals_alt;
  for (int shift = 0; shift < pc.sort_bits; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(kSortBlocks), dim3(kSortThreads), 0, st, kin, count_p,
                       shift, a.hist);
    hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(kScanThreads), 0, st, a.hist);
    hipLaunchKernelGGL(k_radix_scatter, dim3(kSortBlocks), dim3(kSortThreads), 0, st, kin, vin,
                       kout, vout, count_p, shift, a.hist);
    const uint32_t* tk = kout;
    const int32_t* tv = vout;
    kout = const_cast<uint32_t*>(kin);
    vout = const_cast<int32_t*>(vin);
    kin = tk;
    vin = tv;
  }
  hipLaunchKernelGGL(k_sort_fold, dim3(kSortBlocks), dim3(kSortThreads), 0, st, a, kin, vin);
  return hip_last("sort step launch failed");
}

}  // namespace snc_amd
