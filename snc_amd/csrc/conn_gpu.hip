// Device half of the sheet connectome's structural step (see conn_gpu.hpp).
//
// Prune selection: 64-bit keys (float bits of |w| << 32 | position), the
// K smallest found by an 8-pass radix select (one byte per pass, most
// significant first, a 256-bin histogram per pass) and marked.
//
// Grow candidates: one wave per freed post scans the bounding boxes of its
// window disc on the grids of layer l-1 and (hidden posts) layer l in id
// order, keeps the cells within the radius (distance in double, as the host),
// compacts them in order with ballots, then computes the window maximum and
// each candidate's rank (fire descending, id ascending) and writes the 16
// features as bf16 (the scorer's input precision); the MLP scores them where
// they lie. The host gets back pres, scores, ranks and window maxima, from
// which it rebuilds the float features of the candidates it grows.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "conn_gpu.hpp"
#include "hip_err.hpp"
#include "mlp.hpp"
#include "wave.hpp"

namespace snc_amd {

namespace {

constexpr int kCandCap = 4096;  // window candidates of one post (LDS list)

template <class T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t count) {
    if (hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
      p = nullptr;
      hip_clear();
      return fail(SNC_E_NOMEM, "structural step: device allocation failed");
    }
    return SNC_OK;
  }
  int upload(const T* h, size_t count) {
    int rc = alloc(count);
    if (!rc && count) {
      const hipError_t e = hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice);
      if (e != hipSuccess) rc = hip_fail(e, "structural step upload");
    }
    return rc;
  }
};

__global__ void k_prune_keys(const float* __restrict__ w, const int32_t* __restrict__ age,
                             int64_t E, int protect, uint64_t* __restrict__ key) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  key[e] = age[e] >= protect ? ((uint64_t)__float_as_uint(fabsf(w[e])) << 32) | (uint64_t)e
                             : ~0ull;
}

// Histogram of byte `shift / 8` of the keys whose bits above it equal prefix.
__global__ void k_prune_hist(const uint64_t* __restrict__ key, int64_t E, uint64_t prefix,
                             int shift, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  for (int b = threadIdx.x; b < 256; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const uint64_t hi_mask = shift >= 56 ? 0ull : ~0ull << (shift + 8);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = key[e];
    if ((k & hi_mask) == prefix) atomicAdd(&h[(k >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 256; b += blockDim.x)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}

__global__ void k_prune_mark(const uint64_t* __restrict__ key, int64_t E, uint64_t last,
                             uint8_t* __restrict__ mask) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < E) mask[e] = key[e] <= last ? 1 : 0;
}

struct CandArgs {
  int n, num_layers;
  float window;
  const int32_t* layer;
  const int32_t* layer_begin;
  const int32_t* layer_count;
  const int32_t* side;
  const float* px;
  const float* py;
  const double* fire;
  const double* rate;
  const float *f5, *f6, *f7, *f8, *f9, *f10, *f14;
  float f4;
  const int32_t* freed;
  const int32_t* offs;   // [n_freed + 1] (pass 2)
  int32_t* count;        // [n_freed] (pass 1)
  int32_t* cand;         // [M] pre of each candidate (pass 2)
  uint16_t* feat;        // [M][16] features as bf16 bits, the scorer's input (pass 2)
  float* rank;           // [M] window rank of each candidate, feature 13 (pass 2)
  double* wmax;          // [n_freed] largest fire count in each post's window (pass 2)
};

__device__ __forceinline__ double dist2d(const CandArgs& a, int q, int p) {
  const double dx = (double)a.px[q] - (double)a.px[p], dy = (double)a.py[q] - (double)a.py[p];
  return sqrt(dx * dx + dy * dy);
}

// Visits the window members of post p in order (layer l-1 then own layer for
// hidden posts; ids ascending), 64 cells per round; emit(q, rank_in_round,
// round_count) runs on every lane of the wave after each round.
template <class Emit>
__device__ void scan_window(const CandArgs& a, int p, Emit emit) {
  const int lane = lane_id();
  const int l = a.layer[p];
  if (l < 1) return;
  const bool hidden = l < a.num_layers - 1;
  for (int ll = l - 1; ll <= (hidden ? l : l - 1); ++ll) {
    const int side = a.side[ll], cnt = a.layer_count[ll], b0 = a.layer_begin[ll];
    const double sc = side > 1 ? (double)(side - 1) : 1.0;
    const double w = (double)a.window;
    const int xlo = max(0, (int)floor(((double)a.px[p] - w) * sc) - 1);
    const int xhi = min(side - 1, (int)ceil(((double)a.px[p] + w) * sc) + 1);
    const int ylo = max(0, (int)floor(((double)a.py[p] - w) * sc) - 1);
    const int yhi = min(side - 1, (int)ceil(((double)a.py[p] + w) * sc) + 1);
    const int bw = xhi - xlo + 1;
    const int cells = bw * (yhi - ylo + 1);
    for (int c0 = 0; c0 < cells; c0 += kWave) {
      const int c = c0 + lane;
      int q = -1;
      if (c < cells) {
        const int k = (ylo + c / bw) * side + xlo + c % bw;
        if (k < cnt) {
          const int qq = b0 + k;
          if (qq != p && dist2d(a, qq, p) <= w) q = qq;
        }
      }
      const uint64_t m = ballot(q >= 0);
      emit(q, rank_below(m), __popcll(m));
    }
  }
}

__global__ __launch_bounds__(kWave) void k_cand_count(CandArgs a) {
  const int p = a.freed[blockIdx.x];
  int total = 0;
  scan_window(a, p, [&](int, int, int cnt) { total += cnt; });
  if (lane_id() == 0) a.count[blockIdx.x] = total;
}

__global__ __launch_bounds__(kWave) void k_cand_features(CandArgs a) {
  __shared__ int s_q[kCandCap];
  const int u = blockIdx.x, p = a.freed[u], lane = lane_id();
  int total = 0;
  scan_window(a, p, [&](int q, int r, int cnt) {
    if (q >= 0 && total + r < kCandCap) s_q[total + r] = q;
    total += cnt;
  });
  wave_lds_sync();
  const int c = min(total, kCandCap);
  double mx = 0.0;
  for (int k = lane; k < c; k += kWave) mx = fmax(mx, a.fire[s_q[k]]);
  for (int d = kWave / 2; d > 0; d >>= 1) mx = fmax(mx, __shfl_xor(mx, d, kWave));
  const int base = a.offs[u];
  if (lane == 0) a.wmax[u] = mx;
  const double rp = a.rate[p];
  for (int k = lane; k < c; k += kWave) {
    const int q = s_q[k];
    const double fq = a.fire[q];
    int before = 0;  // window members ordered before q: more firing, or as many and lower id
    for (int j = 0; j < c; ++j) {
      const int o = s_q[j];
      const double fo = a.fire[o];
      before += (fo > fq || (fo == fq && o < q)) ? 1 : 0;
    }
    const double rank = c > 1 ? (double)before / (double)(c - 1) : 0.0;
    const double rq = a.rate[q];
    float f[16];
    f[0] = (float)rq;
    f[1] = (float)rp;
    f[2] = (float)(rq * rp * 100.0);
    f[3] = (float)dist2d(a, q, p);
    f[4] = a.f4;
    f[5] = a.f5[p];
    f[6] = a.f6[q];
    f[7] = a.f7[p];
    f[8] = a.f8[q];
    f[9] = a.f9[p];
    f[10] = a.f10[q];
    f[11] = a.f10[p];
    f[12] = mx > 0 ? (float)(fq / mx) : 0.0f;
    f[13] = (float)rank;
    f[14] = a.f14[p];
    f[15] = 1.0f;
    // the scorer rounds every feature to bf16 (round to nearest even) anyway
    uint32_t h[8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
      h[i] = (uint32_t)bf16_rne(f[2 * i]) | ((uint32_t)bf16_rne(f[2 * i + 1]) << 16);
    uint4* dst = reinterpret_cast<uint4*>(a.feat + (size_t)(base + k) * 16);
    dst[0] = make_uint4(h[0], h[1], h[2], h[3]);
    dst[1] = make_uint4(h[4], h[5], h[6], h[7]);
    a.rank[base + k] = f[13];
    a.cand[base + k] = q;
  }
}

}  // namespace

int conn_gpu_prune_select(const float* w, const int32_t* age, int64_t E, int protect, int K,
                          uint8_t* mask) {
  std::memset(mask, 0, (size_t)E);
  if (K <= 0 || E <= 0) return SNC_OK;
  hip_clear();
  DevBuf<float> dw;
  DevBuf<int32_t> dage;
  DevBuf<uint64_t> dkey;
  DevBuf<uint32_t> dhist;
  DevBuf<uint8_t> dmask;
  int rc = dw.upload(w, (size_t)E);
  if (!rc) rc = dage.upload(age, (size_t)E);
  if (!rc) rc = dkey.alloc((size_t)E);
  if (!rc) rc = dhist.alloc(256);
  if (!rc) rc = dmask.alloc((size_t)E);
  if (rc) return rc;
  const int blocks = (int)std::min<int64_t>((E + 255) / 256, 1024);
  hipLaunchKernelGGL(k_prune_keys, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, nullptr, dw.p,
                     dage.p, E, protect, dkey.p);
  rc = hip_last("prune key launch");
  uint64_t prefix = 0;
  int64_t need = K;  // keys still to take at or below the prefix
  for (int shift = 56; shift >= 0 && !rc; shift -= 8) {
    uint32_t h[256];
    hipError_t e = hipMemset(dhist.p, 0, 256 * sizeof(uint32_t));
    if (e != hipSuccess) {
      rc = hip_fail(e, "prune histogram clear");
      break;
    }
    hipLaunchKernelGGL(k_prune_hist, dim3(blocks), dim3(256), 0, nullptr, dkey.p, E, prefix,
                       shift, dhist.p);
    rc = hip_last("prune histogram launch");
    if (rc) break;
    e = hipMemcpy(h, dhist.p, sizeof(h), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      rc = hip_fail(e, "prune histogram download");
      break;
    }
    int digit = 0;
    for (; digit < 256; ++digit) {
      if ((int64_t)h[digit] >= need) break;
      need -= h[digit];
    }
    if (digit == 256) return fail(SNC_E_INVALID, "prune selection: fewer eligible synapses than K");
    prefix |= (uint64_t)digit << shift;
  }
  if (rc) return rc;
  // prefix is now the K-th smallest key; keys are distinct below ~0.
  hipLaunchKernelGGL(k_prune_mark, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, nullptr,
                     dkey.p, E, prefix, dmask.p);
  rc = hip_last("prune mark launch");
  if (rc) return rc;
  const hipError_t e = hipMemcpy(mask, dmask.p, (size_t)E, hipMemcpyDeviceToHost);
  return e == hipSuccess ? SNC_OK : hip_fail(e, "prune mask download");
}

int conn_gpu_candidates(const ConnSheetView& v, const std::vector<int>& freed, float window,
                        snc_hip_mlp* model, std::vector<int>& offs, std::vector<int>& all_q,
                        std::vector<float>& rank, std::vector<double>& wmax,
                        std::vector<float>& score, double* score_ms) {
  const int nf = (int)freed.size();
  offs.assign((size_t)nf + 1, 0);
  all_q.clear();
  rank.clear();
  wmax.assign((size_t)nf, 0.0);
  score.clear();
  if (score_ms) *score_ms = 0.0;
  if (nf == 0) return SNC_OK;
  hip_clear();
  DevBuf<int32_t> d_layer, d_lb, d_lc, d_side, d_freed, d_offs, d_count, d_cand;
  DevBuf<float> d_px, d_py, d_f5, d_f6, d_f7, d_f8, d_f9, d_f10, d_f14, d_rank, d_score;
  DevBuf<uint16_t> d_feat;
  DevBuf<double> d_fire, d_rate, d_wmax;
  const size_t n = (size_t)v.n, L = (size_t)v.num_layers;
  int rc = d_layer.upload(v.layer.data(), n);
  if (!rc) rc = d_lb.upload(v.layer_begin.data(), L);
  if (!rc) rc = d_lc.upload(v.layer_count.data(), L);
  if (!rc) rc = d_side.upload(v.side.data(), L);
  if (!rc) rc = d_freed.upload(freed.data(), (size_t)nf);
  if (!rc) rc = d_count.alloc((size_t)nf);
  if (!rc) rc = d_px.upload(v.px.data(), n);
  if (!rc) rc = d_py.upload(v.py.data(), n);
  if (!rc) rc = d_fire.upload(v.fire.data(), n);
  if (!rc) rc = d_rate.upload(v.rate.data(), n);
  if (!rc) rc = d_f5.upload(v.f5.data(), n);
  if (!rc) rc = d_f6.upload(v.f6.data(), n);
  if (!rc) rc = d_f7.upload(v.f7.data(), n);
  if (!rc) rc = d_f8.upload(v.f8.data(), n);
  if (!rc) rc = d_f9.upload(v.f9.data(), n);
  if (!rc) rc = d_f10.upload(v.f10.data(), n);
  if (!rc) rc = d_f14.upload(v.f14.data(), n);
  if (rc) return rc;
  for (int p : freed)
    if (p < 0 || p >= v.n) return fail(SNC_E_INVALID, "freed post out of range");
  CandArgs a{};
  a.n = v.n;
  a.num_layers = v.num_layers;
  a.window = window;
  a.layer = d_layer.p;
  a.layer_begin = d_lb.p;
  a.layer_count = d_lc.p;
  a.side = d_side.p;
  a.px = d_px.p;
  a.py = d_py.p;
  a.fire = d_fire.p;
  a.rate = d_rate.p;
  a.f5 = d_f5.p;
  a.f6 = d_f6.p;
  a.f7 = d_f7.p;
  a.f8 = d_f8.p;
  a.f9 = d_f9.p;
  a.f10 = d_f10.p;
  a.f14 = d_f14.p;
  a.f4 = v.f4;
  a.freed = d_freed.p;
  a.count = d_count.p;
  hipLaunchKernelGGL(k_cand_count, dim3(nf), dim3(kWave), 0, nullptr, a);
  rc = hip_last("candidate count launch");
  if (rc) return rc;
  std::vector<int32_t> cnt((size_t)nf);
  hipError_t e = hipMemcpy(cnt.data(), d_count.p, cnt.size() * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "candidate count download");
  for (int u = 0; u < nf; ++u) {
    if (cnt[u] > kCandCap) return fail(SNC_E_UNSUPPORTED, "grow window holds more than 4096 candidates");
    offs[u + 1] = offs[u] + cnt[u];
  }
  const int64_t M = offs[nf];
  if (M == 0) return SNC_OK;
  rc = d_offs.upload(offs.data(), offs.size());
  if (!rc) rc = d_cand.alloc((size_t)M);
  if (!rc) rc = d_feat.alloc((size_t)M * 16);
  if (!rc) rc = d_rank.alloc((size_t)M);
  if (!rc) rc = d_wmax.alloc((size_t)nf);
  if (!rc) rc = d_score.alloc((size_t)M);
  if (rc) return rc;
  a.offs = d_offs.p;
  a.cand = d_cand.p;
  a.feat = d_feat.p;
  a.rank = d_rank.p;
  a.wmax = d_wmax.p;
  hipLaunchKernelGGL(k_cand_features, dim3(nf), dim3(kWave), 0, nullptr, a);
  rc = hip_last("candidate feature launch");
  if (rc) return rc;
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "candidate features");
  rc = mlp_score_device(model, d_feat.p, M, d_score.p, score_ms);
  if (rc) return rc;
  all_q.resize((size_t)M);
  rank.resize((size_t)M);
  score.resize((size_t)M);
  e = hipMemcpy(all_q.data(), d_cand.p, (size_t)M * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(rank.data(), d_rank.p, (size_t)M * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(wmax.data(), d_wmax.p, (size_t)nf * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(score.data(), d_score.p, (size_t)M * 4, hipMemcpyDeviceToHost);
  return e == hipSuccess ? SNC_OK : hip_fail(e, "candidate download");
}

}  // namespace snc_amd
