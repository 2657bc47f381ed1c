// hip-pull kernels and the sliced-column build (see pull.hpp).

#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "fold.hpp"
#include "hip_err.hpp"
#include "pull.hpp"
#include "wave.hpp"

namespace snc_amd {

namespace {

constexpr int kPullWalk = 16;      // column entries a lane has in flight per round
constexpr int kPullLifGroup = 8;   // samples whose state loads together
constexpr int kPullHubPerThread = kPullHubChunk / kPullBlock;
static_assert(kPullHubChunk % kPullBlock == 0, "hub chunks split evenly over a block");

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

template <class T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (!count) return SNC_OK;
  const hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    hip_clear();
    return fail(SNC_E_NOMEM, "device allocation failed (pull columns)");
  }
  account_alloc(count * sizeof(T));
  return SNC_OK;
}

template <class T>
int dupload(T** p, const std::vector<T>& h) {
  int rc = dalloc(p, h.size());
  if (!rc && !h.empty()) {
    const hipError_t e = hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_fail(e, "pull column upload");
  }
  return rc;
}

// One thread per pre row: every synapse goes to its target's slice (light)
// or its hub column, at its rank k in the target's column order (cpos is the
// padded column position, ccol the column start).
__global__ void k_pull_build(int n, const int32_t* __restrict__ row_ptr,
                             const int32_t* __restrict__ post, const float* __restrict__ wsig,
                             const uint8_t* __restrict__ delay, const int32_t* __restrict__ cpos,
                             const int32_t* __restrict__ ccol, const int32_t* __restrict__ hub_of,
                             const int64_t* __restrict__ hub_off,
                             const int64_t* __restrict__ slice_ptr, uint32_t* __restrict__ slab,
                             float* __restrict__ slab_w, uint32_t* __restrict__ hub_e,
                             float* __restrict__ hub_w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
    const int j = post[e];
    const int64_t k = cpos[e] - ccol[j];
    const uint32_t packed = ((uint32_t)i << 8) | (uint32_t)delay[e];
    const int h = hub_of[j];
    if (h >= 0) {
      hub_e[hub_off[h] + k] = packed;
      hub_w[hub_off[h] + k] = wsig[e];
    } else {
      const int64_t q = slice_ptr[j >> 6] + k * 64 + (j & 63);
      slab[q] = packed;
      slab_w[q] = wsig[e];
    }
  }
}

// F word of the synapse's pre at its origin step t - d, for arrival step t
// (tm = t % L; 1 <= d < L, so the slot is (tm - d) mod L).
__device__ __forceinline__ uint32_t fired_word(const PullArgs& a, uint32_t packed, int tm) {
  int s = tm - (int)(packed & 0xFFu);
  if (s < 0) s += a.L;
  return a.F[(size_t)s * a.n + (packed >> 8)] & a.sample_mask;
}

// LDS of one block: per-thread accumulator rows of `pitch` floats (B or B+1,
// odd so neighbouring rows start in different banks), then the light block's
// injection pool, per-target list heads and per-sample counters.
__host__ __device__ inline int acc_pitch(int B) { return B | 1; }
__host__ __device__ inline size_t pull_lds_bytes(int B, int pool) {
  return (size_t)kPullBlock * acc_pitch(B) * 4 + (size_t)pool * 12 + (size_t)kPullBlock * 4 +
         (size_t)kPullMaxSamples * 16 + 16;
}

// Light block: 4 waves, wave w steps slice 4 * block + w, lane = target.
__device__ void pull_light(const PullArgs& a, int t, int blk, unsigned char* lds) {
  const int pitch = acc_pitch(a.B);
  float* s_acc = reinterpret_cast<float*>(lds);  // [kPullBlock][pitch]
  int2* s_pool = reinterpret_cast<int2*>(lds + (size_t)kPullBlock * pitch * 4);  // (local | b << 16, amp)
  int* s_next = reinterpret_cast<int*>(s_pool + a.pool);
  int* s_head = s_next + a.pool;  // [kPullBlock]
  unsigned long long* s_spk = reinterpret_cast<unsigned long long*>(s_head + kPullBlock);
  unsigned long long* s_ev = s_spk + kPullMaxSamples;
  int* s_pool_n = reinterpret_cast<int*>(s_ev + kPullMaxSamples);

  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int j = blk * kPullBlock + tid;
  const int base = blk * kPullBlock;
  const bool mine = j < a.n && a.is_hub[j] == 0;
  float* acc = s_acc + tid * pitch;
  for (int b = 0; b < a.B; ++b) acc[b] = 0.0f;
  s_head[tid] = -1;
  if (tid < kPullMaxSamples) {
    s_spk[tid] = 0;
    s_ev[tid] = 0;
  }
  if (tid == 0) *s_pool_n = 0;
  __syncthreads();

  // Injections of step t of every sample into the LDS pool, linked per target.
  for (int b = wave; b < a.B; b += kPullBlock / kWave) {
    const int32_t* bp = a.inj_blk_ptr + (size_t)b * a.inj_stride + (size_t)t * (a.n_light_blocks + 1) + blk;
    const int k0 = bp[0], cnt = bp[1] - k0;
    if (cnt <= 0) continue;
    int at = 0;
    if (lane == 0) at = atomicAdd(s_pool_n, cnt);
    at = __shfl(at, 0, kWave);
    for (int q = lane; q < cnt; q += kWave) {
      const int slot = at + q;
      if (slot >= a.pool) break;  // the plan build sizes the pool for every (step, block)
      const int loc = a.blk_nid[k0 + q] - base;
      s_pool[slot] = make_int2(loc | (b << 16), __float_as_int(a.blk_amp[k0 + q]));
      s_next[slot] = atomicExch(&s_head[loc], slot);
    }
  }

  // The column walk, kPullWalk entries per round: the next round's entries
  // and values load while this round's fired words are gathered, so a round
  // costs about one memory round trip. Additions stay in column order.
  const int s = blk * (kPullBlock / kWave) + wave;
  const int tm = t % a.L;
  if (mine) {
    const int len = a.slice_len[s];
    const uint32_t* col = a.slab + a.slice_ptr[s] + lane;
    const float* colw = a.slab_w + a.slice_ptr[s] + lane;
    uint32_t pk[kPullWalk];
    float w[kPullWalk];
#pragma unroll
    for (int u = 0; u < kPullWalk; ++u) {
      const bool in = u < len;
      pk[u] = in ? col[(size_t)u * 64] : 0u;
      w[u] = in ? colw[(size_t)u * 64] : 0.0f;
    }
    for (int k0 = 0; k0 < len; k0 += kPullWalk) {
      uint32_t wd[kPullWalk], pn[kPullWalk];
      float wn[kPullWalk];
#pragma unroll
      for (int u = 0; u < kPullWalk; ++u) wd[u] = pk[u] & 0xFFu ? fired_word(a, pk[u], tm) : 0u;
#pragma unroll
      for (int u = 0; u < kPullWalk; ++u) {
        const int k = k0 + kPullWalk + u;
        const bool in = k < len;
        pn[u] = in ? col[(size_t)k * 64] : 0u;
        wn[u] = in ? colw[(size_t)k * 64] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < kPullWalk; ++u) {
        uint32_t m = wd[u];
        while (m) {
          const int b = __builtin_ctz(m);
          m &= m - 1u;
          acc[b] = __fadd_rn(acc[b], w[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < kPullWalk; ++u) {
        pk[u] = pn[u];
        w[u] = wn[u];
      }
    }
  }
  __syncthreads();

  // Integrate/fire every sample of this target, kPullLifGroup samples per
  // round of loads.
  uint32_t word = 0;
  if (mine) {
    uint32_t inj = 0;
    for (int q = s_head[tid]; q >= 0; q = s_next[q]) inj |= 1u << (s_pool[q].x >> 16);
    const int deg = a.row_ptr[j + 1] - a.row_ptr[j];
    for (int b0 = 0; b0 < a.B; b0 += kPullLifGroup) {
      float vv[kPullLifGroup];
      int rr[kPullLifGroup];
#pragma unroll
      for (int u = 0; u < kPullLifGroup; ++u) {
        const int b = b0 + u;
        vv[u] = b < a.B ? a.v[(size_t)b * a.n + j] : 0.0f;
        rr[u] = b < a.B ? a.refr[(size_t)b * a.n + j] : 0;
      }
#pragma unroll
      for (int u = 0; u < kPullLifGroup; ++u) {
        const int b = b0 + u;
        if (b >= a.B) break;
        const size_t sj = (size_t)b * a.n + j;
        float ext = 0.0f;
        if ((inj >> b) & 1u)
          for (int q = s_head[tid]; q >= 0; q = s_next[q])
            if ((s_pool[q].x >> 16) == b) ext = __int_as_float(s_pool[q].y);
        const bool f = lif_step(vv[u], rr[u], ext, acc[b], a.lif);
        a.v[sj] = vv[u];
        a.refr[sj] = rr[u];
        if (f) {
          word |= 1u << b;
          atomicAdd(a.nspk + sj, 1);
          atomicAdd(&s_spk[b], 1ull);
          atomicAdd(&s_ev[b], (unsigned long long)deg);
        }
      }
    }
    a.F[(size_t)tm * a.n + j] = word;
  }
  if (a.raster) {  // sample b's raster words of this wave's 64 targets
    for (int b = 0; b < a.B; ++b) {
      const uint64_t m = ballot((word >> b) & 1u);
      if (lane == 0) {
        uint32_t* row = a.raster + (size_t)b * a.raster_stride + (size_t)t * a.words;
        const int w0 = (base + wave * kWave) >> 5;
        if (w0 < a.words) row[w0] = (uint32_t)m;
        if (w0 + 1 < a.words) row[w0 + 1] = (uint32_t)(m >> 32);
      }
    }
  }
  __syncthreads();
  if (tid < a.B && s_spk[tid]) {
    atomicAdd(a.blk_spikes + (size_t)tid * a.blk_stride + blk, s_spk[tid]);
    atomicAdd(a.blk_events + (size_t)tid * a.blk_stride + blk, s_ev[tid]);
  }
}

// Hub chunk block: per-sample partial sums of one chunk of a hub column.
// Thread q takes entries q, q + 256, ... (at most kPullHubPerThread), loads
// them and their fired words in two rounds, sums them into its own LDS row;
// rows are then added in a fixed tree order.
__device__ void pull_hub_chunk(const PullArgs& a, int t, int c, unsigned char* lds) {
  const int pitch = acc_pitch(a.B);
  float* s_acc = reinterpret_cast<float*>(lds);
  const int tid = threadIdx.x;
  float* acc = s_acc + tid * pitch;
  for (int b = 0; b < a.B; ++b) acc[b] = 0.0f;
  const int64_t e0 = a.chunk_beg[c], e1 = a.chunk_beg[c + 1];
  const int tm = t % a.L;
  uint32_t pk[kPullHubPerThread], wd[kPullHubPerThread];
  float w[kPullHubPerThread];
#pragma unroll
  for (int u = 0; u < kPullHubPerThread; ++u) {
    const int64_t e = e0 + tid + (int64_t)u * kPullBlock;
    const bool in = e < e1;
    pk[u] = in ? a.hub_e[e] : 0u;
    w[u] = in ? a.hub_w[e] : 0.0f;
  }
#pragma unroll
  for (int u = 0; u < kPullHubPerThread; ++u) wd[u] = pk[u] & 0xFFu ? fired_word(a, pk[u], tm) : 0u;
#pragma unroll
  for (int u = 0; u < kPullHubPerThread; ++u) {
    uint32_t m = wd[u];
    while (m) {
      const int b = __builtin_ctz(m);
      m &= m - 1u;
      acc[b] = __fadd_rn(acc[b], w[u]);
    }
  }
  __syncthreads();
  for (int half = kPullBlock / 2; half > 0; half >>= 1) {
    if (tid < half) {
      const float* other = s_acc + (tid + half) * pitch;
      for (int b = 0; b < a.B; ++b) acc[b] = __fadd_rn(acc[b], other[b]);
    }
    __syncthreads();
  }
  if (tid < a.B) a.hub_part[(size_t)c * kPullMaxSamples + tid] = s_acc[tid];
}

__global__ __launch_bounds__(kPullBlock) void k_pull_step(PullArgs a, int t) {
  extern __shared__ __align__(16) unsigned char s_dyn[];
  if ((int)blockIdx.x < a.n_chunks)  // hub chunks first
    pull_hub_chunk(a, t, blockIdx.x, s_dyn);
  else
    pull_light(a, t, (int)blockIdx.x - a.n_chunks, s_dyn);
}

// One wave per hub, lane b = sample b: the hub's partials in chunk order,
// then integrate/fire.
__global__ __launch_bounds__(kWave) void k_pull_hub_step(PullArgs a, int t) {
  const int hidx = blockIdx.x, b = lane_id();
  const int h = a.hub_ids[hidx];
  bool f = false;
  if (b < a.B) {
    float slot = 0.0f;
    for (int c = a.hub_chunk_ptr[hidx]; c < a.hub_chunk_ptr[hidx + 1]; ++c)
      slot = __fadd_rn(slot, a.hub_part[(size_t)c * kPullMaxSamples + b]);
    const size_t sj = (size_t)b * a.n + h;
    float v = a.v[sj];
    int r = a.refr[sj];
    const float ext = a.inj_hub[(size_t)b * a.hinj_stride + (size_t)t * a.n_hub + hidx];
    f = lif_step(v, r, ext, slot, a.lif);
    a.v[sj] = v;
    a.refr[sj] = r;
    if (f) {
      atomicAdd(a.nspk + sj, 1);
      const size_t slot_id = (size_t)b * a.blk_stride + a.n_light_blocks + hidx;
      atomicAdd(a.blk_spikes + slot_id, 1ull);
      atomicAdd(a.blk_events + slot_id, (unsigned long long)(a.row_ptr[h + 1] - a.row_ptr[h]));
      if (a.raster)
        atomicOr(a.raster + (size_t)b * a.raster_stride + (size_t)t * a.words + (h >> 5),
                 1u << (h & 31));
    }
  }
  const uint64_t m = ballot(f);
  if (b == 0) a.F[(size_t)(t % a.L) * a.n + h] = (uint32_t)m;
}

}  // namespace

void PullGraph::release() {
  dfree(slab);
  dfree(slab_w);
  dfree(slice_ptr);
  dfree(slice_len);
  dfree(hub_ids);
  dfree(hub_e);
  dfree(hub_w);
  dfree(chunk_hub);
  dfree(chunk_beg);
  dfree(hub_chunk_ptr);
  dfree(is_hub);
  hub_host.clear();
  cap = -1;
  n_slices = n_hub = n_chunks = 0;
  entries = 0;
}

int pull_prepare(const DevGraph& g, int cap, PullGraph& pg) {
  if (pg.cap == cap) return SNC_OK;
  pg.release();
  const int n = g.n;
  if (n >= (1 << 24)) return fail(SNC_E_UNSUPPORTED, "hip-pull packs pre ids in 24 bits");
  std::vector<uint8_t> is_hub((size_t)n, 0);
  std::vector<int32_t> hub_of((size_t)n, -1), hubs;
  for (int j = 0; j < n; ++j)
    if (g.in_degree[j] > cap) {
      hub_of[j] = (int32_t)hubs.size();
      hubs.push_back(j);
      is_hub[j] = 1;
    }
  const int n_slices = (n + 63) / 64;
  std::vector<int64_t> slice_ptr((size_t)n_slices, 0);
  std::vector<int32_t> slice_len((size_t)n_slices, 0);
  int64_t entries = 0;
  for (int s = 0; s < n_slices; ++s) {
    int len = 0;
    for (int j = s * 64; j < std::min(n, s * 64 + 64); ++j)
      if (!is_hub[j]) len = std::max(len, g.in_degree[j]);
    slice_ptr[s] = entries;
    slice_len[s] = len;
    entries += (int64_t)len * 64;
  }
  std::vector<int64_t> hub_off(hubs.size() + 1, 0);
  std::vector<int32_t> chunk_hub, hub_chunk_ptr(hubs.size() + 1, 0);
  std::vector<int64_t> chunk_beg;
  for (size_t h = 0; h < hubs.size(); ++h) {
    const int64_t deg = g.in_degree[hubs[h]];
    hub_off[h + 1] = hub_off[h] + deg;
    for (int64_t q = 0; q < deg; q += kPullHubChunk) {
      chunk_hub.push_back((int32_t)h);
      chunk_beg.push_back(hub_off[h] + q);
    }
    hub_chunk_ptr[h + 1] = (int32_t)chunk_hub.size();
  }
  chunk_beg.push_back(hub_off[hubs.size()]);
  pg.n_slices = n_slices;
  pg.n_hub = (int)hubs.size();
  pg.n_chunks = (int)chunk_hub.size();
  pg.entries = entries;
  pg.hub_host = hubs;
  int rc = SNC_OK;
  int32_t* d_hub_of = nullptr;
  int64_t* d_hub_off = nullptr;
  if (!rc) rc = dalloc(&pg.slab, (size_t)std::max<int64_t>(entries, 1));
  if (!rc) rc = dalloc(&pg.slab_w, (size_t)std::max<int64_t>(entries, 1));
  if (!rc) rc = dupload(&pg.slice_ptr, slice_ptr);
  if (!rc) rc = dupload(&pg.slice_len, slice_len);
  if (!rc) rc = dupload(&pg.hub_ids, hubs);
  if (!rc) rc = dalloc(&pg.hub_e, (size_t)std::max<int64_t>(hub_off.back(), 1));
  if (!rc) rc = dalloc(&pg.hub_w, (size_t)std::max<int64_t>(hub_off.back(), 1));
  if (!rc) rc = dupload(&pg.chunk_hub, chunk_hub);
  if (!rc) rc = dupload(&pg.chunk_beg, chunk_beg);
  if (!rc) rc = dupload(&pg.hub_chunk_ptr, hub_chunk_ptr);
  if (!rc) rc = dupload(&pg.is_hub, is_hub);
  if (!rc) rc = dupload(&d_hub_of, hub_of);
  if (!rc) rc = dupload(&d_hub_off, hub_off);
  if (!rc) {
    hipError_t e = hipMemset(pg.slab, 0, (size_t)std::max<int64_t>(entries, 1) * 4);
    if (e == hipSuccess) e = hipMemset(pg.slab_w, 0, (size_t)std::max<int64_t>(entries, 1) * 4);
    if (e != hipSuccess) rc = hip_fail(e, "pull column clear");
  }
  if (!rc && n > 0) {
    hipLaunchKernelGGL(k_pull_build, dim3((n + 255) / 256), dim3(256), 0, nullptr, n, g.row_ptr,
                       g.post, g.wsig, g.delay, g.cpos, g.ccol, d_hub_of, d_hub_off,
                       pg.slice_ptr, pg.slab, pg.slab_w, pg.hub_e, pg.hub_w);
    rc = hip_last("pull column build launch");
    if (!rc) {
      const hipError_t e = hipDeviceSynchronize();
      if (e != hipSuccess) rc = hip_fail(e, "pull column build");
    }
  }
  dfree(d_hub_of);
  dfree(d_hub_off);
  if (rc) {
    pg.release();
    return rc;
  }
  pg.cap = cap;
  return SNC_OK;
}

int pull_launch_step(const PullArgs& a, int t, hipStream_t st) {
  hipLaunchKernelGGL(k_pull_step, dim3(a.n_chunks + a.n_light_blocks), dim3(kPullBlock),
                     (unsigned)pull_lds_bytes(a.B, a.pool), st, a, t);
  int rc = hip_last("pull step launch failed");
  if (!rc && a.n_hub > 0) {
    hipLaunchKernelGGL(k_pull_hub_step, dim3(a.n_hub), dim3(kWave), 0, st, a, t);
    rc = hip_last("pull hub step launch failed");
  }
  return rc;
}

int pull_set_lds(int B, int pool) {
  const size_t bytes = pull_lds_bytes(B, pool);
  if (bytes > kPullMaxLds) return fail(SNC_E_UNSUPPORTED, "hip-pull: injection pool exceeds LDS");
  // The ceiling, not this plan's size: plans with other pools share the kernel.
  const hipError_t e = hipFuncSetAttribute((const void*)k_pull_step,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)kPullMaxLds);
  return e == hipSuccess ? SNC_OK : hip_fail(e, "hip-pull LDS request");
}

void pull_free(PullGraph* pg) { delete pg; }

}  // namespace snc_amd
