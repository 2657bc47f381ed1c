// hip-pull kernels and the halo-indexed column build (see pull.hpp).

#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "fold.hpp"
#include "hip_err.hpp"
#include "pull.hpp"
#include "wave.hpp"

namespace snc_amd {

namespace {

constexpr int kHubRounds = kPullHubChunk / kWave;
constexpr int kHubThreads = 256;  // chunk waves per hub-chunk block: 4
constexpr int kBatch = 4;         // light targets whose loads go out together
constexpr int kHaloRound = 8;     // halo words a thread loads per round
constexpr int kHubGroup = 4;      // hub chunks whose list loads go out together
constexpr int kHubPre = 8;        // list values per chunk loaded with its count
static_assert(kPullHubChunk % kWave == 0, "hub chunks are whole rounds of a wave");
static_assert(kPullWaveTargets % kBatch == 0, "whole batches per wave");

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

template <class T>
int dupload(T** p, const std::vector<T>& h) {
  *p = nullptr;
  if (h.empty()) return SNC_OK;
  const size_t bytes = h.size() * sizeof(T);
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    hip_clear();
    return fail(SNC_E_NOMEM, "device allocation failed (pull columns)");
  }
  account_alloc(bytes);
  e = hipMemcpy(*p, h.data(), bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? SNC_OK : hip_fail(e, "pull column upload");
}

template <class T>
int dload(std::vector<T>& h, const T* d, size_t count) {
  h.resize(count);
  if (!count) return SNC_OK;
  const hipError_t e = hipMemcpy(h.data(), d, count * sizeof(T), hipMemcpyDeviceToHost);
  return e == hipSuccess ? SNC_OK : hip_fail(e, "pull column readback");
}

// Loads and stores at a 32-bit element offset from a base pointer: the
// offset stays one VGPR and the base an SGPR pair (the saddr form), which
// saves the 64-bit address arithmetic per lane. Every hip-pull array it
// indexes holds less than 4 GiB (checked when the plan is built).
template <class T>
__device__ __forceinline__ T ld32(const T* base, uint32_t i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                     (size_t)(i * (uint32_t)sizeof(T)));
}
template <class T>
__device__ __forceinline__ void st32(T* base, uint32_t i, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (size_t)(i * (uint32_t)sizeof(T))) = v;
}

// The fired word of a direct-form entry ((pre << 8) | delay) for arrival
// step t (tm = t % L; 1 <= d < L, so the origin slot is (tm - d) mod L).
template <class Word>
__device__ __forceinline__ Word direct_word(const PullArgs& a, uint32_t pk, int tm) {
  int s = tm - (int)(pk & 0xFFu);
  if (s < 0) s += a.L;
  return ld32(static_cast<const Word*>(a.F), (uint32_t)s * (uint32_t)a.n + ((pk >> 8) & 0x7FFFFFu));
}

// A light entry's fired word: its LDS halo slot, or (DIRECT kernels only)
// F itself. Lanes past the column's end read the halo's zero slot, which
// follows the block's words. o is the entry's place in the light stream.
template <class Word, bool DIRECT>
__device__ __forceinline__ Word entry_word(const PullArgs& a, const Word* s_halo, bool in,
                                           uint32_t h, uint32_t o, int tm) {
  if constexpr (DIRECT) {
    Word w = s_halo[in && h != kHaloDirect ? h : (uint32_t)a.halo];
    if (in && h == kHaloDirect) w = direct_word<Word>(a, ld32(a.ldir, o), tm);
    return w;
  } else {
    return s_halo[in ? h : (uint32_t)a.halo];
  }
}

// Adds, in entry order, the values (lane k holds entry k's) of the entries
// set in m (bit i: entry first + i) to acc. Every lane stays active and lanes
// with nothing left add +0, which leaves a sum started from +0 as it is.
__device__ __forceinline__ float walk_half(float acc, uint32_t m, int first, float w) {
  while (ballot(m != 0)) {
    const int k = m ? first + __builtin_ctz(m) : 0;
    const float vk = __int_as_float(__builtin_amdgcn_ds_bpermute(k << 2, __float_as_int(w)));
    acc = __fadd_rn(acc, m ? vk : 0.0f);
    m &= m - 1;
  }
  return acc;
}

// Adds one round of 64 column entries (lane k holds entry k's fired word and
// value) to every sample's sum, in column order. The fired bits are
// transposed through the wave's 64 LDS masks (s_mask, all zero on entry and
// on return; sample b's mask is two 32-bit words, entries 0-31 and 32-63):
// entry lane k ors its bit into the mask of every sample its word has, then
// sample lane b walks its mask in entry order and adds those entries'
// values (+0 on the rounds where it has none left). The work follows the set
// bits, not the entries with any bit set, and every bit loop runs on 32-bit
// halves (one ffbl and one and per bit, no 64-bit emulation).
template <class Word>
__device__ __forceinline__ unsigned long long transpose_round(Word word,
                                                              unsigned long long* s_mask) {
  const int lane = lane_id();
  uint32_t* m32 = reinterpret_cast<uint32_t*>(s_mask);
  const uint32_t me = 1u << (lane & 31);
  uint32_t* mine = m32 + (lane >> 5);  // this entry's half of every sample's mask
  for (uint32_t x = (uint32_t)word; x; x &= x - 1) atomicOr(mine + 2 * __builtin_ctz(x), me);
  if constexpr (sizeof(Word) == 8)
    for (uint32_t x = (uint32_t)(word >> 32); x; x &= x - 1)
      atomicOr(mine + 2 * (32 + __builtin_ctz(x)), me);
  wave_lds_sync();
  const unsigned long long both = s_mask[lane];
  s_mask[lane] = 0ull;
  return both;
}

template <class Word>
__device__ __forceinline__ float add_round(Word word, float w, float acc,
                                           unsigned long long* s_mask) {
  if (ballot(word != (Word)0) == 0) return acc;
  const unsigned long long both = transpose_round<Word>(word, s_mask);
  acc = walk_half(acc, (uint32_t)both, 0, w);  // entries 0-31, then 32-63: column order
  return walk_half(acc, (uint32_t)(both >> 32), 32, w);
}

// As walk_half, but appends the values to this lane's list (row pos of lst,
// lane's column) instead of adding them; returns the new count.
__device__ __forceinline__ int list_half(int pos, uint32_t m, int first, float w, float* lst,
                                         int lstride) {
  const int lane = lane_id();
  while (ballot(m != 0)) {
    const int k = m ? first + __builtin_ctz(m) : 0;
    const float vk = __int_as_float(__builtin_amdgcn_ds_bpermute(k << 2, __float_as_int(w)));
    if (m) {
      lst[(size_t)pos * lstride + lane] = vk;
      ++pos;
    }
    m &= m - 1;
  }
  return pos;
}

// One round of 64 hub entries: every sample's arriving values, in column
// order, appended to its list (lane b: sample b, pos its count so far).
template <class Word>
__device__ __forceinline__ int list_round(Word word, float w, int pos, unsigned long long* s_mask,
                                          float* lst, int lstride) {
  if (ballot(word != (Word)0) == 0) return pos;
  const unsigned long long both = transpose_round<Word>(word, s_mask);
  pos = list_half(pos, (uint32_t)both, 0, w, lst, lstride);
  return list_half(pos, (uint32_t)(both >> 32), 32, w, lst, lstride);
}

// LDS of a step block: the halo's fired words, every wave's 64 sample masks,
// then the injection pool, its per-entry links and per-target list heads.
__host__ __device__ inline size_t halo_bytes(int word_bytes, int halo) {
  return ((size_t)(halo + 1) * word_bytes + 15) / 16 * 16;
}
__host__ __device__ inline size_t mask_bytes(int blk) {
  return (size_t)(blk / kPullWaveTargets) * kWave * 8;
}
__host__ __device__ inline size_t lds_bytes(int word_bytes, int halo, int pool, int blk) {
  return halo_bytes(word_bytes, halo) + mask_bytes(blk) + (size_t)pool * 12 + (size_t)blk * 4 + 16;
}

// Hub chunks of step t: wave c lists chunk c's arrivals per sample, in
// column order, and their counts.
template <class Word>
__global__ __launch_bounds__(kHubThreads) void k_pull_hub_chunks(PullArgs a, int t) {
  __shared__ unsigned long long s_masks[kHubThreads];
  const int lane = lane_id();
  const int c = blockIdx.x * (kHubThreads / kWave) + (threadIdx.x >> 6);
  unsigned long long* s_mask = s_masks + (threadIdx.x & ~(kWave - 1));
  s_mask[lane] = 0ull;
  wave_lds_sync();
  if (c >= a.n_chunks) return;
  const int tm = t % a.L;
  const int p0 = a.chunk_pos[c], len = a.chunk_len[c];
  uint32_t pk[kHubRounds];
  float w[kHubRounds];
  Word wd[kHubRounds];
#pragma unroll
  for (int u = 0; u < kHubRounds; ++u) {
    const int k = u * kWave + lane;
    const uint32_t o = (uint32_t)(k < len ? p0 + k : p0);  // in the chunk either way
    pk[u] = ld32(a.hub_pk, o);
    w[u] = ld32(a.hub_w, o);
    if (k >= len) pk[u] = kHaloNone;
  }
  asm volatile("" ::: "memory");  // the value loads are issued here, not where first used
  // the fired summary bit of each entry's (origin step, pre) first: most are
  // clear, and a clear bit needs no word
  uint32_t sb[kHubRounds];
#pragma unroll
  for (int u = 0; u < kHubRounds; ++u) {
    int s = tm - (int)(pk[u] & 0xFFu);
    if (s < 0) s += a.L;
    const uint32_t pre = (pk[u] >> 8) & 0x7FFFFFu;
    const uint32_t q = pk[u] != kHaloNone ? (uint32_t)s * (uint32_t)a.sum_words + (pre >> 4) : 0u;
    sb[u] = (uint32_t)ld32(a.fsum, q) >> (pre & 15u);
  }
#pragma unroll
  for (int u = 0; u < kHubRounds; ++u)
    wd[u] = pk[u] != kHaloNone && (sb[u] & 1u) ? direct_word<Word>(a, pk[u], tm) : (Word)0;
  float* lst = a.hub_list + (size_t)p0 * a.lstride;  // the chunk's list rows
  int pos = 0;
#pragma unroll
  for (int u = 0; u < kHubRounds; ++u) pos = list_round<Word>(wd[u], w[u], pos, s_mask, lst, a.lstride);
  if (lane < a.lstride) a.hub_cnt[(size_t)c * a.lstride + lane] = (uint16_t)pos;
}

// A light target's first two rounds of entries and its state.
struct Tgt {
  uint32_t h0, h1;
  float w0, w1, v;
  int r, len, c0;
};

// One step t: the block's 16-target waves, in batches of kBatch targets
// whose loads go out together, then the hubs among each wave's targets
// (partials from the chunk launch). Eight waves per SIMD: two blocks share a
// CU, halos included.
template <class Word, bool DIRECT>
__global__ __launch_bounds__(kPullMaxThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void
k_pull_step(PullArgs a, int t) {
  extern __shared__ __align__(16) unsigned char s_dyn[];
  Word* s_halo = reinterpret_cast<Word*>(s_dyn);
  unsigned long long* s_masks =
      reinterpret_cast<unsigned long long*>(s_dyn + halo_bytes((int)sizeof(Word), a.halo));
  int2* s_pool = reinterpret_cast<int2*>(reinterpret_cast<unsigned char*>(s_masks) + mask_bytes(a.blk));
  int* s_next = reinterpret_cast<int*>(s_pool + a.pool);
  int* s_head = s_next + a.pool;  // [a.blk]
  int* s_pool_n = s_head + a.blk;

  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int nthr = (int)blockDim.x, waves = nthr >> 6;
  // Blocks are dealt to the 8 XCDs round-robin; give each XCD a contiguous
  // run of target blocks so the halos it stages overlap in its own L2, the
  // blocks with the most hub chunks to walk first (blk_order).
  const int R = a.n_blocks, x = blockIdx.x;
  const int q8 = R >> 3, r8 = R & 7, xcd = x & 7;
  const int blk = a.blk_order[xcd * q8 + min(xcd, r8) + (x >> 3)];
  const int base = blk * a.blk;
  const int tm = t % a.L;

  // Halo: the fired word of every (pre, delay) the block's light columns
  // use; each thread's loads go out together.
  {
    const int64_t h0 = a.halo_ptr[blk];
    const int nh = (int)(a.halo_ptr[blk + 1] - h0);
    for (int k0 = 0; k0 < nh; k0 += kHaloRound * nthr) {
      uint32_t pk[kHaloRound];
      Word wd[kHaloRound];
#pragma unroll
      for (int u = 0; u < kHaloRound; ++u) pk[u] = a.halo_pk[h0 + min(k0 + tid + u * nthr, nh - 1)];
#pragma unroll
      for (int u = 0; u < kHaloRound; ++u) wd[u] = direct_word<Word>(a, pk[u], tm);
#pragma unroll
      for (int u = 0; u < kHaloRound; ++u)
        if (k0 + tid + u * nthr < nh) s_halo[k0 + tid + u * nthr] = wd[u];
    }
  }
  unsigned long long* s_mask = s_masks + wave * kWave;
  s_mask[lane] = 0ull;
  if (tid == 0) s_halo[a.halo] = (Word)0;  // the zero slot
  if (tid < a.blk) s_head[tid] = -1;
  if (tid == 0) *s_pool_n = 0;
  __syncthreads();
  // Injections of step t of every sample into the pool, linked per target.
  // Wave w stages samples w, w + waves, ...: four samples' list bounds load
  // together, then their first 64 entries each.
  for (int b0 = wave; b0 < a.B; b0 += 4 * waves) {
    constexpr int kPer = 4;
    int k0[kPer], cnt[kPer], nid[kPer];
    float amp[kPer];
#pragma unroll
    for (int s = 0; s < kPer; ++s) {
      const int b = b0 + s * waves;
      k0[s] = 0;
      cnt[s] = 0;
      if (b < a.B) {
        const int32_t* bp =
            a.inj_blk_ptr + (size_t)b * a.inj_stride + (size_t)t * (a.n_blocks + 1) + blk;
        k0[s] = bp[0];
        cnt[s] = bp[1] - k0[s];
      }
    }
#pragma unroll
    for (int s = 0; s < kPer; ++s) {
      nid[s] = base;
      amp[s] = 0.0f;
      if (lane < cnt[s]) {
        nid[s] = a.blk_nid[k0[s] + lane];
        amp[s] = a.blk_amp[k0[s] + lane];
      }
    }
#pragma unroll
    for (int s = 0; s < kPer; ++s) {
      if (cnt[s] <= 0) continue;
      const int b = b0 + s * waves;
      int at = 0;
      if (lane == 0) at = atomicAdd(s_pool_n, cnt[s]);
      at = __shfl(at, 0, kWave);
      for (int q = lane; q < cnt[s]; q += kWave) {
        const int slot = at + q;
        if (slot >= a.pool) break;  // the plan build sizes the pool for every (step, block)
        const int id = q == lane ? nid[s] : a.blk_nid[k0[s] + q];
        const float am = q == lane ? amp[s] : a.blk_amp[k0[s] + q];
        const int loc = id - base;
        s_pool[slot] = make_int2(loc | (b << 16), __float_as_int(am));
        s_next[slot] = atomicExch(&s_head[loc], slot);
      }
    }
  }
  __syncthreads();

  // This wave's 16 targets: light ones sum their columns, hubs their chunk
  // partials.
  const int jb = base + wave * kPullWaveTargets;
  int my_c0 = 0, my_len = -1, my_hub = -1;
  if (lane < kPullWaveTargets && jb + lane < a.n) {
    my_c0 = a.loff[jb + lane];
    my_len = a.lcnt[jb + lane];
    my_hub = a.hub_of[jb + lane];
  }
  const bool smp = lane < a.B;
  const uint32_t lane_b = (uint32_t)min(lane, a.B - 1);
  const uint32_t uB = (uint32_t)a.B;
  Word* Fw = static_cast<Word*>(a.F);
  uint32_t rbits = 0, sbits = 0;
  for (int i0 = 0; i0 < kPullWaveTargets; i0 += kBatch) {
    // Every load of the batch before any is used (no branches: lanes past a
    // column's end read the next entries of the stream, which is padded past
    // its end), so they are in flight together.
    Tgt p[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int i = i0 + u;
      p[u].c0 = __builtin_amdgcn_readlane(my_c0, i);
      p[u].len = __builtin_amdgcn_readlane(my_len, i);
      const uint32_t o0 = (uint32_t)(p[u].c0 + lane);
      p[u].h0 = ld32(a.lh, o0);
      p[u].w0 = ld32(a.lw, o0);
      p[u].h1 = ld32(a.lh, o0 + kWave);
      p[u].w1 = ld32(a.lw, o0 + kWave);
      const uint32_t so = (uint32_t)min(jb + i, a.n - 1) * uB + lane_b;
      p[u].v = ld32(a.v, so);
      p[u].r = ld32(a.refr, so);
    }
    // The loads above are issued here, not where their values are first used
    // (inside branches the compiler would otherwise sink them into, each then
    // waiting out a full memory latency).
    asm volatile("" ::: "memory");
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int i = i0 + u;
      const int j = jb + i;
      const int len = p[u].len;
      if (len < 0) continue;  // a hub (stepped below) or past the last neuron
      const uint32_t o0 = (uint32_t)(p[u].c0 + lane);
      float acc = 0.0f, ext = 0.0f;
      // both rounds' fired words looked up before either is used
      const Word wd0 = entry_word<Word, DIRECT>(a, s_halo, lane < len, p[u].h0, o0, tm);
      const Word wd1 =
          entry_word<Word, DIRECT>(a, s_halo, lane + kWave < len, p[u].h1, o0 + kWave, tm);
      acc = add_round<Word>(wd0, p[u].w0, acc, s_mask);
      acc = add_round<Word>(wd1, p[u].w1, acc, s_mask);
      for (int k0 = 2 * kWave; k0 < len; k0 += kWave) {  // longer columns
        const uint32_t o = o0 + (uint32_t)k0;
        acc = add_round<Word>(
            entry_word<Word, DIRECT>(a, s_halo, k0 + lane < len, ld32(a.lh, o), o, tm),
            ld32(a.lw, o), acc, s_mask);
      }
      for (int q = s_head[j - base]; q >= 0; q = s_next[q]) {
        const int2 e = s_pool[q];
        if ((e.x >> 16) == lane) ext = __int_as_float(e.y);
      }
      bool f = false;
      if (smp) {
        const uint32_t so = (uint32_t)j * uB + (uint32_t)lane;
        float v = p[u].v;
        int r = p[u].r;
        f = lif_step(v, r, ext, acc, a.lif);
        st32(a.v, so, v);
        if (r != p[u].r) st32(a.refr, so, (uint8_t)r);
        // a no-return atomic: nothing waits for it (a load-add-store here
        // waited out a memory latency whenever any sample fired)
        if (f) atomicAdd(a.nspk + so, 1);
      }
      const uint64_t m = ballot(f);
      if (lane == 0) st32(Fw, (uint32_t)tm * (uint32_t)a.n + (uint32_t)j, (Word)m);
      rbits |= (f ? 1u : 0u) << i;
      sbits |= (m != 0 ? 1u : 0u) << i;
    }
  }
  // The wave's hubs: each one's chunk lists in chunk order, lane b adding
  // sample b's arrivals one by one (counts and first values of kHubGroup
  // chunks loaded together; the lists are padded past their end, and values
  // past a count are loaded but not added), then integrate/fire.
  for (uint64_t hm = ballot(my_hub >= 0); hm; hm &= hm - 1) {
    const int i = __builtin_ctzll(hm);
    const int hub = __builtin_amdgcn_readlane(my_hub, i);
    const int j = jb + i;
    const int c0 = a.hub_chunk_ptr[hub], c1 = a.hub_chunk_ptr[hub + 1];
    const uint32_t so = (uint32_t)j * uB + lane_b;
    float v = ld32(a.v, so);
    const int r0 = ld32(a.refr, so);
    const uint32_t ls = (uint32_t)a.lstride;
    float acc = 0.0f;
    for (int cc = c0; cc < c1; cc += kHubGroup) {
      int cnt[kHubGroup];
      uint32_t row[kHubGroup];
      float x[kHubGroup][kHubPre];
#pragma unroll
      for (int q = 0; q < kHubGroup; ++q) {
        const bool in = cc + q < c1;
        row[q] = in ? (uint32_t)a.chunk_pos[cc + q] : 0u;
        cnt[q] = ld32(a.hub_cnt, (uint32_t)(in ? cc + q : c0) * ls + (uint32_t)lane);
        if (!in || lane >= a.lstride) cnt[q] = 0;
#pragma unroll
        for (int k = 0; k < kHubPre; ++k)
          x[q][k] = ld32(a.hub_list, (row[q] + (uint32_t)k) * ls + (uint32_t)lane);
      }
#pragma unroll
      for (int q = 0; q < kHubGroup; ++q) {
#pragma unroll
        for (int k = 0; k < kHubPre; ++k)
          if (k < cnt[q]) acc = __fadd_rn(acc, x[q][k]);
        for (int k0 = kHubPre; k0 < cnt[q]; k0 += kHubPre) {  // longer lists, kHubPre at a time
          float y[kHubPre];
#pragma unroll
          for (int k = 0; k < kHubPre; ++k)
            y[k] = ld32(a.hub_list, (row[q] + (uint32_t)(k0 + k)) * ls + (uint32_t)lane);
#pragma unroll
          for (int k = 0; k < kHubPre; ++k)
            if (k0 + k < cnt[q]) acc = __fadd_rn(acc, y[k]);
        }
      }
    }
    bool f = false;
    if (smp) {
      const float ext = a.inj_hub[(size_t)lane * a.hinj_stride + (size_t)t * a.n_hub + hub];
      int r = r0;
      f = lif_step(v, r, ext, acc, a.lif);
      st32(a.v, so, v);
      if (r != r0) st32(a.refr, so, (uint8_t)r);
      if (f) atomicAdd(a.nspk + so, 1);
    }
    const uint64_t m = ballot(f);
    if (lane == 0) st32(Fw, (uint32_t)tm * (uint32_t)a.n + (uint32_t)j, (Word)m);
    rbits |= (f ? 1u : 0u) << i;
    sbits |= (m != 0 ? 1u : 0u) << i;
  }
  // The fired summary half-word of the wave's 16 targets.
  if (lane == 0 && jb < a.n)
    st32(a.fsum, (uint32_t)tm * (uint32_t)a.sum_words + (uint32_t)(jb >> 4), (uint16_t)sbits);
  // Sample b's raster bits of the wave's 16 targets (half a raster word, so
  // the word is or-ed).
  if (a.raster && smp && rbits && jb < a.n)
    atomicOr(a.raster + (size_t)lane * a.raster_stride + (size_t)t * a.words + (jb >> 5),
             rbits << (jb & 31));
}

template <class Word, bool DIRECT>
void set_step_lds() {
  (void)hipFuncSetAttribute((const void*)k_pull_step<Word, DIRECT>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPullMaxLds);
}

}  // namespace

void PullGraph::release() {
  dfree(lw);
  dfree(lh);
  dfree(ldir);
  dfree(loff);
  dfree(lcnt);
  dfree(hub_of);
  dfree(halo_ptr);
  dfree(halo);
  dfree(hub_ids);
  dfree(chunk_hub);
  dfree(hub_pk);
  dfree(hub_w);
  dfree(chunk_pos);
  dfree(chunk_len);
  dfree(hub_chunk_ptr);
  dfree(blk_order);
  hub_host.clear();
  cap = -1;
  n_blocks = n_hub = n_chunks = halo_words = 0;
  hub_entries = 0;
  direct_entries = 0;
  direct = false;
}

bool pull_block_ok(int blk) {
  return blk >= kPullMinBlock && blk <= kPullMaxBlock && blk % kPullWaveTargets == 0;
}

int pull_prepare(const DevGraph& g, int cap, int blk, PullGraph& pg) {
  if (pg.cap == cap && pg.blk == blk) return SNC_OK;
  if (!pull_block_ok(blk)) return fail(SNC_E_INVALID, "hip-pull block size");
  pg.release();
  const int n = g.n;
  if (n >= (1 << 23)) return fail(SNC_E_UNSUPPORTED, "hip-pull packs pre ids in 23 bits");
  if (g.L > 256) return fail(SNC_E_UNSUPPORTED, "hip-pull packs delays in 8 bits");
  // Host copies of the column layout: each position's (pre << 8) | delay.
  std::vector<int32_t> row_ptr, cpos, ccol;
  std::vector<uint8_t> cdel;
  int rc = dload(row_ptr, g.row_ptr, (size_t)n + 1);
  if (!rc) rc = dload(cpos, g.cpos, (size_t)g.S);
  if (!rc) rc = dload(ccol, g.ccol, (size_t)n + 1);
  if (!rc) rc = dload(cdel, g.cdel, (size_t)g.P);
  if (rc) return rc;
  std::vector<float> cvh;
  rc = dload(cvh, g.cval, (size_t)g.P);
  if (rc) return rc;
  std::vector<uint32_t> cpk((size_t)g.P, 0u);
  for (int i = 0; i < n; ++i)
    for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
      const int p = cpos[(size_t)e];
      cpk[(size_t)p] = ((uint32_t)i << 8) | cdel[(size_t)p];
    }

  const int R = std::max(1, (n + blk - 1) / blk);
  std::vector<int32_t> lcnt((size_t)n, 0), hub_of((size_t)n, -1), hubs, chunk_hub, chunk_pos,
      chunk_len;
  std::vector<int32_t> hub_chunk_ptr(1, 0);
  std::vector<uint32_t> hub_pk;
  std::vector<float> hub_w;
  for (int j = 0; j < n; ++j) {
    const int deg = g.in_degree[j];
    if (deg <= cap) {
      lcnt[j] = deg;
      continue;
    }
    lcnt[j] = -1;
    const int h = (int)hubs.size();
    hub_of[j] = h;
    hubs.push_back(j);
    // the hub's entries in column order: (delay descending, pre, CSR position)
    const int32_t e0 = (int32_t)hub_pk.size();
    for (int k = 0; k < deg; ++k) {
      hub_pk.push_back(cpk[(size_t)ccol[j] + k]);
      hub_w.push_back(cvh[(size_t)ccol[j] + k]);
    }
    for (int k = 0; k < deg; k += kPullHubChunk) {
      chunk_hub.push_back(h);
      chunk_pos.push_back(e0 + k);
      chunk_len.push_back(std::min(kPullHubChunk, deg - k));
    }
    hub_chunk_ptr.push_back((int32_t)chunk_hub.size());
  }
  // The light stream: every light target's entries in column order, back to
  // back. Per block: the distinct (delay, pre) of its light columns, sorted;
  // the first kPullHaloCap are staged, the rest read F directly.
  int64_t n_light = 0;
  std::vector<int32_t> loff((size_t)n, 0);
  for (int j = 0; j < n; ++j) {
    loff[j] = (int32_t)n_light;
    n_light += std::max(0, lcnt[j]);
  }
  if (n_light + kPullStreamPad >= (1ll << 30)) return fail(SNC_E_UNSUPPORTED, "hip-pull: too many light entries");
  std::vector<float> lw((size_t)(n_light + kPullStreamPad), 0.0f);
  std::vector<uint16_t> lh((size_t)(n_light + kPullStreamPad), kHaloDirect);
  std::vector<uint32_t> ldir;
  std::vector<int64_t> halo_ptr((size_t)R + 1, 0);
  std::vector<uint32_t> halo;
  std::vector<uint32_t> keys;
  int halo_max = 0;
  int64_t direct = 0;
  auto key_of = [](uint32_t pk) { return ((pk & 0xFFu) << 24) | (pk >> 8); };  // (delay, pre)
  for (int q = 0; q < R; ++q) {
    const int j0 = q * blk, j1 = std::min(n, j0 + blk);
    keys.clear();
    for (int j = j0; j < j1; ++j)
      for (int k = 0; k < lcnt[j]; ++k) keys.push_back(key_of(cpk[(size_t)ccol[j] + k]));
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const int keep = std::min<int>((int)keys.size(), kPullHaloCap);
    halo_max = std::max(halo_max, keep);
    for (int k = 0; k < keep; ++k) halo.push_back(((keys[k] & 0xFFFFFFu) << 8) | (keys[k] >> 24));
    halo_ptr[q + 1] = (int64_t)halo.size();
    for (int j = j0; j < j1; ++j)
      for (int k = 0; k < lcnt[j]; ++k) {
        const size_t p = (size_t)ccol[j] + k;
        const size_t o = (size_t)loff[j] + k;
        const int rank = (int)(std::lower_bound(keys.begin(), keys.end(), key_of(cpk[p])) - keys.begin());
        lw[o] = cvh[p];
        if (rank < keep) {
          lh[o] = (uint16_t)rank;
        } else {
          if (ldir.empty()) ldir.assign(lw.size(), 0u);
          lh[o] = kHaloDirect;
          ldir[o] = cpk[p];
          ++direct;
        }
      }
  }
  if (ldir.empty()) ldir.assign(1, 0u);
  // Launch order: each XCD's run of target blocks (as the step kernel deals
  // them), the blocks whose hubs have the most chunks first -- a hub's wave
  // walks its chunks' lists one after another, so those blocks run longest
  // and should not start in the launch's last round.
  std::vector<int32_t> blk_order((size_t)R), blk_chunks((size_t)R, 0);
  for (size_t h = 0; h < hubs.size(); ++h)
    blk_chunks[(size_t)(hubs[h] / blk)] += hub_chunk_ptr[h + 1] - hub_chunk_ptr[h];
  for (int xcd = 0, q8 = R >> 3, r8 = R & 7; xcd < 8; ++xcd) {
    const int b0 = xcd * q8 + std::min(xcd, r8), b1 = b0 + q8 + (xcd < r8 ? 1 : 0);
    for (int q = b0; q < b1; ++q) blk_order[(size_t)q] = q;
    std::stable_sort(blk_order.begin() + b0, blk_order.begin() + b1,
                     [&](int32_t u, int32_t v) { return blk_chunks[(size_t)u] > blk_chunks[(size_t)v]; });
  }
  pg.n_blocks = R;
  pg.n_hub = (int)hubs.size();
  pg.n_chunks = (int)chunk_hub.size();
  pg.hub_entries = (int64_t)hub_pk.size();
  pg.halo_words = std::max(1, halo_max);
  pg.direct_entries = direct;
  pg.direct = direct > 0;
  pg.hub_host = hubs;
  if (!rc) rc = dupload(&pg.lw, lw);
  if (!rc) rc = dupload(&pg.lh, lh);
  if (!rc) rc = dupload(&pg.ldir, ldir);
  if (!rc) rc = dupload(&pg.loff, loff);
  if (!rc) rc = dupload(&pg.hub_of, hub_of);
  if (!rc) rc = dupload(&pg.lcnt, lcnt);
  if (!rc) rc = dupload(&pg.halo_ptr, halo_ptr);
  if (!rc) rc = dupload(&pg.halo, halo.empty() ? std::vector<uint32_t>(1, 0u) : halo);
  if (!rc) rc = dupload(&pg.hub_ids, hubs);
  if (!rc) rc = dupload(&pg.chunk_hub, chunk_hub);
  if (!rc) rc = dupload(&pg.hub_pk, hub_pk);
  if (!rc) rc = dupload(&pg.hub_w, hub_w);
  if (!rc) rc = dupload(&pg.chunk_pos, chunk_pos);
  if (!rc) rc = dupload(&pg.chunk_len, chunk_len);
  if (!rc) rc = dupload(&pg.hub_chunk_ptr, hub_chunk_ptr);
  if (!rc) rc = dupload(&pg.blk_order, blk_order);
  if (rc) {
    pg.release();
    return rc;
  }
  pg.cap = cap;
  pg.blk = blk;
  return SNC_OK;
}

size_t pull_lds_bytes(int B, int halo, int pool, int blk) {
  return lds_bytes(pull_word_bytes(B), halo, pool, blk);
}

int pull_launch_step(const PullArgs& a, int t, hipStream_t st) {
  const bool wide = a.B > 32;
  if (a.n_chunks > 0) {
    const int hb = (a.n_chunks + kHubThreads / kWave - 1) / (kHubThreads / kWave);
    if (wide)
      hipLaunchKernelGGL(k_pull_hub_chunks<uint64_t>, dim3(hb), dim3(kHubThreads), 0, st, a, t);
    else
      hipLaunchKernelGGL(k_pull_hub_chunks<uint32_t>, dim3(hb), dim3(kHubThreads), 0, st, a, t);
    const int rc = hip_last("pull hub chunk launch failed");
    if (rc) return rc;
  }
  const size_t lds = pull_lds_bytes(a.B, a.halo, a.pool, a.blk);
  const dim3 grid(a.n_blocks), block(a.blk / kPullWaveTargets * kWave);
  if (wide && a.direct)
    hipLaunchKernelGGL((k_pull_step<uint64_t, true>), grid, block, lds, st, a, t);
  else if (wide)
    hipLaunchKernelGGL((k_pull_step<uint64_t, false>), grid, block, lds, st, a, t);
  else if (a.direct)
    hipLaunchKernelGGL((k_pull_step<uint32_t, true>), grid, block, lds, st, a, t);
  else
    hipLaunchKernelGGL((k_pull_step<uint32_t, false>), grid, block, lds, st, a, t);
  return hip_last("pull step launch failed");
}

int pull_set_lds(int B, int halo, int pool, int blk) {
  const size_t bytes = pull_lds_bytes(B, halo, pool, blk);
  if (bytes > kPullMaxLds) return fail(SNC_E_UNSUPPORTED, "hip-pull: halo and injection pool exceed LDS");
  // The ceiling, not this plan's size: plans of other sizes share the kernels.
  set_step_lds<uint32_t, false>();
  set_step_lds<uint32_t, true>();
  set_step_lds<uint64_t, false>();
  set_step_lds<uint64_t, true>();
  return hip_last("hip-pull LDS request");
}

void pull_free(PullGraph* pg) { delete pg; }

}  // namespace snc_amd
