// Learned grow proposal: a 16 -> 64 -> 64 -> 1 MLP scoring candidate
// synapses for the structural grow step (the hook replaces the candidate
// choice of Connectome::structural_update, src/structure/connectome.cpp:164-211).
//
// Model (BF16 weights, FP32 accumulation):
//   x  = bf16(features)                         16
//   h1 = relu(W1 x + b1),  h1b = bf16(h1)        64
//   h2 = relu(W2 h1b + b2)                       64
//   s  = W3 h2 + b3                              logit of P(synapse survives)
// Parameters are kept as an FP32 master copy (trained with Adam on BCE loss)
// and quantised to BF16 (round to nearest even) for both passes.
//
// Inference runs on the matrix cores: every wave scores tiles of 16
// candidates with v_mfma_f32_16x16x16 bf16 in the transposed form
// H^T = W X^T -- the weight block is operand A, the activation tile operand B --
// so the accumulator layout of one layer (lane l: rows 4*(l/16)+i, column
// l%16) is exactly the B-operand layout of the next layer. Training runs one
// example per wave on the VALU (lane j = hidden unit j) and writes
// per-example gradients example-major, reduced per parameter in example
// order (deterministic).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "hip_err.hpp"
#include "mlp.hpp"
#include "wave.hpp"

namespace snc_amd {

typedef short bf16x4 __attribute__((ext_vector_type(4)));

constexpr int kScoreDepth = 4;  // candidate tiles whose loads a wave keeps in flight
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float f) { return bf16_rne(f); }
__host__ __device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
__device__ __forceinline__ float q(float f) { return bf16_to_f32(f32_to_bf16(f)); }

// ---------------------------------------------------------------------------
// inference (MFMA)
// ---------------------------------------------------------------------------

// X holds the candidates' features as bf16 bits, [M][16]: the layer-1 MFMA
// takes its B operand in bf16, so 32 bytes per candidate is all it reads.
__global__ __launch_bounds__(256) void k_mlp_score(const MlpDev p, const uint16_t* __restrict__ X,
                                                   int64_t M, float* __restrict__ out) {
  const int lane = lane_id();
  const int r = lane & 15;          // A row / B,D column within a 16-block
  const int kq = (lane >> 4) * 4;   // k quad of the lane (A, B) / D row quad
  // Weight blocks as A operands, resident in registers for the whole launch.
  bf16x4 a1[4], a2[4][4];
#pragma unroll
  for (int ob = 0; ob < 4; ++ob)
#pragma unroll
    for (int i = 0; i < 4; ++i) a1[ob][i] = (short)p.w1b[(ob * 16 + r) * kMlpIn + kq + i];
#pragma unroll
  for (int ob = 0; ob < 4; ++ob)
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        a2[ob][kb][i] = (short)p.w2b[(ob * 16 + r) * kMlpH + kb * 16 + kq + i];
  float b1v[4][4], b2v[4][4], w3v[4][4];
#pragma unroll
  for (int ob = 0; ob < 4; ++ob)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      b1v[ob][i] = p.b1[ob * 16 + kq + i];
      b2v[ob][i] = p.b2[ob * 16 + kq + i];
      w3v[ob][i] = bf16_to_f32(p.w3b[ob * 16 + kq + i]);
    }
  const float b3 = p.b3[0];

  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t tiles = (M + 15) / 16;
  // Each wave takes kScoreDepth tiles per round and issues all their feature
  // loads before the first MFMA, so one memory latency covers the round.
  for (int64_t base = wave; base < tiles; base += kScoreDepth * nwaves) {
    uint2 xs[kScoreDepth];  // features kq..kq+3 of candidate r: four bf16
#pragma unroll
    for (int u = 0; u < kScoreDepth; ++u) {
      const int64_t cand = (base + u * nwaves) * 16 + r;
      xs[u] = cand < M ? *reinterpret_cast<const uint2*>(X + cand * kMlpIn + kq) : make_uint2(0u, 0u);
    }
#pragma unroll
    for (int u = 0; u < kScoreDepth; ++u) {
      const int64_t tile = base + u * nwaves;
      if (tile >= tiles) break;  // uniform across the wave
      const int64_t cand = tile * 16 + r;
      // B operand of layer 1: X^T, lane holds features kq..kq+3 of candidate r.
      bf16x4 bx;
      bx[0] = (short)(xs[u].x & 0xffffu);
      bx[1] = (short)(xs[u].x >> 16);
      bx[2] = (short)(xs[u].y & 0xffffu);
      bx[3] = (short)(xs[u].y >> 16);
      bf16x4 h1[4];
#pragma unroll
      for (int ob = 0; ob < 4; ++ob) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1[ob], bx, c, 0, 0, 0);
        // ReLU outputs are finite and >= 0, where the hardware's packed
        // conversion (v_cvt_pk_bf16_f32, round to nearest even) equals
        // f32_to_bf16; it converts two values per instruction.
#pragma unroll
        for (int i = 0; i < 4; ++i)
          h1[ob][i] = __builtin_bit_cast(short, (__bf16)fmaxf(c[i] + b1v[ob][i], 0.0f));
      }
      float part = 0.0f;
#pragma unroll
      for (int ob = 0; ob < 4; ++ob) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[ob][kb], h1[kb], c, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) part += w3v[ob][i] * fmaxf(c[i] + b2v[ob][i], 0.0f);
      }
      // Lanes r, r+16, r+32, r+48 hold the four row quads of candidate r.
      part += __shfl_xor(part, 16, kWave);
      part += __shfl_xor(part, 32, kWave);
      if (lane < 16 && cand < M) out[cand] = part + b3;
    }
  }
}

// ---------------------------------------------------------------------------
// training (VALU, one example per lane)
// ---------------------------------------------------------------------------

// Forward + backward of one example per wave (lane j owns hidden unit j of
// both layers); writes the example's gradient to G[e][param] and its loss to
// loss[e]. BF16 rounding is treated as identity in the backward pass
// (straight-through).
__global__ __launch_bounds__(256) void k_mlp_example_grad(const MlpDev p,
                                                          const float* __restrict__ X,
                                                          const float* __restrict__ y, int64_t M,
                                                          float* __restrict__ G,
                                                          float* __restrict__ loss) {
  constexpr int kWaves = 256 / kWave;
  __shared__ float s_w1[kMlpH * kMlpIn], s_w2[kMlpH * kMlpH], s_w3[kMlpH], s_b1[kMlpH],
      s_b2[kMlpH];
  __shared__ float s_h1[kWaves][kMlpH], s_d2[kWaves][kMlpH];
  for (int k = threadIdx.x; k < kMlpH * kMlpIn; k += blockDim.x) s_w1[k] = bf16_to_f32(p.w1b[k]);
  for (int k = threadIdx.x; k < kMlpH * kMlpH; k += blockDim.x) s_w2[k] = bf16_to_f32(p.w2b[k]);
  for (int k = threadIdx.x; k < kMlpH; k += blockDim.x) {
    s_w3[k] = bf16_to_f32(p.w3b[k]);
    s_b1[k] = p.b1[k];
    s_b2[k] = p.b2[k];
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, j = lane_id();
  const float b3 = p.b3[0];
  for (int64_t e = (int64_t)blockIdx.x * kWaves + wave; e < M; e += (int64_t)gridDim.x * kWaves) {
    float x[kMlpIn];
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) x[k] = q(X[e * kMlpIn + k]);
    float s1 = s_b1[j];
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) s1 += s_w1[j * kMlpIn + k] * x[k];
    const float h1 = q(fmaxf(s1, 0.0f));
    s_h1[wave][j] = h1;
    wave_lds_sync();
    float s2 = s_b2[j];
#pragma unroll 16
    for (int k = 0; k < kMlpH; ++k) s2 += s_w2[j * kMlpH + k] * s_h1[wave][k];
    const float h2 = fmaxf(s2, 0.0f);
    float out = s_w3[j] * h2;  // fixed-order wave sum
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) out += __shfl_xor(out, off, kWave);
    out += b3;
    // mean binary cross-entropy with logits
    const float t = y[e];
    const float sg = 1.0f / (1.0f + expf(-out));
    const float dout = (sg - t) / (float)M;
    float* g = G + e * kMlpParams;
    if (j == 0) {
      loss[e] = fmaxf(out, 0.0f) - out * t + log1pf(expf(-fabsf(out)));
      g[kMlpOffB3] = dout;
    }
    g[kMlpOffW3 + j] = dout * h2;
    const float d2 = h2 > 0.0f ? dout * s_w3[j] : 0.0f;
    g[kMlpOffB2 + j] = d2;
    s_d2[wave][j] = d2;
    wave_lds_sync();
    // dW2[r][c] = d2[r] * h1[c]: lane j writes column j of every row (coalesced rows)
#pragma unroll 8
    for (int r = 0; r < kMlpH; ++r) g[kMlpOffW2 + r * kMlpH + j] = s_d2[wave][r] * h1;
    float dh1 = 0.0f;
#pragma unroll 16
    for (int r = 0; r < kMlpH; ++r) dh1 += s_d2[wave][r] * s_w2[r * kMlpH + j];
    const float d1 = h1 > 0.0f ? dh1 : 0.0f;
    g[kMlpOffB1 + j] = d1;
#pragma unroll
    for (int f = 0; f < kMlpIn; ++f) g[kMlpOffW1 + j * kMlpIn + f] = d1 * x[f];
    wave_lds_sync();  // s_h1 / s_d2 are rewritten by the next example
  }
}

// grad[p] = sum_e G[e][p] (examples in order); loss_out = sum_e loss[e] / M.
__global__ void k_mlp_reduce(const float* __restrict__ G, const float* __restrict__ loss,
                             int64_t M, float* __restrict__ grad, float* __restrict__ loss_out) {
  const int pidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx < kMlpParams) {
    float s = 0.0f;
    for (int64_t e = 0; e < M; ++e) s += G[e * kMlpParams + pidx];
    grad[pidx] = s;
  }
  if (pidx == kMlpParams) {
    double s = 0.0;
    for (int64_t e = 0; e < M; ++e) s += loss[e];
    loss_out[0] = (float)(s / (double)M);
  }
}

// Adam on the FP32 master parameters, then BF16 copies of the weights.
__global__ void k_mlp_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float lr, float b1, float b2, float eps,
                           float c1, float c2, MlpDev p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kMlpParams) return;
  m[i] = b1 * m[i] + (1.0f - b1) * g[i];
  v[i] = b2 * v[i] + (1.0f - b2) * g[i] * g[i];
  w[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
  if (i < kMlpOffB1) p.w1b[i] = f32_to_bf16(w[i]);
  else if (i >= kMlpOffW2 && i < kMlpOffB2) p.w2b[i - kMlpOffW2] = f32_to_bf16(w[i]);
  else if (i >= kMlpOffW3 && i < kMlpOffB3) p.w3b[i - kMlpOffW3] = f32_to_bf16(w[i]);
}

__global__ void k_mlp_quantize(const float* __restrict__ w, MlpDev p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kMlpParams) return;
  if (i < kMlpOffB1) p.w1b[i] = f32_to_bf16(w[i]);
  else if (i >= kMlpOffW2 && i < kMlpOffB2) p.w2b[i - kMlpOffW2] = f32_to_bf16(w[i]);
  else if (i >= kMlpOffW3 && i < kMlpOffB3) p.w3b[i - kMlpOffW3] = f32_to_bf16(w[i]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

int mlp_launch_score(const MlpDev& p, const uint16_t* X, int64_t M, float* out, int blocks,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_score, dim3(blocks), dim3(256), 0, st, p, X, M, out);
  return hip_last("MLP score launch failed");
}

int mlp_launch_grad(const MlpDev& p, const float* X, const float* y, int64_t M, float* G,
                    float* loss, float* grad, float* loss_out, hipStream_t st) {
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((M + 3) / 4, 4096));
  hipLaunchKernelGGL(k_mlp_example_grad, dim3(blocks), dim3(256), 0, st, p, X, y, M, G, loss);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, st, G, loss,
                     M, grad, loss_out);
  return hip_last("MLP grad launch failed");
}

int mlp_launch_adam(float* w, float* m, float* v, const float* g, float lr, int step,
                    const MlpDev& p, hipStream_t st) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, w, m, v, g, lr,
                     b1, b2, eps, c1, c2, p);
  return hip_last("MLP update launch failed");
}

int mlp_launch_quantize(const float* w, const MlpDev& p, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_quantize, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, w, p);
  return hip_last("MLP quantize launch failed");
}

uint16_t mlp_bf16(float f) { return f32_to_bf16(f); }

void mlp_features_bf16(const float* x, int64_t count, uint16_t* out) {
  for (int64_t k = 0; k < count; ++k) out[k] = f32_to_bf16(x[k]);
}

}  // namespace snc_amd
