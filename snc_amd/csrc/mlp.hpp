// Learned grow proposal MLP (16 -> 64 -> 64 -> 1): layout and launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace snc_amd {

constexpr int kMlpIn = 16;
constexpr int kMlpH = 64;
// Parameter vector layout (FP32 master copy): W1[64][16], b1[64], W2[64][64],
// b2[64], W3[64], b3.
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpH * kMlpIn;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpH;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpH * kMlpH;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpH;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpH;
constexpr int kMlpParams = kMlpOffB3 + 1;

// bf16 bits of a float, rounded to nearest even (a NaN stays a quiet NaN):
// the precision of every feature and weight the MLP's MFMAs take.
__host__ __device__ __forceinline__ uint16_t bf16_rne(float f) {
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

struct MlpDev {
  uint16_t* w1b;    // [64][16] bf16
  uint16_t* w2b;    // [64][64] bf16
  uint16_t* w3b;    // [64] bf16
  const float* b1;  // into the FP32 master vector
  const float* b2;
  const float* b3;
};

// X: features as bf16 bits [M][16] (mlp_features_bf16).
int mlp_launch_score(const MlpDev& p, const uint16_t* X, int64_t M, float* out, int blocks,
                     hipStream_t st);
int mlp_launch_grad(const MlpDev& p, const float* X, const float* y, int64_t M, float* G,
                    float* loss, float* grad, float* loss_out, hipStream_t st);
int mlp_launch_adam(float* w, float* m, float* v, const float* g, float lr, int step,
                    const MlpDev& p, hipStream_t st);
int mlp_launch_quantize(const float* w, const MlpDev& p, hipStream_t st);
uint16_t mlp_bf16(float f);
// Round-to-nearest-even bf16 bits of count floats (what the scorer and the
// gradient kernels round every feature to).
void mlp_features_bf16(const float* x, int64_t count, uint16_t* out);

}  // namespace snc_amd

struct snc_hip_mlp;

namespace snc_amd {

// Scores M candidates whose features are already on the device (X [M][16]
// bf16 bits) into out [M], on the model's stream, and waits for them.
int mlp_score_device(snc_hip_mlp* m, const uint16_t* X, int64_t M, float* out, double* kernel_ms);

}  // namespace snc_amd
