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

struct MlpDev {
  uint16_t* w1b;    // [64][16] bf16
  uint16_t* w2b;    // [64][64] bf16
  uint16_t* w3b;    // [64] bf16
  const float* b1;  // into the FP32 master vector
  const float* b2;
  const float* b3;
};

int mlp_launch_score(const MlpDev& p, const float* X, int64_t M, float* out, int blocks,
                     hipStream_t st);
int mlp_launch_grad(const MlpDev& p, const float* X, const float* y, int64_t M, float* G,
                    float* loss, float* grad, float* loss_out, hipStream_t st);
int mlp_launch_adam(float* w, float* m, float* v, const float* g, float lr, int step,
                    const MlpDev& p, hipStream_t st);
int mlp_launch_quantize(const float* w, const MlpDev& p, hipStream_t st);
uint16_t mlp_bf16(float f);

}  // namespace snc_amd

struct snc_hip_mlp;

namespace snc_amd {

// Scores M candidates whose features are already on the device (X [M][16])
// into out [M], on the model's stream, and waits for them.
int mlp_score_device(snc_hip_mlp* m, const float* X, int64_t M, float* out, double* kernel_ms);

}  // namespace snc_amd
