// C ABI of the learned grow proposal MLP. Host code; kernels in mlp.hip.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "common.hpp"
#include "hip_err.hpp"
#include "mlp.hpp"
#include "rng.hpp"
#include "snc_hip.h"

using namespace snc_amd;

static_assert(kMlpParams == SNC_MLP_PARAMS, "parameter count mismatch");
static_assert(kMlpIn == SNC_MLP_IN, "feature count mismatch");

struct snc_hip_mlp {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float* scratch = nullptr;  // this model's device's copy of the fixed scratch array
  float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // FP32 master, Adam state, grad
  uint16_t *w1b = nullptr, *w2b = nullptr, *w3b = nullptr;
  float *X = nullptr, *y = nullptr, *out = nullptr, *G = nullptr, *loss = nullptr, *loss1 = nullptr;
  uint16_t* Xb = nullptr;  // scoring input: features as bf16 bits [cap_xb][16]
  int64_t cap_x = 0, cap_g = 0, cap_xb = 0;
  int adam_step = 0;
  MlpDev dev() const { return {w1b, w2b, w3b, w + kMlpOffB1, w + kMlpOffB2, w + kMlpOffB3}; }
  ~snc_hip_mlp() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
    for (void* p : {(void*)w, (void*)m, (void*)v, (void*)g, (void*)w1b, (void*)w2b, (void*)w3b,
                    (void*)X, (void*)y, (void*)out, (void*)G, (void*)loss, (void*)loss1,
                    (void*)Xb})
      if (p) (void)hipFree(p);
  }
};

namespace {

// Fixed scratch of snc_hip_mlp_grad_noalloc: what snc_hip_mlp_grad keeps in
// buffers it allocates -- the staged examples X [kNoAllocMax][16] and labels,
// the per-example losses and the per-example gradients G [kNoAllocMax][params]
// -- in one file-scope device array. Every offset is a multiple of 64 floats,
// so each part starts 256-byte aligned.
constexpr int64_t kNoAllocMax = SNC_MLP_NOALLOC_MAX_EXAMPLES;
constexpr int64_t kScrX = 0;
constexpr int64_t kScrY = kScrX + kNoAllocMax * kMlpIn;
constexpr int64_t kScrLoss = kScrY + kNoAllocMax;
constexpr int64_t kScrG = kScrLoss + kNoAllocMax;
constexpr int64_t kScrFloats = kScrG + kNoAllocMax * kMlpParams;
static_assert(kNoAllocMax % 64 == 0, "scratch parts must stay 256-byte aligned");

}  // namespace

static __device__ __align__(256) float scratch[kScrFloats];

namespace {

// The device address of `scratch` on each device, resolved once per device
// with that device current. In host code the array's name is only a host-side
// stand-in; kernels and copies get these addresses.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_addr[kMaxDevices] = {};

int scratch_address(float** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "scratch: no current device");
  if (dev < 0 || dev >= kMaxDevices) return fail(SNC_E_UNSUPPORTED, "scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "scratch: hipGetSymbolAddress");
    if (!p) return fail(SNC_E_HIP, "scratch: hipGetSymbolAddress returned null");
    g_scratch_addr[dev] = (float*)p;
  }
  *out = g_scratch_addr[dev];
  return SNC_OK;
}

int grow(void** p, int64_t& cap, int64_t need, size_t elem) {
  if (need <= cap) return SNC_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (hipMalloc(p, (size_t)need * elem) != hipSuccess) {
    cap = 0;
    return fail(SNC_E_NOMEM, "MLP buffer allocation failed");
  }
  cap = need;
  return SNC_OK;
}

int ensure_x(snc_hip_mlp* m, int64_t M) {
  if (M <= m->cap_x) return SNC_OK;
  int64_t c0 = m->cap_x, c1 = m->cap_x, c2 = m->cap_x, c3 = m->cap_x;
  int rc = grow((void**)&m->X, c0, M * kMlpIn, sizeof(float));
  if (!rc) rc = grow((void**)&m->y, c1, M, sizeof(float));
  if (!rc) rc = grow((void**)&m->out, c2, M, sizeof(float));
  if (!rc) rc = grow((void**)&m->loss, c3, M, sizeof(float));
  if (!rc) m->cap_x = M;
  return rc;
}

// Largest example chunk of one gradient pass (bounds the G buffer).
constexpr int64_t kGradChunk = 8192;

int grad_device(snc_hip_mlp* m, int64_t off, int64_t M, std::vector<float>& grad_acc,
                double& loss_acc) {
  // G is [chunk][params]; chunks of at most kGradChunk examples, combined on
  // the host in chunk order weighted by chunk size.
  int rc = grow((void**)&m->G, m->cap_g, (int64_t)kMlpParams * std::min(M, kGradChunk),
                sizeof(float));
  if (rc) return rc;
  grad_acc.assign(kMlpParams, 0.0f);
  loss_acc = 0.0;
  std::vector<float> part(kMlpParams);
  for (int64_t c = 0; c < M; c += kGradChunk) {
    const int64_t mc = std::min(kGradChunk, M - c);
    rc = mlp_launch_grad(m->dev(), m->X + (off + c) * kMlpIn, m->y + off + c, mc, m->G, m->loss,
                         m->g, m->loss1, m->stream);
    if (rc) return rc;
    float l = 0.0f;
    if (hipMemcpyAsync(part.data(), m->g, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost,
                       m->stream) != hipSuccess ||
        hipMemcpyAsync(&l, m->loss1, sizeof(float), hipMemcpyDeviceToHost, m->stream) !=
            hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
      return fail(SNC_E_HIP, "MLP gradient readback failed");
    const float wgt = (float)mc / (float)M;  // chunk means -> overall mean
    for (int k = 0; k < kMlpParams; ++k) grad_acc[k] += part[k] * wgt;
    loss_acc += (double)l * mc / (double)M;
  }
  return SNC_OK;
}

}  // namespace

namespace snc_amd {

int mlp_score_device(snc_hip_mlp* m, const uint16_t* X, int64_t M, float* out, double* kernel_ms) {
  if (M <= 0) return SNC_OK;
  int dev = 0, cus = 256;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    cus = prop.multiProcessorCount;
  const int64_t tiles = (M + 15) / 16;
  // 4 waves per block, several rounds of prefetched tiles per wave; about
  // as many blocks as stay resident (three per CU at the kernel's register use).
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((tiles + 15) / 16, (int64_t)cus * 3));
  (void)hipEventRecord(m->ev0, m->stream);
  const int rc = mlp_launch_score(m->dev(), X, M, out, blocks, m->stream);
  if (rc) return rc;
  (void)hipEventRecord(m->ev1, m->stream);
  const hipError_t e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) return hip_fail(e, "scoring");
  if (kernel_ms) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, m->ev0, m->ev1);
    *kernel_ms = ms;
  }
  return SNC_OK;
}

}  // namespace snc_amd

extern "C" {

int snc_hip_mlp_create(uint64_t seed, snc_hip_mlp** out) {
  if (!out) return fail(SNC_E_INVALID, "null argument");
  *out = nullptr;
  std::unique_ptr<snc_hip_mlp> m(new snc_hip_mlp());
  int rc = SNC_OK;
  auto al = [&](void** p, size_t bytes) {
    if (!rc && hipMalloc(p, bytes) != hipSuccess) rc = fail(SNC_E_NOMEM, "MLP allocation failed");
  };
  al((void**)&m->w, sizeof(float) * kMlpParams);
  al((void**)&m->m, sizeof(float) * kMlpParams);
  al((void**)&m->v, sizeof(float) * kMlpParams);
  al((void**)&m->g, sizeof(float) * kMlpParams);
  al((void**)&m->loss1, sizeof(float) * 4);
  al((void**)&m->w1b, 2 * kMlpH * kMlpIn);
  al((void**)&m->w2b, 2 * kMlpH * kMlpH);
  al((void**)&m->w3b, 2 * kMlpH);
  if (rc) return rc;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess)
    return fail(SNC_E_HIP, "stream/event creation failed");
  // The allocation-free gradient's scratch on this device (once per device).
  rc = scratch_address(&m->scratch);
  if (rc) return rc;
  // He-uniform: U(-sqrt(6 / fan_in), +sqrt(6 / fan_in)); zero biases.
  std::vector<float> p(kMlpParams, 0.0f);
  SplitMix64 r(seed);
  auto fill = [&](int off, int count, int fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int k = 0; k < count; ++k) p[off + k] = (float)((2.0 * r.uniform() - 1.0) * a);
  };
  fill(kMlpOffW1, kMlpH * kMlpIn, kMlpIn);
  fill(kMlpOffW2, kMlpH * kMlpH, kMlpH);
  fill(kMlpOffW3, kMlpH, kMlpH);
  *out = m.release();
  return snc_hip_mlp_set_params(*out, p.data());
}

void snc_hip_mlp_destroy(snc_hip_mlp* m) { delete m; }

int snc_hip_mlp_get_params(const snc_hip_mlp* m, float* params) {
  if (!m || !params) return fail(SNC_E_INVALID, "null argument");
  if (hipStreamSynchronize(m->stream) != hipSuccess ||
      hipMemcpy(params, m->w, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(SNC_E_HIP, "parameter readback failed");
  return SNC_OK;
}

int snc_hip_mlp_set_params(snc_hip_mlp* m, const float* params) {
  if (!m || !params) return fail(SNC_E_INVALID, "null argument");
  if (hipMemcpy(m->w, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice) != hipSuccess)
    return fail(SNC_E_HIP, "parameter upload failed");
  std::vector<float> z(kMlpParams, 0.0f);
  (void)hipMemcpy(m->m, z.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice);
  (void)hipMemcpy(m->v, z.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice);
  m->adam_step = 0;
  int rc = mlp_launch_quantize(m->w, m->dev(), m->stream);
  if (!rc && hipStreamSynchronize(m->stream) != hipSuccess) rc = fail(SNC_E_HIP, "quantize failed");
  return rc;
}

int snc_hip_mlp_score(snc_hip_mlp* m, const float* features, int64_t M, float* scores,
                      double* kernel_ms) {
  if (!m || (M > 0 && (!features || !scores)) || M < 0) return fail(SNC_E_INVALID, "bad arguments");
  if (M == 0) return SNC_OK;
  hip_clear();
  int rc = ensure_x(m, M);
  if (!rc) rc = grow((void**)&m->Xb, m->cap_xb, M * kMlpIn, sizeof(uint16_t));
  if (rc) return rc;
  // the features go up as the bf16 values the scorer computes with
  std::vector<uint16_t> xb((size_t)M * kMlpIn);
  mlp_features_bf16(features, M * kMlpIn, xb.data());
  if (hipMemcpyAsync(m->Xb, xb.data(), sizeof(uint16_t) * xb.size(), hipMemcpyHostToDevice,
                     m->stream) != hipSuccess)
    return fail(SNC_E_HIP, "feature upload failed");
  rc = mlp_score_device(m, m->Xb, M, m->out, kernel_ms);
  if (rc) return rc;
  if (hipMemcpy(scores, m->out, sizeof(float) * M, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(SNC_E_HIP, "score download failed");
  return SNC_OK;
}

int snc_hip_mlp_grad(snc_hip_mlp* m, const float* features, const float* labels, int64_t M,
                     float* grad, double* loss) {
  if (!m || !features || !labels || !grad || M < 1) return fail(SNC_E_INVALID, "bad arguments");
  hip_clear();
  int rc = ensure_x(m, M);
  if (rc) return rc;
  if (hipMemcpy(m->X, features, sizeof(float) * M * kMlpIn, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->y, labels, sizeof(float) * M, hipMemcpyHostToDevice) != hipSuccess)
    return fail(SNC_E_HIP, "example upload failed");
  std::vector<float> gacc;
  double l = 0.0;
  rc = grad_device(m, 0, M, gacc, l);
  if (rc) return rc;
  std::copy(gacc.begin(), gacc.end(), grad);
  if (loss) *loss = l;
  return SNC_OK;
}

int snc_hip_mlp_grad_noalloc(snc_hip_mlp* m, const float* features, const float* labels,
                             int64_t M, float* grad, double* loss) {
  if (!m || !features || !labels || !grad || M < 1) return fail(SNC_E_INVALID, "bad arguments");
  if (M > kNoAllocMax)
    return fail(SNC_E_INVALID, "snc_hip_mlp_grad_noalloc: M = " + std::to_string(M) +
                                   " exceeds SNC_MLP_NOALLOC_MAX_EXAMPLES (" +
                                   std::to_string(kNoAllocMax) + ")");
  if (!m->scratch) return fail(SNC_E_HIP, "snc_hip_mlp_grad_noalloc: scratch not resolved");
  hip_clear();
  // Every G row and loss entry of the M examples is written by the gradient
  // kernel before the reduction reads it, so nothing left by an earlier call
  // needs clearing.
  float* X = m->scratch + kScrX;
  float* y = m->scratch + kScrY;
  float* lossv = m->scratch + kScrLoss;
  float* G = m->scratch + kScrG;
  if (hipMemcpyAsync(X, features, sizeof(float) * M * kMlpIn, hipMemcpyHostToDevice, m->stream) !=
          hipSuccess ||
      hipMemcpyAsync(y, labels, sizeof(float) * M, hipMemcpyHostToDevice, m->stream) != hipSuccess)
    return fail(SNC_E_HIP, "example upload failed");
  int rc = mlp_launch_grad(m->dev(), X, y, M, G, lossv, m->g, m->loss1, m->stream);
  if (rc) return rc;
  float l = 0.0f;
  if (hipMemcpyAsync(grad, m->g, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m->stream) !=
          hipSuccess ||
      hipMemcpyAsync(&l, m->loss1, sizeof(float), hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
      hipStreamSynchronize(m->stream) != hipSuccess)
    return fail(SNC_E_HIP, "MLP gradient readback failed");
  if (loss) *loss = (double)l;
  return SNC_OK;
}

int snc_hip_mlp_train(snc_hip_mlp* m, const float* features, const float* labels, int64_t M,
                      int32_t epochs, int32_t batch, float lr, uint64_t seed, double* final_loss) {
  if (!m || !features || !labels || M < 1 || epochs < 0 || batch < 1)
    return fail(SNC_E_INVALID, "bad arguments");
  hip_clear();
  int rc = ensure_x(m, M);
  if (rc) return rc;
  // Shuffle on the host each epoch and upload the permuted examples once per
  // epoch; minibatches are then contiguous slices.
  std::vector<int64_t> order(M);
  std::iota(order.begin(), order.end(), 0);
  std::vector<float> Xp((size_t)M * kMlpIn), yp((size_t)M);
  SplitMix64 r(seed);
  double last = 0.0;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int64_t i = M - 1; i > 0; --i) std::swap(order[i], order[(int64_t)r.below((uint64_t)i + 1)]);
    for (int64_t i = 0; i < M; ++i) {
      std::memcpy(&Xp[(size_t)i * kMlpIn], features + order[i] * kMlpIn, sizeof(float) * kMlpIn);
      yp[i] = labels[order[i]];
    }
    if (hipMemcpy(m->X, Xp.data(), sizeof(float) * Xp.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->y, yp.data(), sizeof(float) * M, hipMemcpyHostToDevice) != hipSuccess)
      return fail(SNC_E_HIP, "example upload failed");
    double ep_loss = 0.0;
    for (int64_t off = 0; off < M; off += batch) {
      const int64_t mb = std::min<int64_t>(batch, M - off);
      rc = grow((void**)&m->G, m->cap_g, (int64_t)kMlpParams * mb, sizeof(float));
      if (!rc) rc = mlp_launch_grad(m->dev(), m->X + off * kMlpIn, m->y + off, mb, m->G, m->loss,
                                    m->g, m->loss1, m->stream);
      if (!rc) rc = mlp_launch_adam(m->w, m->m, m->v, m->g, lr, ++m->adam_step, m->dev(), m->stream);
      if (rc) return rc;
      float l = 0.0f;
      if (hipMemcpyAsync(&l, m->loss1, sizeof(float), hipMemcpyDeviceToHost, m->stream) !=
              hipSuccess ||
          hipStreamSynchronize(m->stream) != hipSuccess)
        return fail(SNC_E_HIP, "training step failed");
      ep_loss += (double)l * mb;
    }
    last = ep_loss / (double)M;
  }
  if (final_loss) *final_loss = last;
  return SNC_OK;
}

}  // extern "C"
