// CPU oracle of the grow-proposal MLP -- TEST INFRASTRUCTURE ONLY.
//
// The MLP has no counterpart in the reference; this is its mathematical
// definition in FP64: s = W3 relu(W2 relu(W1 x + b1) + b2) + b3 with the
// parameter layout W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3, loss =
// mean binary cross-entropy with logits. With quantize != 0 the weights, the
// input and the first hidden layer are rounded to BF16 (nearest even) as the
// device model does, and the backward pass treats that rounding as identity
// (straight-through). With quantize == 0 it is the plain FP64 network whose
// analytic gradient finite differences can check.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int IN = 16, H = 64;
constexpr int OW1 = 0, OB1 = OW1 + H * IN, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H,
              OB3 = OW3 + H, NP = OB3 + 1;

double bf16(double x) {
  float f = (float)x;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  std::memcpy(&f, &u, 4);
  return f;
}

struct Net {
  std::vector<double> w;
  bool qz;
  double W(int off, int k) const { return qz && off != OB1 && off != OB2 && off != OB3 ? bf16(w[off + k]) : w[off + k]; }
};

double forward(const Net& n, const float* xin, double* x, double* h1, double* h2) {
  for (int k = 0; k < IN; ++k) x[k] = n.qz ? bf16(xin[k]) : (double)xin[k];
  for (int j = 0; j < H; ++j) {
    double s = n.W(OB1, j);
    for (int k = 0; k < IN; ++k) s += n.W(OW1, j * IN + k) * x[k];
    s = s > 0 ? s : 0.0;
    h1[j] = n.qz ? bf16(s) : s;
  }
  double out = n.W(OB3, 0);
  for (int j = 0; j < H; ++j) {
    double s = n.W(OB2, j);
    for (int k = 0; k < H; ++k) s += n.W(OW2, j * H + k) * h1[k];
    h2[j] = s > 0 ? s : 0.0;
    out += n.W(OW3, j) * h2[j];
  }
  return out;
}

}  // namespace

extern "C" {

int oracle_mlp_nparams() { return NP; }

void oracle_mlp_forward(const float* params, const float* X, int64_t M, int32_t quantize,
                        double* out) {
  Net n{std::vector<double>(params, params + NP), quantize != 0};
  double x[IN], h1[H], h2[H];
  for (int64_t e = 0; e < M; ++e) out[e] = forward(n, X + e * IN, x, h1, h2);
}

// Mean BCE-with-logits loss and its gradient (FP64).
double oracle_mlp_grad(const float* params, const float* X, const float* y, int64_t M,
                       int32_t quantize, double* grad) {
  Net n{std::vector<double>(params, params + NP), quantize != 0};
  std::vector<double> g(NP, 0.0);
  double x[IN], h1[H], h2[H], dh1[H];
  double loss = 0.0;
  for (int64_t e = 0; e < M; ++e) {
    const double o = forward(n, X + e * IN, x, h1, h2);
    const double t = y[e];
    loss += std::fmax(o, 0.0) - o * t + std::log1p(std::exp(-std::fabs(o)));
    const double d = (1.0 / (1.0 + std::exp(-o)) - t) / (double)M;
    g[OB3] += d;
    for (int k = 0; k < H; ++k) dh1[k] = 0.0;
    for (int j = 0; j < H; ++j) {
      g[OW3 + j] += d * h2[j];
      const double d2 = h2[j] > 0 ? d * n.W(OW3, j) : 0.0;
      g[OB2 + j] += d2;
      for (int k = 0; k < H; ++k) {
        g[OW2 + j * H + k] += d2 * h1[k];
        dh1[k] += d2 * n.W(OW2, j * H + k);
      }
    }
    for (int k = 0; k < H; ++k) {
      const double d1 = h1[k] > 0 ? dh1[k] : 0.0;
      g[OB1 + k] += d1;
      for (int f = 0; f < IN; ++f) g[OW1 + k * IN + f] += d1 * x[f];
    }
  }
  std::memcpy(grad, g.data(), sizeof(double) * NP);
  return loss / (double)M;
}

}  // extern "C"
