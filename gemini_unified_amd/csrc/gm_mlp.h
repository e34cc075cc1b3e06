// Step-proposer network layout and in-kernel inference.
#pragma once
#include <hip/hip_runtime.h>

#define GM_MLP_IN 9
#define GM_MLP_H 64
#define GM_MLP_OUT 2
#define GM_MLP_NPARAM (GM_MLP_IN * GM_MLP_H + GM_MLP_H + 2 * (GM_MLP_H * GM_MLP_H + GM_MLP_H) + GM_MLP_H * GM_MLP_OUT + GM_MLP_OUT)

// feature / target normalisation and control of the learned proposer
struct GmProposer {
  float mu[GM_MLP_IN], isd[GM_MLP_IN];   // x_norm = (x - mu) * isd
  float omu[GM_MLP_OUT], osd[GM_MLP_OUT]; // y = y_norm * osd + omu
  float safety;                           // proposed step = safety * predicted step
  int kcheck;                             // 1: first error test at the predicted try - 1
  int every;                              // consult the network every `every` accepted steps
  int pad;
};

// One forward pass for one lane.  Every lane reads the same weight at the same
// time (uniform address), so the loads go through the scalar data cache and the
// FMAs take the weight as a scalar operand.  Hidden activations stay in VGPRs.
__device__ __forceinline__ void gmMlpInfer(const float* __restrict__ P, const float* x, float* y) {
  const float* W1 = P;
  const float* b1 = W1 + GM_MLP_IN * GM_MLP_H;
  const float* W2 = b1 + GM_MLP_H;
  const float* b2 = W2 + GM_MLP_H * GM_MLP_H;
  const float* W3 = b2 + GM_MLP_H;
  const float* b3 = W3 + GM_MLP_H * GM_MLP_H;
  const float* W4 = b3 + GM_MLP_H;
  const float* b4 = W4 + GM_MLP_H * GM_MLP_OUT;
  float h1[GM_MLP_H], h2[GM_MLP_H];
#pragma unroll
  for (int o = 0; o < GM_MLP_H; ++o) {
    float a = b1[o];
#pragma unroll
    for (int i = 0; i < GM_MLP_IN; ++i) a += x[i] * W1[i * GM_MLP_H + o];
    h1[o] = a > 0.0f ? a : 0.0f;
  }
#pragma unroll
  for (int o = 0; o < GM_MLP_H; ++o) {
    float a = b2[o];
#pragma unroll
    for (int i = 0; i < GM_MLP_H; ++i) a += h1[i] * W2[i * GM_MLP_H + o];
    h2[o] = a > 0.0f ? a : 0.0f;
  }
#pragma unroll
  for (int o = 0; o < GM_MLP_H; ++o) {
    float a = b3[o];
#pragma unroll
    for (int i = 0; i < GM_MLP_H; ++i) a += h2[i] * W3[i * GM_MLP_H + o];
    h1[o] = a > 0.0f ? a : 0.0f;
  }
#pragma unroll
  for (int o = 0; o < GM_MLP_OUT; ++o) {
    float a = b4[o];
#pragma unroll
    for (int i = 0; i < GM_MLP_H; ++i) a += h1[i] * W4[i * GM_MLP_OUT + o];
    y[o] = a;
  }
}
