// Step-proposer network layout and in-kernel inference.
#pragma once
#include <hip/hip_runtime.h>

#define GM_MLP_IN 9
#define GM_MLP_H 64
#define GM_MLP_OUT 2
#define GM_MLP_NPARAM (GM_MLP_IN * GM_MLP_H + GM_MLP_H + 2 * (GM_MLP_H * GM_MLP_H + GM_MLP_H) + GM_MLP_H * GM_MLP_OUT + GM_MLP_OUT)

// host-side entry of the batched forward pass on device buffers (gm_mlp.hip)
#define GM_MLP_CHUNK (1 << 19)
size_t gm_mlp_forward_work(int n);
const float* gm_mlp_forward_dev(hipStream_t s, const float* params, const float* X, int n, float* work);

// feature / target normalisation and control of the learned proposer
struct GmProposer {
  float mu[GM_MLP_IN], isd[GM_MLP_IN];   // x_norm = (x - mu) * isd
  float omu[GM_MLP_OUT], osd[GM_MLP_OUT]; // y = y_norm * osd + omu
  float safety;                           // proposed step = safety * predicted step
  int kcheck;                             // 1: first error test at the predicted try - 1
  int every;                              // consult the network every `every` accepted steps
  int first_step;                         // per-instance policy: also apply the network's first step
};

// One forward pass for one lane.  Every lane reads the same weight at the same
// time, and no kernel writes the weights while it runs, so the pointer is made
// wave-uniform and constant-address-space: the weights then come in through
// scalar loads and the FMAs take them as scalar operands (a per-lane generic
// pointer costs one vector load per four weights, about 2,300 per proposal).
// Hidden activations stay in VGPRs.
typedef const __attribute__((address_space(4))) float* GmConstF;
__device__ __forceinline__ void gmMlpInfer(const float* __restrict__ Pg, const float* x, float* y) {
  const unsigned long long pa = (unsigned long long)Pg;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)pa);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(pa >> 32));
  GmConstF P = (GmConstF)((((unsigned long long)hi) << 32) | lo);
  GmConstF W1 = P;
  GmConstF b1 = W1 + GM_MLP_IN * GM_MLP_H;
  GmConstF W2 = b1 + GM_MLP_H;
  GmConstF b2 = W2 + GM_MLP_H * GM_MLP_H;
  GmConstF W3 = b2 + GM_MLP_H;
  GmConstF b3 = W3 + GM_MLP_H * GM_MLP_H;
  GmConstF W4 = b3 + GM_MLP_H;
  GmConstF b4 = W4 + GM_MLP_H * GM_MLP_OUT;
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
