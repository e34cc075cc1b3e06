// Step-proposer network: 9 -> 64 -> 64 -> 64 -> 2 multilayer perceptron in
// FP32 with ReLU hidden activations, trained on the GPU with these kernels
// (forward, backward, Adam).  Parameters are one contiguous FP32 array:
//   W1[9][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][2] b4[2]
// (row-major, input index first).  Activations of a batch are stored
// sample-major: X[b][i].
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <vector>
#include <cmath>
#include "gm_common.h"
#include "gm_mlp.h"

namespace gm_mlp {

// C[M][N] (+)= sum_k A'[m][k] B'[k][n]; A' = A or A^T, B' = B or B^T.
// 16x16 tiles staged in LDS, one thread per output element.
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const float* __restrict__ A, int lda, int ta,
                                              const float* __restrict__ B, int ldb, int tb, float* __restrict__ Cm,
                                              int ldc, int accumulate) {
  __shared__ float sA[16][17], sB[16][17];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
  float acc = 0.0f;
  for (int k0 = 0; k0 < K; k0 += 16) {
    int ka = k0 + tx, kb = k0 + ty;
    float a = 0.0f, b = 0.0f;
    if (row < M && ka < K) a = ta ? A[(size_t)ka * lda + row] : A[(size_t)row * lda + ka];
    if (kb < K && col < N) b = tb ? B[(size_t)col * ldb + kb] : B[(size_t)kb * ldb + col];
    sA[ty][tx] = a;
    sB[ty][tx] = b;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) acc += sA[ty][k] * sB[k][tx];
    __syncthreads();
  }
  if (row < M && col < N) {
    float* c = &Cm[(size_t)row * ldc + col];
    *c = accumulate ? *c + acc : acc;
  }
}

// Z[b][o] += bias[o]; Y = relu(Z) (or identity for the output layer)
__global__ void k_bias_act(int Bn, int O, const float* __restrict__ bias, float* __restrict__ Z, float* __restrict__ Y,
                           int relu) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Bn * O) return;
  float z = Z[t] + bias[t % O];
  Z[t] = z;
  Y[t] = relu ? (z > 0.0f ? z : 0.0f) : z;
}

// dZ = dY * relu'(Z) in place on dY
__global__ void k_relu_back(int n, const float* __restrict__ Z, float* __restrict__ dY) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && !(Z[t] > 0.0f)) dY[t] = 0.0f;
}

// column sums: db[o] = sum_b dZ[b][o]
__global__ void k_colsum(int Bn, int O, const float* __restrict__ dZ, float* __restrict__ db) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= O) return;
  float s = 0.0f;
  for (int b = 0; b < Bn; ++b) s += dZ[(size_t)b * O + o];
  db[o] = s;
}

// loss = mean_b sum_j w_j (y_bj - t_bj)^2 ; dY = 2 w_j (y - t) / B
__global__ void k_loss(int Bn, const float* __restrict__ Y, const float* __restrict__ Tg, float w0, float w1,
                       float* __restrict__ dY, float* __restrict__ lossacc) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (b < Bn) {
    float e0 = Y[2 * b] - Tg[2 * b], e1 = Y[2 * b + 1] - Tg[2 * b + 1];
    dY[2 * b] = 2.0f * w0 * e0 / Bn;
    dY[2 * b + 1] = 2.0f * w1 * e1 / Bn;
    l = (w0 * e0 * e0 + w1 * e1 * e1) / Bn;
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(lossacc, l);
}

__global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float gt = g[t];
  float a = b1 * m1[t] + (1.0f - b1) * gt;
  float b = b2 * m2[t] + (1.0f - b2) * gt * gt;
  m1[t] = a;
  m2[t] = b;
  p[t] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// gather a minibatch: rows idx[b] of X (9 wide) and T (2 wide)
__global__ void k_gather(int Bn, const int* __restrict__ idx, const float* __restrict__ X, const float* __restrict__ T,
                         float* __restrict__ xb, float* __restrict__ tb) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Bn) return;
  int r = idx[t];
  for (int i = 0; i < GM_MLP_IN; ++i) xb[t * GM_MLP_IN + i] = X[(size_t)r * GM_MLP_IN + i];
  tb[2 * t] = T[(size_t)r * 2];
  tb[2 * t + 1] = T[(size_t)r * 2 + 1];
}

struct Net {
  const float *W1, *b1, *W2, *b2, *W3, *b3, *W4, *b4;
};
static Net view(const float* p) {
  Net n;
  n.W1 = p; n.b1 = n.W1 + GM_MLP_IN * GM_MLP_H; n.W2 = n.b1 + GM_MLP_H; n.b2 = n.W2 + GM_MLP_H * GM_MLP_H;
  n.W3 = n.b2 + GM_MLP_H; n.b3 = n.W3 + GM_MLP_H * GM_MLP_H; n.W4 = n.b3 + GM_MLP_H; n.b4 = n.W4 + GM_MLP_H * GM_MLP_OUT;
  return n;
}

static void gemm(hipStream_t s, int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb, int tb,
                 float* C, int ldc, int acc) {
  dim3 g((N + 15) / 16, (M + 15) / 16);
  hipLaunchKernelGGL(k_gemm, g, dim3(256), 0, s, M, N, K, A, lda, ta, B, ldb, tb, C, ldc, acc);
}

// workspace of one batch: Z1,Y1,Z2,Y2,Z3,Y3,Y4 and the matching gradients
struct Work {
  float *z1, *y1, *z2, *y2, *z3, *y3, *y4, *d1, *d2, *d3, *d4;
};

// forward of a batch X[Bn][9] -> Y4[Bn][2]
static void forward(hipStream_t s, const float* P, const float* X, int Bn, Work& w) {
  Net n = view(P);
  const int H = GM_MLP_H, T = 256;
  gemm(s, Bn, H, GM_MLP_IN, X, GM_MLP_IN, 0, n.W1, H, 0, w.z1, H, 0);
  hipLaunchKernelGGL(k_bias_act, dim3((Bn * H + T - 1) / T), dim3(T), 0, s, Bn, H, n.b1, w.z1, w.y1, 1);
  gemm(s, Bn, H, H, w.y1, H, 0, n.W2, H, 0, w.z2, H, 0);
  hipLaunchKernelGGL(k_bias_act, dim3((Bn * H + T - 1) / T), dim3(T), 0, s, Bn, H, n.b2, w.z2, w.y2, 1);
  gemm(s, Bn, H, H, w.y2, H, 0, n.W3, H, 0, w.z3, H, 0);
  hipLaunchKernelGGL(k_bias_act, dim3((Bn * H + T - 1) / T), dim3(T), 0, s, Bn, H, n.b3, w.z3, w.y3, 1);
  gemm(s, Bn, GM_MLP_OUT, H, w.y3, H, 0, n.W4, GM_MLP_OUT, 0, w.y4, GM_MLP_OUT, 0);
  hipLaunchKernelGGL(k_bias_act, dim3((Bn * GM_MLP_OUT + T - 1) / T), dim3(T), 0, s, Bn, GM_MLP_OUT, n.b4, w.y4, w.y4, 0);
}

// backward from dY4 (in w.d4) to parameter gradients G (same layout as P)
static void backward(hipStream_t s, const float* P, const float* X, int Bn, Work& w, float* G) {
  Net n = view(P);
  Net g = view(G);
  const int H = GM_MLP_H, T = 256;
  float* gW1 = const_cast<float*>(g.W1); float* gb1 = const_cast<float*>(g.b1);
  float* gW2 = const_cast<float*>(g.W2); float* gb2 = const_cast<float*>(g.b2);
  float* gW3 = const_cast<float*>(g.W3); float* gb3 = const_cast<float*>(g.b3);
  float* gW4 = const_cast<float*>(g.W4); float* gb4 = const_cast<float*>(g.b4);
  // layer 4 (linear)
  gemm(s, H, GM_MLP_OUT, Bn, w.y3, H, 1, w.d4, GM_MLP_OUT, 0, gW4, GM_MLP_OUT, 0);
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(64), 0, s, Bn, GM_MLP_OUT, w.d4, gb4);
  gemm(s, Bn, H, GM_MLP_OUT, w.d4, GM_MLP_OUT, 0, n.W4, GM_MLP_OUT, 1, w.d3, H, 0);
  // layer 3
  hipLaunchKernelGGL(k_relu_back, dim3((Bn * H + T - 1) / T), dim3(T), 0, s, Bn * H, w.z3, w.d3);
  gemm(s, H, H, Bn, w.y2, H, 1, w.d3, H, 0, gW3, H, 0);
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(64), 0, s, Bn, H, w.d3, gb3);
  gemm(s, Bn, H, H, w.d3, H, 0, n.W3, H, 1, w.d2, H, 0);
  // layer 2
  hipLaunchKernelGGL(k_relu_back, dim3((Bn * H + T - 1) / T), dim3(T), 0, s, Bn * H, w.z2, w.d2);
  gemm(s, H, H, Bn, w.y1, H, 1, w.d2, H, 0, gW2, H, 0);
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(64), 0, s, Bn, H, w.d2, gb2);
  gemm(s, Bn, H, H, w.d2, H, 0, n.W2, H, 1, w.d1, H, 0);
  // layer 1
  hipLaunchKernelGGL(k_relu_back, dim3((Bn * H + T - 1) / T), dim3(T), 0, s, Bn * H, w.z1, w.d1);
  gemm(s, GM_MLP_IN, H, Bn, X, GM_MLP_IN, 1, w.d1, H, 0, gW1, H, 0);
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(64), 0, s, Bn, H, w.d1, gb1);
}

static size_t work_floats(int Bn) { return (size_t)Bn * (7 * GM_MLP_H + 2 * GM_MLP_OUT) + (size_t)Bn * 3 * GM_MLP_H; }
static void carve(float* base, int Bn, Work& w) {
  const size_t H = (size_t)Bn * GM_MLP_H, O = (size_t)Bn * GM_MLP_OUT;
  w.z1 = base; w.y1 = w.z1 + H; w.z2 = w.y1 + H; w.y2 = w.z2 + H; w.z3 = w.y2 + H; w.y3 = w.z3 + H;
  w.y4 = w.y3 + H; w.d4 = w.y4 + O; w.d3 = w.d4 + O; w.d2 = w.d3 + H; w.d1 = w.d2 + H;
}

}  // namespace gm_mlp

using namespace gm_mlp;

// forward of n device-resident samples on stream s (inference inside a batch
// call): X[n][9] -> the returned Y[n][2] inside work (gm_mlp_forward_work(n)
// floats).  The GEMM grid's second dimension is n/16 blocks, so callers pass
// at most GM_MLP_CHUNK samples at a time.
size_t gm_mlp_forward_work(int n) { return work_floats(n); }
const float* gm_mlp_forward_dev(hipStream_t s, const float* params, const float* X, int n, float* work) {
  Work w;
  carve(work, n, w);
  forward(s, params, X, n, w);
  return w.y4;
}

// ---------------------------------------------------------------------------
// C ABI (declared in include/gemini_fk.h)
// ---------------------------------------------------------------------------
#define MCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { fprintf(stderr, "gm_mlp: %s: %s\n", #x, hipGetErrorString(e_)); return GM_ERR_HIP; } } while (0)

namespace {
// device buffer released on every exit path (hipFree waits for pending work)
template <class T>
struct DevBuf {
  T* p = nullptr;
  hipError_t alloc(size_t n) { return hipMalloc(&p, sizeof(T) * n); }
  ~DevBuf() { if (p) (void)hipFree(p); }
  operator T*() const { return p; }
};
struct StreamGuard {
  hipStream_t s = nullptr;
  ~StreamGuard() { if (s) (void)hipStreamDestroy(s); }
};
}  // namespace

extern "C" int gm_mlp_nparams(void) { return GM_MLP_NPARAM; }

// forward of n samples (host arrays): X[n][9] -> Y[n][2]
extern "C" int gm_mlp_forward(int device, const float* params, int n, const float* X, float* Y) {
  if (n <= 0 || !params || !X || !Y) return GM_ERR_ARG;
  MCHK(hipSetDevice(device));
  DevBuf<float> dp, dx, dw;
  MCHK(dp.alloc(GM_MLP_NPARAM));
  MCHK(dx.alloc((size_t)n * GM_MLP_IN));
  MCHK(dw.alloc(work_floats(n)));
  MCHK(hipMemcpy(dp, params, sizeof(float) * GM_MLP_NPARAM, hipMemcpyHostToDevice));
  MCHK(hipMemcpy(dx, X, sizeof(float) * (size_t)n * GM_MLP_IN, hipMemcpyHostToDevice));
  Work w;
  carve(dw, n, w);
  forward(0, dp, dx, n, w);
  MCHK(hipGetLastError());
  MCHK(hipMemcpy(Y, w.y4, sizeof(float) * (size_t)n * GM_MLP_OUT, hipMemcpyDeviceToHost));
  return GM_OK;
}

// loss and parameter gradient of one batch (for gradient checks)
extern "C" int gm_mlp_grad(int device, const float* params, int n, const float* X, const float* T, float w0, float w1,
                           float* loss, float* grad) {
  if (n <= 0 || !params || !X || !T || !grad) return GM_ERR_ARG;
  MCHK(hipSetDevice(device));
  DevBuf<float> dp, dx, dt, dw, dg, dl;
  MCHK(dp.alloc(GM_MLP_NPARAM));
  MCHK(dg.alloc(GM_MLP_NPARAM));
  MCHK(dx.alloc((size_t)n * GM_MLP_IN));
  MCHK(dt.alloc((size_t)n * 2));
  MCHK(dw.alloc(work_floats(n)));
  MCHK(dl.alloc(1));
  MCHK(hipMemcpy(dp, params, sizeof(float) * GM_MLP_NPARAM, hipMemcpyHostToDevice));
  MCHK(hipMemcpy(dx, X, sizeof(float) * (size_t)n * GM_MLP_IN, hipMemcpyHostToDevice));
  MCHK(hipMemcpy(dt, T, sizeof(float) * (size_t)n * 2, hipMemcpyHostToDevice));
  float zero = 0.0f;
  MCHK(hipMemcpy(dl, &zero, sizeof(float), hipMemcpyHostToDevice));
  Work w;
  carve(dw, n, w);
  forward(0, dp, dx, n, w);
  hipLaunchKernelGGL(k_loss, dim3((n + 255) / 256), dim3(256), 0, 0, n, w.y4, dt, w0, w1, w.d4, dl);
  backward(0, dp, dx, n, w, dg);
  MCHK(hipGetLastError());
  MCHK(hipMemcpy(grad, dg, sizeof(float) * GM_MLP_NPARAM, hipMemcpyDeviceToHost));
  if (loss) MCHK(hipMemcpy(loss, dl, sizeof(float), hipMemcpyDeviceToHost));
  return GM_OK;
}

// Adam training on n samples X[n][9], T[n][2]; params in/out; losses[epoch] out
extern "C" int gm_mlp_train(int device, float* params, int n, const float* X, const float* T, float w0, float w1,
                            int epochs, int batch, float lr, unsigned seed, float* losses) {
  if (n <= 0 || batch <= 0 || epochs <= 0 || !params || !X || !T) return GM_ERR_ARG;
  if (batch > n) batch = n;
  MCHK(hipSetDevice(device));
  // buffers are declared before the stream so that the stream is destroyed first on every exit
  DevBuf<float> dp, dg, dm1, dm2, dx, dt, xb, tb, dw, dl;
  DevBuf<int> didx;
  StreamGuard sg;
  MCHK(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
  hipStream_t s = sg.s;
  MCHK(dp.alloc(GM_MLP_NPARAM));
  MCHK(dg.alloc(GM_MLP_NPARAM));
  MCHK(dm1.alloc(GM_MLP_NPARAM));
  MCHK(dm2.alloc(GM_MLP_NPARAM));
  MCHK(dx.alloc((size_t)n * GM_MLP_IN));
  MCHK(dt.alloc((size_t)n * 2));
  MCHK(xb.alloc((size_t)batch * GM_MLP_IN));
  MCHK(tb.alloc((size_t)batch * 2));
  MCHK(dw.alloc(work_floats(batch)));
  MCHK(dl.alloc((size_t)epochs + 1));
  MCHK(didx.alloc((size_t)n));
  MCHK(hipMemcpyAsync(dp, params, sizeof(float) * GM_MLP_NPARAM, hipMemcpyHostToDevice, s));
  MCHK(hipMemcpyAsync(dx, X, sizeof(float) * (size_t)n * GM_MLP_IN, hipMemcpyHostToDevice, s));
  MCHK(hipMemcpyAsync(dt, T, sizeof(float) * (size_t)n * 2, hipMemcpyHostToDevice, s));
  std::vector<float> zeros(std::max<size_t>(GM_MLP_NPARAM, (size_t)epochs + 1), 0.0f);
  MCHK(hipMemcpyAsync(dm1, zeros.data(), sizeof(float) * GM_MLP_NPARAM, hipMemcpyHostToDevice, s));
  MCHK(hipMemcpyAsync(dm2, zeros.data(), sizeof(float) * GM_MLP_NPARAM, hipMemcpyHostToDevice, s));
  MCHK(hipMemcpyAsync(dl, zeros.data(), sizeof(float) * (epochs + 1), hipMemcpyHostToDevice, s));
  Work w;
  carve(dw, batch, w);
  std::vector<int> perm(n);
  unsigned st = seed ? seed : 1u;
  auto rnd = [&]() { st ^= st << 13; st ^= st >> 17; st ^= st << 5; return st; };
  const float b1 = 0.9f, b2 = 0.999f, ae = 1e-8f;
  int step = 0;
  const int nb = n / batch;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = n - 1; i > 0; --i) { int j = (int)(rnd() % (unsigned)(i + 1)); std::swap(perm[i], perm[j]); }
    MCHK(hipMemcpyAsync(didx, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s));
    for (int b = 0; b < nb; ++b) {
      ++step;
      hipLaunchKernelGGL(k_gather, dim3((batch + 255) / 256), dim3(256), 0, s, batch, didx + (size_t)b * batch, dx, dt, xb, tb);
      forward(s, dp, xb, batch, w);
      hipLaunchKernelGGL(k_loss, dim3((batch + 255) / 256), dim3(256), 0, s, batch, w.y4, tb, w0, w1, w.d4, dl + ep);
      backward(s, dp, xb, batch, w, dg);
      float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
      hipLaunchKernelGGL(k_adam, dim3((GM_MLP_NPARAM + 255) / 256), dim3(256), 0, s, GM_MLP_NPARAM, dp, dg, dm1, dm2,
                         lr, b1, b2, ae, c1, c2);
    }
    // sync the per-epoch RNG upload (perm is reused on the host)
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(s));
  }
  MCHK(hipMemcpyAsync(params, dp, sizeof(float) * GM_MLP_NPARAM, hipMemcpyDeviceToHost, s));
  std::vector<float> hl(epochs);
  MCHK(hipMemcpyAsync(hl.data(), dl, sizeof(float) * epochs, hipMemcpyDeviceToHost, s));
  MCHK(hipStreamSynchronize(s));
  if (losses) for (int ep = 0; ep < epochs; ++ep) losses[ep] = hl[ep] / (float)std::max(nb, 1);
  return GM_OK;
}
