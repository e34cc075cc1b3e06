// Green-FK batch that allocates nothing: gm_gfk_batch_noalloc.
//
// Every buffer of the batch beyond the caller's inputs and outputs (spline
// tables and LDS images, the instance plan, minors, unit-jump Green functions,
// counters) lives in one file-scope device array, `scratch`, of
// GM_NOALLOC_BYTES per GPU, held until the process ends.  Its address on each
// device is taken once, when the first context on that device is created
// (gm_noalloc_init); host code uses that address, never the array's name.
// The plan that gm_gfk_batch builds in host vectors is built here on the
// device by one block, so the call touches neither device nor host heap.
// All calls on one GPU share the array: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>
#include <mutex>
#include "gm_common.h"
#include "gm_ctx.h"
#include "../../include/gemini_fk.h"

static __device__ __align__(256) unsigned char scratch[GM_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;
}  // namespace

int gm_noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return GM_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return GM_OK;
  void* p = nullptr;
  // the caller has made `device` current
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GM_ERR_HIP;
  g_scratch[device] = p;
  return GM_OK;
}

void* gm_noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[device];
}

// Plan of a batch (the host plan of gm_gfk_batch): instance i's frequency slot
// q[i] and degree l[i], omega of every frequency, and the work units of at most
// 256 consecutive degrees of one frequency, highest frequency first.  One block
// of GM_NOALLOC_MAXF threads, thread f owning frequency f.
__global__ void __launch_bounds__(GM_NOALLOC_MAXF) k_noalloc_plan(int nfreq, const double* __restrict__ freqs,
                                                                  const int* __restrict__ lmin,
                                                                  const int* __restrict__ nl, int* __restrict__ q,
                                                                  int* __restrict__ l, double* __restrict__ omre,
                                                                  int* __restrict__ uq, int* __restrict__ us,
                                                                  int* __restrict__ ue) {
  __shared__ int ioff[GM_NOALLOC_MAXF + 1], uoff[GM_NOALLOC_MAXF + 1], order[GM_NOALLOC_MAXF];
  const int f = threadIdx.x;
  int rank = 0;
  if (f < nfreq) {
    // position of f when the frequencies are ordered by decreasing value (ties by index)
    for (int g = 0; g < nfreq; ++g) rank += (freqs[g] > freqs[f] || (freqs[g] == freqs[f] && g < f)) ? 1 : 0;
    order[rank] = f;
  }
  __syncthreads();
  if (f == 0) {
    ioff[0] = 0;
    for (int g = 0; g < nfreq; ++g) ioff[g + 1] = ioff[g] + nl[g];
    uoff[0] = 0;
    for (int r = 0; r < nfreq; ++r) uoff[r + 1] = uoff[r] + (nl[order[r]] + 255) / 256;
  }
  __syncthreads();
  if (f >= nfreq) return;
  omre[f] = 2.0 * (double)3.1415927410125732 * freqs[f];
  const int a = ioff[f], n = nl[f];
  for (int u = 0, k = uoff[rank]; u * 256 < n; ++u, ++k) {
    uq[k] = f;
    us[k] = a + 256 * u;
    ue[k] = a + (256 * (u + 1) < n ? 256 * (u + 1) : n);
  }
  for (int i = 0; i < n; ++i) { q[a + i] = f; l[a + i] = lmin[f] + i; }
}

// zero the histogram, the step-log counter and the failed-status summary
// (a fill kernel rather than a memset, so that the call can be captured)
__global__ void k_noalloc_zero(unsigned long long* __restrict__ hist, int* __restrict__ logc, int* __restrict__ smax) {
  const int t = threadIdx.x;
  if (t < 48) hist[t] = 0ULL;
  if (t < 4) logc[t] = 0;
  if (t == 0) *smax = 0;
}

// largest status of the batch (0: every instance succeeded)
__global__ void k_noalloc_status(int n, const int* __restrict__ status, int* __restrict__ smax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && status[i]) atomicMax(smax, status[i]);
}

void gm_noalloc_plan(hipStream_t s, int nfreq, const double* freqs, const int* lmin, const int* nl, int* q, int* l,
                     double* omre, int* uq, int* us, int* ue, unsigned long long* hist, int* logc, int* smax) {
  hipLaunchKernelGGL(k_noalloc_zero, dim3(1), dim3(64), 0, s, hist, logc, smax);
  hipLaunchKernelGGL(k_noalloc_plan, dim3(1), dim3(GM_NOALLOC_MAXF), 0, s, nfreq, freqs, lmin, nl, q, l, omre, uq, us,
                     ue);
}

void gm_noalloc_status(hipStream_t s, int n, const int* status, int* smax) {
  if (n > 0) hipLaunchKernelGGL(k_noalloc_status, dim3((n + 255) / 256), dim3(256), 0, s, n, status, smax);
}

extern "C" int gm_gfk_batch_noalloc(gm_ctx* c, int nfreq, const double* freqs, const int* lmin, const int* nl,
                                    double sigma, double eps, int attmode, int jre, int istyp, double* gf,
                                    float* gfre, float* gfim, int* status, long long* stats) {
  if (!c) return GM_ERR_ARG;
  void* base = gm_noalloc_scratch(c->device);
  if (!base) return gm_fail(c, GM_ERR_HIP, "no scratch address on device %d", c->device);
  GmNoAlloc na{base, (size_t)GM_NOALLOC_BYTES};
  return gm_run_batch(c, nfreq, freqs, lmin, nl, 0.0, sigma, eps, attmode, jre, istyp, gf, gfre, gfim, status, stats,
                      &na);
}
