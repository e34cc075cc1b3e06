// Context of the Green-FK engine (private to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "gm_common.h"

struct gm_ctx {
  int device = 0;
  int variant = 1;
  int persistent = 0;      // 1: substep-scheduled persistent kernels, 0: one lane per job (default, faster here)
  int unit = 256;          // instances per work unit of the persistent kernels (at most one 256-thread block)
  int waves = 1;           // waves per SIMD the persistent kernels are compiled for (1 or 2)
  int sph_split = 4;       // blocks per work unit of the one-lane-per-job spheroidal kernel
  int ncu = 0;
  int max_lds = 65536;      // LDS bytes per block the device allows
  hipStream_t stream = nullptr;
  std::string err;
  GmModel hmodel;
  bool has_model = false, has_nodes = false;
  GmModel* dmodel = nullptr;
  // growable device scratch
  void* dbuf = nullptr;
  size_t dbuf_size = 0;
  hipEvent_t ev[7] = {};
  double timing[6] = {0};
  unsigned long long last_hist[48] = {0};
  // learned step proposer
  float* dmlp = nullptr;
  GmProposer prop;
  int prop_enabled = 0;
  int prop_every = 1;      // network consulted every prop_every accepted steps
  int order_target = 6;    // target try count of the step rule without a network (6: the reference's)
  int adj_pair = 0;        // 1: two adjoint integrations per leg give the four jumps (GM_OPT_ADJOINT_PAIR)
  // hipGraph of the kernel stages of the last batch (GM_OPT_GRAPH): replayed while the batch's
  // kernel arguments (GmBatch by value and the launch shape) are unchanged
  int graph = 0;
  hipGraph_t g_graph = nullptr;
  hipGraphExec_t g_exec = nullptr;
  std::vector<unsigned char> g_key;
  long long graph_captures = 0, graph_replays = 0;
  // step logging of the next run (gm_collect_steplog)
  int log_max = 0;
  float* log_host = nullptr;
  int log_count = 0;
  // synthetics half: growable device scratch, stage events and timing
  void* sbuf = nullptr;
  size_t sbuf_size = 0;
  double* dtw = nullptr;   // twiddle table of the last transform length
  int tw_n = 0;
  hipEvent_t sev[6] = {};
  double stiming[6] = {0};
};

// records the message of a failure on c (or of a failed context creation) and returns code
int gm_fail(gm_ctx* c, int code, const char* fmt, ...);
#define HIPCHK(c, x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return gm_fail((c), GM_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); } while (0)

// Batch without allocation (gm_noalloc.hip): every buffer lives in the
// per-device file-scope array of GM_NOALLOC_BYTES; at most GM_NOALLOC_MAXF
// frequencies per call (the plan is built by one block).
#define GM_NOALLOC_BYTES (128u << 20)
#define GM_NOALLOC_MAXF 1024
struct GmNoAlloc {
  void* base;       // the array's address on the context's device
  size_t bytes;     // its size
};
int gm_noalloc_init(int device);               // takes the array's address on the current device (once per device)
void* gm_noalloc_scratch(int device);
void gm_noalloc_plan(hipStream_t s, int nfreq, const double* freqs, const int* lmin, const int* nl, int* q, int* l,
                     double* omre, int* uq, int* us, int* ue, unsigned long long* hist, int* logc, int* smax);
void gm_noalloc_status(hipStream_t s, int n, const int* status, int* smax);
// the Green-FK column batch of gm_gfk_batch (device outputs) in the buffers of na
int gm_run_batch(gm_ctx* c, int nfreq, const double* freqs, const int* lmin, const int* nl, double dwn, double sigma,
                 double eps, int attmode, int jre, int istyp, double* gf, float* gfre, float* gfim, int* status,
                 long long* stats, const GmNoAlloc* na);
