// C ABI of the Green-FK engine: context, model upload, batch orchestration.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <cstdarg>
#include <cmath>
#include <string>
#include <vector>
#include <algorithm>
#include "gm_common.h"
#include "gm_ctx.h"
#include "../../include/gemini_fk.h"

namespace gm_exact {
void launch_node_props(const GmModel*, const double*, int, GmKnot*, unsigned char*, int, unsigned char*, int, double*,
                       double*, double*, hipStream_t);
void launch_stage(int, const GmModel*, const GmBatch&, int, const double*, int, GmKnot*, hipStream_t, int, size_t, int);
int persist_blocks_per_cu(size_t, int);
void launch_policy(const GmModel*, const GmBatch&, hipStream_t, size_t, float*, const float*, float*, float*, double, int);
}
namespace gm_fast {
void launch_node_props(const GmModel*, const double*, int, GmKnot*, unsigned char*, int, unsigned char*, int, double*,
                       double*, double*, hipStream_t);
void launch_stage(int, const GmModel*, const GmBatch&, int, const double*, int, GmKnot*, hipStream_t, int, size_t, int);
int persist_blocks_per_cu(size_t, int);
void launch_policy(const GmModel*, const GmBatch&, hipStream_t, size_t, float*, const float*, float*, float*, double, int);
}

// per-system constants of the algorithmic work model
static const int kN[5] = {5, 2, 2, 2, 4};
static const int kNNZ[5] = {19, 4, 3, 4, 14};
static const int kESYS[5] = {219, 148, 156, 196, 206};

static std::string g_create_err;

int gm_fail(gm_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf; else g_create_err = buf;
  return code;
}

extern "C" int gm_ctx_create(int device, gm_ctx** out) {
  if (!out) return GM_ERR_ARG;
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return gm_fail(nullptr, GM_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  if (device < 0 || device >= n) return gm_fail(nullptr, GM_ERR_ARG, "device %d out of range", device);
  gm_ctx* c = new gm_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->dmodel, sizeof(GmModel)) != hipSuccess) {
    delete c;
    return gm_fail(nullptr, GM_ERR_HIP, "context creation failed on device %d", device);
  }
  bool ok = true;
  for (int i = 0; i < 7; ++i) ok = ok && hipEventCreate(&c->ev[i]) == hipSuccess;
  for (int i = 0; i < 6; ++i) ok = ok && hipEventCreate(&c->sev[i]) == hipSuccess;
  ok = ok && hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess;
  ok = ok && hipDeviceGetAttribute(&c->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess;
  if (!ok) {
    gm_ctx_destroy(c);
    return gm_fail(nullptr, GM_ERR_HIP, "context creation failed on device %d", device);
  }
  if (c->ncu <= 0) c->ncu = 1;
  // the address of the batch-without-allocation array on this device (once per device)
  if (gm_noalloc_init(device) != GM_OK) {
    gm_ctx_destroy(c);
    return gm_fail(nullptr, GM_ERR_HIP, "no address of the scratch array on device %d", device);
  }
  memset(&c->hmodel, 0, sizeof(GmModel));
  *out = c;
  return GM_OK;
}

extern "C" int gm_ctx_destroy(gm_ctx* c) {
  if (!c) return GM_OK;
  // teardown: release everything even if one call reports an error
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->dbuf) (void)hipFree(c->dbuf);
  if (c->dmodel) (void)hipFree(c->dmodel);
  if (c->dmlp) (void)hipFree(c->dmlp);
  for (int i = 0; i < 7; ++i) if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  for (int i = 0; i < 6; ++i) if (c->sev[i]) (void)hipEventDestroy(c->sev[i]);
  if (c->sbuf) (void)hipFree(c->sbuf);
  if (c->dtw) (void)hipFree(c->dtw);
  if (c->g_exec) (void)hipGraphExecDestroy(c->g_exec);
  if (c->g_graph) (void)hipGraphDestroy(c->g_graph);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return GM_OK;
}

extern "C" int gm_set_option(gm_ctx* c, int option, int value) {
  if (!c) return GM_ERR_ARG;
  if (option == GM_OPT_VARIANT) {
    if (value != 0 && value != 1) return gm_fail(c, GM_ERR_ARG, "variant must be 0 or 1");
    c->variant = value;
    return GM_OK;
  }
  if (option == GM_OPT_PERSISTENT) {
    if (value != 0 && value != 1) return gm_fail(c, GM_ERR_ARG, "persistent must be 0 or 1");
    c->persistent = value;
    return GM_OK;
  }
  if (option == GM_OPT_WAVES) {
    if (value != 1 && value != 2) return gm_fail(c, GM_ERR_ARG, "waves must be 1 or 2");
    c->waves = value;
    return GM_OK;
  }
  if (option == GM_OPT_SPH_SPLIT) {
    if (value != 1 && value != 2 && value != 4) return gm_fail(c, GM_ERR_ARG, "split must be 1, 2 or 4");
    c->sph_split = value;
    return GM_OK;
  }
  if (option == GM_OPT_PROPOSE_EVERY) {
    if (value < 1 || value > 1000) return gm_fail(c, GM_ERR_ARG, "proposal interval must be in [1, 1000]");
    c->prop_every = value;
    c->prop.every = value;
    return GM_OK;
  }
  if (option == GM_OPT_ORDER_TARGET) {
    if (value < 3 || value > 7) return gm_fail(c, GM_ERR_ARG, "target try count must be in [3, 7]");
    c->order_target = value;
    return GM_OK;
  }
  if (option == GM_OPT_ADJOINT_PAIR) {
    if (value != 0 && value != 1) return gm_fail(c, GM_ERR_ARG, "adjoint pair must be 0 or 1");
    c->adj_pair = value;
    return GM_OK;
  }
  if (option == GM_OPT_GRAPH) {
    if (value != 0 && value != 1) return gm_fail(c, GM_ERR_ARG, "graph must be 0 or 1");
    c->graph = value;
    return GM_OK;
  }
  if (option == GM_OPT_UNIT) {
    // the set-up, adjoint and assembly kernels cover one work unit with one 256-thread block
    if (value < 64 || value > 256) return gm_fail(c, GM_ERR_ARG, "unit size must be in [64, 256]");
    c->unit = value;
    return GM_OK;
  }
  return gm_fail(c, GM_ERR_ARG, "unknown option %d", option);
}

extern "C" int gm_last_error(gm_ctx* c, char* buf, int len) {
  if (!buf || len <= 0) return GM_ERR_ARG;
  const std::string& s = c ? c->err : g_create_err;
  snprintf(buf, (size_t)len, "%s", s.c_str());
  return GM_OK;
}

extern "C" int gm_last_timing(gm_ctx* c, double* ms) {
  if (!c || !ms) return GM_ERR_ARG;
  for (int i = 0; i < 6; ++i) ms[i] = c->timing[i];
  return GM_OK;
}

// Structural checks of a node earth model before anything indexes it: the
// spline kernel and the knot search of every evaluation trust iktop, the
// layer flags and the knot radii.  Layer 1 (the central sphere or halfspace,
// constant) needs one knot, every other layer at least two with strictly
// increasing radii (splineEarthmodelCoefficients.f90:59-94).
static int check_model(gm_ctx* c, int nk, const double* rk, const double* ro, const double* vpv, const double* vph,
                       const double* vsv, const double* vsh, const double* eta, const double* qkinv,
                       const double* qminv, int nlay, const int* iktop, const double* rb, const int* iflso, int nic,
                       int noc, int equidis, double rearth) {
  if (!rk || !ro || !vpv || !vph || !vsv || !vsh || !eta || !qkinv || !qminv || !iktop || !rb || !iflso)
    return gm_fail(c, GM_ERR_ARG, "model arrays must not be NULL");
  if (equidis != 0 && equidis != 1) return gm_fail(c, GM_ERR_ARG, "equidistant flag must be 0 or 1");
  if (!(rearth > 0.0) || !std::isfinite(rearth)) return gm_fail(c, GM_ERR_ARG, "earth radius must be positive");
  if (nic < 0 || noc < nic || noc > nlay)
    return gm_fail(c, GM_ERR_ARG, "layer indices must satisfy 0 <= nic=%d <= noc=%d <= nlay=%d", nic, noc, nlay);
  for (int i = 0; i < nk; ++i) {
    const double v[9] = {rk[i], ro[i], vpv[i], vph[i], vsv[i], vsh[i], eta[i], qkinv[i], qminv[i]};
    for (double x : v)
      if (!std::isfinite(x)) return gm_fail(c, GM_ERR_ARG, "knot %d holds a non-finite value", i + 1);
    if (!(ro[i] > 0.0)) return gm_fail(c, GM_ERR_ARG, "knot %d: density must be positive", i + 1);
  }
  if (iktop[0] < 1) return gm_fail(c, GM_ERR_ARG, "iktop(1)=%d must be >= 1", iktop[0]);
  for (int nl = 1; nl < nlay; ++nl)
    if (iktop[nl] - iktop[nl - 1] < 2)
      return gm_fail(c, GM_ERR_ARG, "layer %d needs at least two knots (iktop %d -> %d)", nl + 1, iktop[nl - 1],
                     iktop[nl]);
  if (iktop[nlay - 1] != nk) return gm_fail(c, GM_ERR_ARG, "iktop(nlay)=%d differs from nk=%d", iktop[nlay - 1], nk);
  for (int nl = 1; nl < nlay; ++nl)
    for (int j = iktop[nl - 1] + 1; j < iktop[nl]; ++j)      // 1-based knots of layer nl + 1
      if (!(rk[j] > rk[j - 1]))
        return gm_fail(c, GM_ERR_ARG, "knot radii of layer %d must increase (knot %d)", nl + 1, j + 1);
  for (int nl = 0; nl < nlay; ++nl) {
    if (iflso[nl] != 0 && iflso[nl] != 1) return gm_fail(c, GM_ERR_ARG, "isfluid of layer %d must be 0 or 1", nl + 1);
    if (!(rb[nl] > 0.0) || (nl > 0 && !(rb[nl] > rb[nl - 1])))
      return gm_fail(c, GM_ERR_ARG, "layer top radii must be positive and increase (layer %d)", nl + 1);
    const double rt = rk[iktop[nl] - 1];
    if (std::fabs(rb[nl] - rt) > 1e-9 * rearth)
      return gm_fail(c, GM_ERR_ARG, "top radius of layer %d (%g) differs from its top knot (%g)", nl + 1, rb[nl], rt);
  }
  if (noc > 0 && iflso[noc - 1] == 1 && noc == nlay)
    return gm_fail(c, GM_ERR_ARG, "the fluid outer core (layer %d) must lie below another layer", noc);
  return GM_OK;
}

extern "C" int gm_model_set(gm_ctx* c, int nk, const double* rk, const double* ro, const double* vpv, const double* vph,
                            const double* vsv, const double* vsh, const double* eta, const double* qkinv,
                            const double* qminv, int nlay, const int* iktop, const double* rb, const int* iflso,
                            int nic, int noc, int fsflag, int equidis, double rearth, double ric, double roc,
                            double fref) {
  if (!c) return GM_ERR_ARG;
  if (nk < 2 || nk > GM_MAXK) return gm_fail(c, GM_ERR_ARG, "nk=%d outside [2,%d]", nk, GM_MAXK);
  if (fsflag != 0 && fsflag != 1) return gm_fail(c, GM_ERR_ARG, "full_space flag must be 0 or 1");
  if (nlay < 2 || nlay > GM_MAXL) return gm_fail(c, GM_ERR_ARG, "nlay=%d outside [2,%d]", nlay, GM_MAXL);
  {
    const int rc = check_model(c, nk, rk, ro, vpv, vph, vsv, vsh, eta, qkinv, qminv, nlay, iktop, rb, iflso, nic, noc,
                               equidis, rearth);
    if (rc) { c->has_model = false; return rc; }
  }
  // the minors kernels stage one frequency's spline table image in LDS
  if (gmLdsBytes(nk, nlay) > (size_t)c->max_lds)
    return gm_fail(c, GM_ERR_UNSUPPORTED, "nk=%d: spline table image of %zu B exceeds the %d B of LDS per block", nk,
                   gmLdsBytes(nk, nlay), c->max_lds);
  GmModel& m = c->hmodel;
  int nnod = m.nnod;
  double rn[GM_MAXN + 1]; int stt[GM_MAXN + 1], lay[GM_MAXN + 1];
  memcpy(rn, m.rnod, sizeof rn); memcpy(stt, m.state, sizeof stt); memcpy(lay, m.layer, sizeof lay);
  memset(&m, 0, sizeof m);
  c->has_model = false;                  // until the new model has passed every check
  m.nnod = nnod; memcpy(m.rnod, rn, sizeof rn); memcpy(m.state, stt, sizeof stt); memcpy(m.layer, lay, sizeof lay);
  m.nk = nk; m.nlay = nlay; m.nic = nic; m.noc = noc; m.fsflag = fsflag; m.equidis = equidis;
  m.rearth = rearth; m.ric = ric; m.roc = roc; m.fref = fref;
  for (int i = 0; i < nk; ++i) {
    m.rk[i + 1] = rk[i]; m.ro[i + 1] = ro[i]; m.vpv[i + 1] = vpv[i]; m.vph[i + 1] = vph[i]; m.vsv[i + 1] = vsv[i];
    m.vsh[i + 1] = vsh[i]; m.eta[i + 1] = eta[i]; m.qkinv[i + 1] = qkinv[i]; m.qminv[i + 1] = qminv[i];
  }
  for (int i = 0; i < nlay; ++i) { m.iktop[i + 1] = iktop[i]; m.rb[i + 1] = rb[i]; m.iflso[i + 1] = iflso[i]; }
  // nodeEarthmodel.f90:134-141
  if (fsflag == 1 && m.iflso[nlay] == 1) return gm_fail(c, GM_ERR_ARG, "Full space and ocean is not allowed");
  if (m.iktop[nlay] != nk) return gm_fail(c, GM_ERR_ARG, "iktop(nlay)=%d differs from nk=%d", m.iktop[nlay], nk);
  m.ikbot[1] = 0;
  for (int nl = 2; nl <= nlay; ++nl) m.ikbot[nl] = m.iktop[nl - 1] + 1;
  for (int nl = 1; nl <= nlay; ++nl) m.drk[nl] = -1.0;
  if (equidis == 1)
    for (int nl = 2; nl <= nlay; ++nl) m.drk[nl] = (m.rk[m.iktop[nl]] - m.rk[m.ikbot[nl]]) / (m.iktop[nl] - m.ikbot[nl]);
  m.hascore = (noc > 0 && m.iflso[noc] == 1) ? 1 : 0;
  m.hasocean = m.iflso[nlay] == 1 ? 1 : 0;
  m.ruc = m.hasocean ? m.rb[nlay - 1] : rearth;
  m.rhs = m.rb[1];
  c->has_model = true;
  return GM_OK;
}

extern "C" int gm_nodes_set(gm_ctx* c, int nnod, const double* rnod, const int* state, const int* layer) {
  if (!c) return GM_ERR_ARG;
  if (nnod < 1 || nnod > GM_MAXN) return gm_fail(c, GM_ERR_ARG, "nnod=%d outside [1,%d]", nnod, GM_MAXN);
  if (!rnod || !state || !layer) return gm_fail(c, GM_ERR_ARG, "node arrays must not be NULL");
  for (int i = 0; i < nnod; ++i) {
    if (!std::isfinite(rnod[i]) || !(rnod[i] > 0.0) || (i > 0 && !(rnod[i] > rnod[i - 1])))
      return gm_fail(c, GM_ERR_ARG, "node radii must be positive and increase (node %d)", i + 1);
    if (state[i] != 0 && state[i] != 1) return gm_fail(c, GM_ERR_ARG, "state of node %d must be 0 or 1", i + 1);
    if (layer[i] < 1 || layer[i] > GM_MAXL) return gm_fail(c, GM_ERR_ARG, "layer of node %d outside [1,%d]", i + 1, GM_MAXL);
  }
  GmModel& m = c->hmodel;
  m.nnod = nnod;
  for (int i = 0; i < nnod; ++i) { m.rnod[i + 1] = rnod[i]; m.state[i + 1] = state[i]; m.layer[i + 1] = layer[i]; }
  c->has_nodes = true;
  return GM_OK;
}

// nodes and model may be set in either order: check the nodes against the
// model before a launch
static int check_nodes(gm_ctx* c) {
  const GmModel& m = c->hmodel;
  for (int k = 1; k <= m.nnod; ++k) {
    if (m.layer[k] < 1 || m.layer[k] > m.nlay)
      return gm_fail(c, GM_ERR_ARG, "layer %d of node %d outside the model's %d layers", m.layer[k], k, m.nlay);
    if (m.rnod[k] > m.rb[m.nlay] * (1.0 + 1e-12))
      return gm_fail(c, GM_ERR_ARG, "node %d at r=%g lies above the model top %g", k, m.rnod[k], m.rb[m.nlay]);
  }
  return GM_OK;
}

// grows the context's device scratch to at least `bytes`
static int grow_dbuf(gm_ctx* c, size_t bytes) {
  if (bytes <= c->dbuf_size) return GM_OK;
  if (c->dbuf) { HIPCHK(c, hipFree(c->dbuf)); c->dbuf = nullptr; c->dbuf_size = 0; }
  if (hipMalloc(&c->dbuf, bytes) != hipSuccess) return gm_fail(c, GM_ERR_NOMEM, "device allocation of %zu bytes failed", bytes);
  c->dbuf_size = bytes;
  return GM_OK;
}

static void derive_stats(const unsigned long long* hist, long long* st, int only_sys = -1) {
  static const int seq[8] = {0, 2, 4, 6, 8, 12, 16, 24};
  static const int nx[8] = {0, 3, 2, 4, 4, 4, 8, 8};
  long long MV[8] = {0}, SE[8] = {0}, NC[8] = {0};
  for (int k = 1; k <= 7; ++k) {
    MV[k] = MV[k - 1] + seq[k] + 1;
    SE[k] = SE[k - 1] + nx[k];
    NC[k] = NC[k - 1] + (k >= 3 ? k * (k - 1) / 2 : 0);
  }
  for (int s = 0; s < 5; ++s) {
    if (only_sys >= 0 && s != only_sys) continue;
    const unsigned long long* h = hist + s * 8;
    for (int k = 3; k <= 8; ++k) {           // k = 8 encodes rejected steps (7 tries)
      long long cnt = (long long)(k <= 7 ? h[k - 3] : h[5]);
      int kk = k <= 7 ? k : 7;
      if (!cnt) continue;
      st[3] += cnt * kk;
      st[4] += cnt * MV[kk];
      st[5] += cnt * SE[kk];
      st[6] += cnt * NC[kk];
      st[7] += cnt * (SE[kk] * kESYS[s] + MV[kk] * 8LL * kNNZ[s] + 4LL * kN[s] * MV[kk] + 8LL * kN[s] * NC[kk]);
      if (k <= 7) st[1] += cnt; else st[2] += cnt;
    }
    st[0] += (long long)h[6];
    st[9] += (long long)h[7];
  }
  if (only_sys < 0) st[8] += (long long)hist[5 * 8];
}

namespace {
struct Plan {
  int ninst = 0, nfreq = 0;
  std::vector<int> q, l;
  std::vector<double> omre;
  std::vector<int> uq, us, ue;      // work units (frequency slot, instance range)
};
}

// common driver of gm_minors_batch / gm_gfk_batch; with na (gm_gfk_batch_noalloc)
// every buffer lives in na's array, the plan is built on the device, outputs
// (status included) are device pointers, and nothing is allocated
static int run(gm_ctx* c, int nfreq, const double* freqs, const int* lmin, const int* nl, double dwn, double sigma,
               double eps, int attmode, int jsl, int jsr, int jre, int istyp, bool full, int flags,
               double* u_tor, double* d_tor, double* dettor, double* u_sph, double* d_sph, double* ulg, double* dlg,
               double* detsph, double* gf, float* gfre, float* gfim, int* status, long long* stats,
               const GmNoAlloc* na = nullptr, int nsnod = -1, const int* jsnod = nullptr,
               double* ujs_out = nullptr, double* ujt_out = nullptr) {
  if (!c) return GM_ERR_ARG;
  if (na && nfreq > GM_NOALLOC_MAXF)
    return gm_fail(c, GM_ERR_ARG, "%d frequencies: the batch without allocation takes at most %d", nfreq,
                   GM_NOALLOC_MAXF);
  if (!c->has_model || !c->has_nodes) return gm_fail(c, GM_ERR_ARG, "model and nodes must be set first");
  if (nfreq < 0 || (nfreq > 0 && (!freqs || !lmin || !nl))) return gm_fail(c, GM_ERR_ARG, "invalid frequency arguments");
  if (attmode < 0 || attmode > 3) return gm_fail(c, GM_ERR_ARG, "invalid attenuation mode %d", attmode);
  if (!(eps > 0.0)) return gm_fail(c, GM_ERR_ARG, "eps must be positive");
  GmModel& m = c->hmodel;
  const int nnod = m.nnod;
  if (jsl < 1 || jsl > nnod || jsr < 1 || jsr > nnod) return gm_fail(c, GM_ERR_NODEIDX, "Invalid index for source node");
  if (const int rc = check_nodes(c)) return rc;
  if (full) {
    if (istyp != 0 && istyp != 1) return gm_fail(c, GM_ERR_ARG, "istyp must be 0 (FORCE) or 1 (MOMENT)");
    // the receiver lies in the solid mantle or the ocean (unitJumpGreenFunctions.f90:119-123, 242),
    // the other nodes anywhere (inner core, fluid core, mantle, ocean)
    if (m.state[jre] != 0 && m.layer[jre] != m.nlay)
      return gm_fail(c, GM_ERR_UNSUPPORTED, "receiver node in the fluid core is not supported");
    if (m.hascore && !(m.rnod[jre] > m.roc))
      return gm_fail(c, GM_ERR_UNSUPPORTED, "receiver node below the core-mantle boundary is not supported");
  }
  // multi-source batch: every source node takes the receiver's place as the jump node of the adjoints
  const bool aski = nsnod >= 0;
  const int nsrc = aski ? nsnod + 1 : 1;
  if (aski)
    for (int js = 0; js < nsnod; ++js) {
      const int jsh = jsnod[js];
      // a source in the ocean is taken (computeGreenFKSpectraForASKI.f90:359-376), one in the fluid core not
      if (m.state[jsh] != 0 && m.layer[jsh] != m.nlay)
        return gm_fail(c, GM_ERR_UNSUPPORTED, "source node %d lies in the fluid core", jsh);
      if (m.hascore && !(m.rnod[jsh] > m.roc))
        return gm_fail(c, GM_ERR_UNSUPPORTED, "source node %d below the core-mantle boundary is not supported", jsh);
    }
  // nodes in a fluid layer or the inner core take the adjoint's end values, inner-core nodes a fifth job
  bool core_nodes = false, ic_nodes = false;
  for (int k = 1; k <= nnod; ++k) {
    if (m.state[k] == 1 || m.rnod[k] < m.ric) core_nodes = true;
    if (m.state[k] == 0 && m.rnod[k] < m.ric) ic_nodes = true;
  }
  // (omega, l) plan
  const int usz = (c->persistent && !na) ? c->unit : 256;   // one-lane-per-job kernels: one unit per block
  size_t ninst = 0;
  int nunits = 0;
  for (int f = 0; f < nfreq; ++f) {
    if (nl[f] < 0 || lmin[f] < 0) return gm_fail(c, GM_ERR_ARG, "negative degree range at frequency %d", f);
    if (dwn > 0.0 && nl[f] > 0 && lmin[f] < 1) return gm_fail(c, GM_ERR_ARG, "wavenumber indices are 1-based (frequency %d)", f);
    if (!(freqs[f] > 0.0)) return gm_fail(c, GM_ERR_ARG, "frequency %d must be positive", f);
    ninst += (size_t)nl[f];
    nunits += (nl[f] + usz - 1) / usz;
  }
  if (ninst == 0) return GM_OK;     // no (omega, l) instance: every output is empty, stats unchanged
  if (ninst > (size_t)INT32_MAX) return gm_fail(c, GM_ERR_ARG, "more than 2^31 - 1 instances in one batch");
  Plan P;
  P.nfreq = nfreq;
  P.ninst = (int)ninst;
  if (!na) {
    P.omre.resize(nfreq);
    for (int f = 0; f < nfreq; ++f) {
      P.omre[f] = 2.0 * (double)3.1415927410125732 * freqs[f];
      for (int i = 0; i < nl[f]; ++i) { P.q.push_back(f); P.l.push_back(lmin[f] + i); }
    }
    // work units: degree ranges of one frequency, highest frequency first and
    // low degrees (the longest integrations) first within a frequency
    std::vector<int> off(nfreq + 1, 0);
    for (int f = 0; f < nfreq; ++f) off[f + 1] = off[f] + nl[f];
    std::vector<int> order(nfreq);
    for (int f = 0; f < nfreq; ++f) order[f] = f;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return freqs[a] > freqs[b]; });
    for (int f : order)
      for (int a = off[f]; a < off[f + 1]; a += usz) {
        P.uq.push_back(f); P.us.push_back(a); P.ue.push_back(std::min(a + usz, off[f + 1]));
      }
  }
  const size_t n = P.ninst;
  // spectra per node: U, V, W (moment 10, force 5), U, V for a receiver in the ocean (8, 4);
  // with stresses U, R, V, S, W, T (20, 10; ocean U, R, V and zero S: 16, 8);
  // multi-source: U, V, dU/dr, dV/dr, W, dW/dr (20, 10)
  const bool ocean = full && m.state[jre] == 1;
  const bool stress = full && !aski && (flags & GM_FLAG_STRESS_SPECTRA) != 0;
  const int njsph = istyp == 1 ? 4 : 2, njtor = istyp == 1 ? 2 : 1, nmask = stress ? 4 : 2;
  const int nsp = aski ? 2 * (2 * njsph + njtor) : nmask * njsph + (ocean ? 0 : (nmask / 2) * njtor);
  const size_t nout = (size_t)nsrc * nsp * nnod * n;
  HIPCHK(c, hipSetDevice(c->device));
  // device buffer layout
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
  size_t o_model = take(sizeof(GmModel));
  size_t o_freqs = take(sizeof(double) * (nfreq + 1));
  size_t o_omre = take(sizeof(double) * (nfreq + 1));
  size_t o_tables = take(sizeof(GmKnot) * (size_t)(m.nk + 1) * (nfreq + 1));
  const size_t lstride = gmImageBytes<GmKnotL>(m.nk, m.nlay), tstride = gmImageBytes<GmKnotT>(m.nk, m.nlay);
  size_t o_ltab = take(lstride * (size_t)(nfreq + 1));
  size_t o_ttab = take(tstride * (size_t)(nfreq + 1));
  size_t o_q = take(sizeof(int) * (n + 1)), o_l = take(sizeof(int) * (n + 1)), o_status = take(sizeof(int) * (n + 1));
  size_t o_rs = take(sizeof(double) * 2 * (n + 1));
  size_t o_ut = take(16 * 2 * nnod * n + 16), o_dt = take(16 * 2 * nnod * n + 16), o_dett = take(16 * nnod * n + 16);
  size_t o_us = take(16 * 6 * nnod * n + 16), o_ds = take(16 * 6 * nnod * n + 16);
  size_t o_ulg = take(8 * nnod * n + 8), o_dlg = take(8 * nnod * n + 8), o_dets = take(16 * nnod * n + 16);
  size_t o_ujs = full ? take(16 * 16 * nnod * n + 16) : 0;
  size_t o_adje = (full && core_nodes) ? take(16 * 8 * n + 16) : 0;
  bool dev = (flags & GM_FLAG_DEVICE_PTRS) != 0;
  size_t o_gf = (full && (gf && !dev)) ? take(16 * nout + 16) : 0;
  size_t o_gfre = (full && (gfre && !dev)) ? take(4 * nout + 4) : 0;
  size_t o_gfim = (full && (gfim && !dev)) ? take(4 * nout + 4) : 0;
  // unit-jump Green functions out (gm_ujgf_batch): [nnod][4][4][k] and [nnod][2][2][k] complex
  const bool ujout = full && ujs_out && ujt_out;
  size_t o_ujo = (ujout && !dev) ? take(16 * 16 * nnod * n + 16) : 0;
  size_t o_ujto = (ujout && !dev) ? take(16 * 4 * nnod * n + 16) : 0;
  size_t o_hist = take(sizeof(unsigned long long) * 6 * 8);
  size_t o_uq = take(sizeof(int) * (nunits + 1)), o_ust = take(sizeof(int) * (nunits + 1));
  size_t o_ue = take(sizeof(int) * (nunits + 1)), o_uc = take(sizeof(int) * GM_NCOUNTERS * nsrc);
  const int ctl_mode = c->log_max > 0 ? 1 : (c->prop_enabled == 1 ? 2 : 0);
  // per-instance policy: the proposer network's (enable 2) or a fixed target try count
  const bool pol_net = ctl_mode == 0 && c->prop_enabled == 2;
  const bool pol = pol_net || (ctl_mode == 0 && c->order_target != 6);
  size_t o_pol = pol ? take(sizeof(float) * 2 * n) : 0;
  size_t o_px = pol_net ? take(sizeof(float) * GM_MLP_IN * n) : 0;
  size_t o_pw = pol_net ? take(sizeof(float) * gm_mlp_forward_work((int)std::min(n, (size_t)GM_MLP_CHUNK))) : 0;
  size_t o_log = ctl_mode == 1 ? take(sizeof(float) * GM_LOGREC * (size_t)c->log_max) : 0;
  size_t o_logc = take(sizeof(int) * 4);
  size_t o_lmin = na ? take(sizeof(int) * (nfreq + 1)) : 0, o_nl = na ? take(sizeof(int) * (nfreq + 1)) : 0;
  size_t o_smax = na ? take(sizeof(int)) : 0;
  char* base;
  if (na) {
    if (off > na->bytes)
      return gm_fail(c, GM_ERR_ARG, "the batch needs %zu bytes, more than the %zu bytes of the batch without allocation",
                     off, na->bytes);
    base = (char*)na->base;
  } else {
    if (const int rc = grow_dbuf(c, off)) return rc;
    base = (char*)c->dbuf;
  }
  hipStream_t s = c->stream;
  GmModel* dM = (GmModel*)(base + o_model);
  HIPCHK(c, hipMemcpyAsync(dM, &m, sizeof(GmModel), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + o_freqs, freqs, sizeof(double) * nfreq, hipMemcpyHostToDevice, s));
  if (na) {
    HIPCHK(c, hipMemcpyAsync(base + o_lmin, lmin, sizeof(int) * nfreq, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(base + o_nl, nl, sizeof(int) * nfreq, hipMemcpyHostToDevice, s));
    gm_noalloc_plan(s, nfreq, (const double*)(base + o_freqs), (const int*)(base + o_lmin), (const int*)(base + o_nl),
                    (int*)(base + o_q), (int*)(base + o_l), (double*)(base + o_omre), (int*)(base + o_uq),
                    (int*)(base + o_ust), (int*)(base + o_ue), (unsigned long long*)(base + o_hist),
                    (int*)(base + o_logc), (int*)(base + o_smax));
  } else {
    HIPCHK(c, hipMemcpyAsync(base + o_omre, P.omre.data(), sizeof(double) * nfreq, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(base + o_q, P.q.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(base + o_l, P.l.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(base + o_uq, P.uq.data(), sizeof(int) * nunits, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(base + o_ust, P.us.data(), sizeof(int) * nunits, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(base + o_ue, P.ue.data(), sizeof(int) * nunits, hipMemcpyHostToDevice, s));
    std::vector<unsigned long long> zero(48, 0ULL);
    HIPCHK(c, hipMemcpyAsync(base + o_hist, zero.data(), sizeof(unsigned long long) * 48, hipMemcpyHostToDevice, s));
    // the work counters of the sources after the first (k_setup clears the first block)
    if (nsrc > 1) {
      std::vector<int> zc((size_t)GM_NCOUNTERS * (nsrc - 1), 0);
      HIPCHK(c, hipMemcpyAsync(base + o_uc + sizeof(int) * GM_NCOUNTERS, zc.data(), sizeof(int) * zc.size(),
                               hipMemcpyHostToDevice, s));
    }
  }
  GmBatch B;
  memset(&B, 0, sizeof B);
  B.n = (int)n; B.nnod = nnod; B.jsl = jsl; B.jsr = jsr; B.jre = jre; B.istyp = istyp;
  B.stress = stress ? 1 : 0;
  B.dwn = dwn;
  B.omim = -sigma; B.eps = eps;
  B.q = (const int*)(base + o_q); B.l = (const int*)(base + o_l); B.omre = (const double*)(base + o_omre);
  B.nkp1 = m.nk + 1; B.tables = (const GmKnot*)(base + o_tables);
  B.ltab = (const unsigned char*)(base + o_ltab); B.lstride = (int)lstride;
  B.ttab = (const unsigned char*)(base + o_ttab); B.tstride = (int)tstride;
  B.tor_split = 4;
  B.status = (na && status) ? status : (int*)(base + o_status); B.rs = (double*)(base + o_rs);
  B.ut = (double*)(base + o_ut); B.dt = (double*)(base + o_dt); B.dett = (double*)(base + o_dett);
  B.us = (double*)(base + o_us); B.ds = (double*)(base + o_ds); B.ulg = (double*)(base + o_ulg);
  B.dlg = (double*)(base + o_dlg); B.dets = (double*)(base + o_dets);
  B.ujs = full ? (double*)(base + o_ujs) : nullptr;
  B.adjend = (full && core_nodes) ? (double*)(base + o_adje) : nullptr;
  // adjoint jobs per instance: one per jump (two with GM_OPT_ADJOINT_PAIR), one for a receiver in the ocean
  B.adj_jumps = ocean ? 1 : (c->adj_pair ? 2 : 4);
  B.adj_jobs = B.adj_jumps + (ic_nodes ? 1 : 0);
  if (full) {
    B.gf = gf ? (dev ? gf : (double*)(base + o_gf)) : nullptr;
    B.gfre = gfre ? (dev ? gfre : (float*)(base + o_gfre)) : nullptr;
    B.gfim = gfim ? (dev ? gfim : (float*)(base + o_gfim)) : nullptr;
    if ((B.gfre == nullptr) != (B.gfim == nullptr)) return gm_fail(c, GM_ERR_ARG, "gfre and gfim must both be given");
    if (ujout) {
      B.ujs_out = dev ? ujs_out : (double*)(base + o_ujo);
      B.ujt_out = dev ? ujt_out : (double*)(base + o_ujto);
    }
  }
  B.hist = (unsigned long long*)(base + o_hist);
  B.nunits = nunits;
  B.unit_q = (const int*)(base + o_uq); B.unit_start = (const int*)(base + o_ust); B.unit_end = (const int*)(base + o_ue);
  B.unit_counter = (int*)(base + o_uc);
  B.ctl_mode = ctl_mode;
  B.sph_split = c->sph_split;
  B.mlp = c->dmlp;
  B.prop = c->prop;
  B.logbuf = ctl_mode == 1 ? (float*)(base + o_log) : nullptr;
  B.logcount = (int*)(base + o_logc);
  B.logmax = c->log_max;
  B.policy = pol ? (const float*)(base + o_pol) : nullptr;
  if (!na) {
    int z4[4] = {0, 0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(base + o_logc, z4, sizeof z4, hipMemcpyHostToDevice, s));
  }
  auto launch = c->variant == 0 ? gm_exact::launch_stage : gm_fast::launch_stage;
  auto bpc = c->variant == 0 ? gm_exact::persist_blocks_per_cu : gm_fast::persist_blocks_per_cu;
  const size_t lds = lstride;
  const int pgrid = std::max(1, std::min(nunits, c->ncu * bpc(lds, c->waves)));
  GmKnot* dT = (GmKnot*)(base + o_tables);
  const double* dF = (const double*)(base + o_freqs);
  // step logging / learned control / a policy / the batch without allocation: one lane per job
  const bool pers = c->persistent != 0 && ctl_mode == 0 && !pol && !na;
  // the kernel stages as one hipGraph (plain control, no policy): the uploads above and the copies
  // below stay on the stream; every stage clears what it accumulates with kernels of its own, so a
  // replay needs no memset node
  const bool use_graph = c->graph && ctl_mode == 0 && !pol && !na && !aski;
  if (use_graph) {
    struct Shape { int nfreq, attmode, pgrid, waves, variant, pers, full; size_t lds; const void* dM; const void* dF;
                   const void* dT; } sh;
    memset(&sh, 0, sizeof sh);
    sh.nfreq = nfreq; sh.attmode = attmode; sh.pgrid = pgrid; sh.waves = c->waves; sh.variant = c->variant;
    sh.pers = pers; sh.full = full; sh.lds = lds; sh.dM = dM; sh.dF = dF; sh.dT = dT;
    std::vector<unsigned char> key(sizeof B + sizeof sh);
    memcpy(key.data(), &B, sizeof B);
    memcpy(key.data() + sizeof B, &sh, sizeof sh);
    if (!c->g_exec || key != c->g_key) {
      if (c->g_exec) { HIPCHK(c, hipGraphExecDestroy(c->g_exec)); c->g_exec = nullptr; }
      if (c->g_graph) { HIPCHK(c, hipGraphDestroy(c->g_graph)); c->g_graph = nullptr; }
      c->g_key.clear();
      HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      launch(0, dM, B, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
      launch(1, dM, B, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
      launch(pers ? 6 : 2, dM, B, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
      launch(pers ? 7 : 3, dM, B, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
      if (full) {
        launch(4, dM, B, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
        launch(5, dM, B, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
      }
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(s, &g);
      if (ec != hipSuccess) return gm_fail(c, GM_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
      c->g_graph = g;
      HIPCHK(c, hipGraphInstantiate(&c->g_exec, c->g_graph, nullptr, nullptr, 0));
      c->g_key.swap(key);
      ++c->graph_captures;
    } else {
      ++c->graph_replays;
    }
    HIPCHK(c, hipEventRecord(c->ev[0], s));
    HIPCHK(c, hipGraphLaunch(c->g_exec, s));
    for (int k = 1; k <= 5; ++k) HIPCHK(c, hipEventRecord(c->ev[k], s));
  } else {
    HIPCHK(c, hipEventRecord(c->ev[0], s));
    launch(0, dM, B, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
    HIPCHK(c, hipEventRecord(c->ev[1], s));
    launch(1, dM, B, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
    if (pol) {
      auto lp = c->variant == 0 ? gm_exact::launch_policy : gm_fast::launch_policy;
      lp(dM, B, s, lds, pol_net ? (float*)(base + o_px) : nullptr, c->dmlp, pol_net ? (float*)(base + o_pw) : nullptr,
         (float*)(base + o_pol), m.rearth, pol_net ? 0 : c->order_target);
    }
    HIPCHK(c, hipEventRecord(c->ev[2], s));
    launch(pers ? 6 : 2, dM, B, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
    HIPCHK(c, hipEventRecord(c->ev[3], s));
    launch(pers ? 7 : 3, dM, B, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
    HIPCHK(c, hipEventRecord(c->ev[4], s));
    if (full && aski) {
      // per source: the adjoints of a jump at its node, then the direct Green functions and their
      // derivatives into its output block (the last source is a FORCE at the receiver node)
      for (int js = 0; js < nsrc; ++js) {
        GmBatch b = B;
        b.jre = js < nsnod ? jsnod[js] : jre;
        // one adjoint job for a jump in the ocean, else one per jump (two with GM_OPT_ADJOINT_PAIR)
        b.adj_jumps = m.state[b.jre] == 1 ? 1 : (c->adj_pair ? 2 : 4);
        b.adj_jobs = b.adj_jumps + (ic_nodes ? 1 : 0);
        b.unit_counter = B.unit_counter + (size_t)GM_NCOUNTERS * js;
        b.aski_js = js;
        b.aski_styp = js < nsnod ? istyp : 0;
        b.aski_nspmax = nsp;
        launch(4, dM, b, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
        launch(8, dM, b, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
      }
    } else if (full) {
      launch(4, dM, B, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
      launch(5, dM, B, nfreq, dF, attmode, dT, s, pgrid, lds, c->waves);
    }
    HIPCHK(c, hipEventRecord(c->ev[5], s));
  }
  HIPCHK(c, hipGetLastError());
  // results
  auto cp = [&](void* dst, size_t o, size_t bytes) -> hipError_t {
    if (!dst) return hipSuccess;
    if (dev) return hipMemcpyAsync(dst, base + o, bytes, hipMemcpyDeviceToDevice, s);
    return hipMemcpyAsync(dst, base + o, bytes, hipMemcpyDeviceToHost, s);
  };
  if (ujout && !dev) {
    HIPCHK(c, cp(ujs_out, o_ujo, 16 * 16 * nnod * n));
    HIPCHK(c, cp(ujt_out, o_ujto, 16 * 4 * nnod * n));
  }
  HIPCHK(c, cp(u_tor, o_ut, 16 * 2 * nnod * n));
  HIPCHK(c, cp(d_tor, o_dt, 16 * 2 * nnod * n));
  HIPCHK(c, cp(dettor, o_dett, 16 * nnod * n));
  HIPCHK(c, cp(u_sph, o_us, 16 * 6 * nnod * n));
  HIPCHK(c, cp(d_sph, o_ds, 16 * 6 * nnod * n));
  HIPCHK(c, cp(ulg, o_ulg, 8 * nnod * n));
  HIPCHK(c, cp(dlg, o_dlg, 8 * nnod * n));
  HIPCHK(c, cp(detsph, o_dets, 16 * nnod * n));
  if (full && !dev) {
    if (gf) HIPCHK(c, hipMemcpyAsync(gf, base + o_gf, 16 * nout, hipMemcpyDeviceToHost, s));
    if (gfre) HIPCHK(c, hipMemcpyAsync(gfre, base + o_gfre, 4 * nout, hipMemcpyDeviceToHost, s));
    if (gfim) HIPCHK(c, hipMemcpyAsync(gfim, base + o_gfim, 4 * nout, hipMemcpyDeviceToHost, s));
  }
  int first = 0;
  std::vector<int> hstatus;
  if (na) {
    gm_noalloc_status(s, (int)n, B.status, (int*)(base + o_smax));
    HIPCHK(c, hipMemcpyAsync(&first, base + o_smax, sizeof(int), hipMemcpyDeviceToHost, s));
  } else {
    hstatus.resize(n + 1);
    HIPCHK(c, hipMemcpyAsync(hstatus.data(), base + o_status, sizeof(int) * n, hipMemcpyDeviceToHost, s));
  }
  unsigned long long hist[48];
  HIPCHK(c, hipMemcpyAsync(hist, base + o_hist, sizeof hist, hipMemcpyDeviceToHost, s));
  if (ctl_mode == 1) {
    int cnt = 0;
    HIPCHK(c, hipMemcpyAsync(&cnt, base + o_logc, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    c->log_count = cnt;
    int ncopy = cnt < c->log_max ? cnt : c->log_max;
    if (ncopy > 0 && c->log_host)
      HIPCHK(c, hipMemcpyAsync(c->log_host, base + o_log, sizeof(float) * GM_LOGREC * (size_t)ncopy, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(c, hipStreamSynchronize(s));
  float ms;
  for (int i = 0; i < 5; ++i) { HIPCHK(c, hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1])); c->timing[i] = ms; }
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[5]));
  c->timing[5] = ms;
  memcpy(c->last_hist, hist, sizeof hist);
  if (status && !na) memcpy(status, hstatus.data(), sizeof(int) * n);
  if (stats) derive_stats(hist, stats);
  if (!na)
    for (size_t i = 0; i < n; ++i) if (hstatus[i]) { first = hstatus[i]; break; }
  if (first) return gm_fail(c, first, "at least one (omega,l) instance failed with error %d", first);
  return GM_OK;
}

int gm_run_batch(gm_ctx* c, int nfreq, const double* freqs, const int* lmin, const int* nl, double dwn, double sigma,
                 double eps, int attmode, int jre, int istyp, double* gf, float* gfre, float* gfim, int* status,
                 long long* stats, const GmNoAlloc* na) {
  return run(c, nfreq, freqs, lmin, nl, dwn, sigma, eps, attmode, jre, jre, jre, istyp, true, GM_FLAG_DEVICE_PTRS,
             nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, gf, gfre, gfim, status, stats, na);
}

extern "C" int gm_minors_batch(gm_ctx* c, int nfreq, const double* freqs, const int* lmin, const int* nl, double sigma,
                               double eps, int attmode, int jsl, int jsr, int flags, double* u_tor, double* d_tor,
                               double* dettor, double* u_sph, double* d_sph, double* ulg, double* dlg,
                               double* detsph, int* status, long long* stats) {
  return run(c, nfreq, freqs, lmin, nl, 0.0, sigma, eps, attmode, jsl, jsr, jsr, 1, false, flags, u_tor, d_tor, dettor,
             u_sph, d_sph, ulg, dlg, detsph, nullptr, nullptr, nullptr, status, stats);
}

extern "C" int gm_minors_batch_wn(gm_ctx* c, int nfreq, const double* freqs, const int* iwnmin, const int* nwn,
                                  double dwn, double sigma, double eps, int attmode, int jsl, int jsr, int flags,
                                  double* u_tor, double* d_tor, double* dettor, double* u_sph, double* d_sph,
                                  double* ulg, double* dlg, double* detsph, int* status, long long* stats) {
  if (!c) return GM_ERR_ARG;
  if (!(dwn > 0.0)) return gm_fail(c, GM_ERR_ARG, "dwn must be positive");
  return run(c, nfreq, freqs, iwnmin, nwn, dwn, sigma, eps, attmode, jsl, jsr, jsr, 1, false, flags, u_tor, d_tor,
             dettor, u_sph, d_sph, ulg, dlg, detsph, nullptr, nullptr, nullptr, status, stats);
}

extern "C" int gm_gfk_batch(gm_ctx* c, int nfreq, const double* freqs, const int* lmin, const int* nl, double sigma,
                            double eps, int attmode, int jre, int istyp, int flags, double* gf, float* gfre,
                            float* gfim, int* status, long long* stats) {
  return run(c, nfreq, freqs, lmin, nl, 0.0, sigma, eps, attmode, jre, jre, jre, istyp, true, flags, nullptr, nullptr,
             nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, gf, gfre, gfim, status, stats);
}

extern "C" int gm_ujgf_batch(gm_ctx* c, int nfreq, const double* freqs, const int* lmin, const int* nl, double sigma,
                             double eps, int attmode, int jre, int flags, double* ujs, double* ujt, int* status,
                             long long* stats) {
  if (!c) return GM_ERR_ARG;
  if (!ujs || !ujt) return gm_fail(c, GM_ERR_ARG, "ujs and ujt are required");
  return run(c, nfreq, freqs, lmin, nl, 0.0, sigma, eps, attmode, jre, jre, jre, 1, true, flags, nullptr, nullptr,
             nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, status, stats, nullptr,
             -1, nullptr, ujs, ujt);
}

extern "C" int gm_ujgf_batch_wn(gm_ctx* c, int nfreq, const double* freqs, const int* iwnmin, const int* nwn,
                                double dwn, double sigma, double eps, int attmode, int jre, int flags, double* ujs,
                                double* ujt, int* status, long long* stats) {
  if (!c) return GM_ERR_ARG;
  if (!(dwn > 0.0)) return gm_fail(c, GM_ERR_ARG, "dwn must be positive");
  if (!ujs || !ujt) return gm_fail(c, GM_ERR_ARG, "ujs and ujt are required");
  return run(c, nfreq, freqs, iwnmin, nwn, dwn, sigma, eps, attmode, jre, jre, jre, 1, true, flags, nullptr, nullptr,
             nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, status, stats, nullptr,
             -1, nullptr, ujs, ujt);
}

extern "C" int gm_gfk_batch_wn(gm_ctx* c, int nfreq, const double* freqs, const int* iwnmin, const int* nwn, double dwn,
                               double sigma, double eps, int attmode, int jre, int istyp, int flags, double* gf,
                               float* gfre, float* gfim, int* status, long long* stats) {
  if (!c) return GM_ERR_ARG;
  if (!(dwn > 0.0)) return gm_fail(c, GM_ERR_ARG, "dwn must be positive");
  return run(c, nfreq, freqs, iwnmin, nwn, dwn, sigma, eps, attmode, jre, jre, jre, istyp, true, flags, nullptr,
             nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, gf, gfre, gfim, status, stats);
}

extern "C" int gm_aski_batch(gm_ctx* c, int nfreq, const double* freqs, const int* lmin, const int* nl, double sigma,
                             double eps, int attmode, int nsnod, const int* jsnod, int jre, int istyp, int flags,
                             double* gf, float* gfre, float* gfim, int* status, long long* stats) {
  if (!c) return GM_ERR_ARG;
  if (!c->has_model || !c->has_nodes) return gm_fail(c, GM_ERR_ARG, "model and nodes must be set first");
  const int nnod = c->hmodel.nnod;
  if (nsnod < 1 || !jsnod) return gm_fail(c, GM_ERR_ARG, "at least one source node is required");
  if (jre < 1 || jre > nnod) return gm_fail(c, GM_ERR_NODEIDX, "Invalid index for receiver node");
  int jsl = nnod + 1, jsr = 0;
  bool jre_in = false;
  for (int js = 0; js < nsnod; ++js) {
    if (jsnod[js] < 1 || jsnod[js] > nnod) return gm_fail(c, GM_ERR_NODEIDX, "Invalid index for source node");
    jsl = std::min(jsl, jsnod[js]);
    jsr = std::max(jsr, jsnod[js]);
    if (jsnod[js] == jre) jre_in = true;
  }
  // computeGreenFKSpectraForASKI.f90:287-290
  if (!jre_in) return gm_fail(c, GM_ERR_ARG, "receiver node must also be a source node");
  return run(c, nfreq, freqs, lmin, nl, 0.0, sigma, eps, attmode, jsl, jsr, jre, istyp, true, flags, nullptr,
             nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, gf, gfre, gfim, status, stats, nullptr,
             nsnod, jsnod);
}

// elasticConstantsExternalRadialNodes (externalRadialNodes.f90:422-433) for
// any frequency and attenuation mode: spline table of f on the device, then
// its evaluation at every node.
extern "C" int gm_node_properties(gm_ctx* c, double f, int attmode, double* ro, double* elcon, double* qinv) {
  if (!c) return GM_ERR_ARG;
  if (!c->has_model || !c->has_nodes) return gm_fail(c, GM_ERR_ARG, "model and nodes must be set first");
  if (!ro || !elcon || !qinv) return gm_fail(c, GM_ERR_ARG, "output arrays must not be NULL");
  if (attmode < 0 || attmode > 3) return gm_fail(c, GM_ERR_ARG, "invalid attenuation mode %d", attmode);
  if (!(f > 0.0) || !std::isfinite(f)) return gm_fail(c, GM_ERR_ARG, "frequency must be positive");
  if (const int rc = check_nodes(c)) return rc;
  const GmModel& m = c->hmodel;
  const int nnod = m.nnod;
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
  const size_t lstride = gmImageBytes<GmKnotL>(m.nk, m.nlay), tstride = gmImageBytes<GmKnotT>(m.nk, m.nlay);
  const size_t o_model = take(sizeof(GmModel)), o_f = take(sizeof(double));
  const size_t o_tab = take(sizeof(GmKnot) * (size_t)(m.nk + 1));
  const size_t o_lt = take(lstride), o_tt = take(tstride);
  const size_t o_ro = take(sizeof(double) * nnod), o_el = take(sizeof(double) * 14 * nnod);
  const size_t o_q = take(sizeof(double) * 2 * nnod);
  HIPCHK(c, hipSetDevice(c->device));
  if (const int rc = grow_dbuf(c, off)) return rc;
  char* base = (char*)c->dbuf;
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(base + o_model, &m, sizeof(GmModel), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + o_f, &f, sizeof(double), hipMemcpyHostToDevice, s));
  auto launch = c->variant == 0 ? gm_exact::launch_node_props : gm_fast::launch_node_props;
  launch((const GmModel*)(base + o_model), (const double*)(base + o_f), attmode, (GmKnot*)(base + o_tab),
         (unsigned char*)(base + o_lt), (int)lstride, (unsigned char*)(base + o_tt), (int)tstride,
         (double*)(base + o_ro), (double*)(base + o_el), (double*)(base + o_q), s);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(ro, base + o_ro, sizeof(double) * nnod, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(elcon, base + o_el, sizeof(double) * 14 * nnod, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(qinv, base + o_q, sizeof(double) * 2 * nnod, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  return GM_OK;
}

extern "C" int gm_last_system_stats(gm_ctx* c, int sys, long long* stats) {
  if (!c || !stats || sys < 0 || sys > 4) return GM_ERR_ARG;
  for (int i = 0; i < GM_NSTAT; ++i) stats[i] = 0;
  derive_stats(c->last_hist, stats, sys);
  return GM_OK;
}

extern "C" int gm_set_proposer(gm_ctx* c, const float* params, const float* norm, float safety, int kcheck, int enable) {
  if (!c) return GM_ERR_ARG;
  if (!enable) { c->prop_enabled = 0; return GM_OK; }
  if (enable != 1 && enable != 2) return gm_fail(c, GM_ERR_ARG, "enable must be 0, 1 (every step) or 2 (per instance)");
  if (!params || !norm || !(safety > 0.0f)) return gm_fail(c, GM_ERR_ARG, "proposer needs parameters, normalisation and safety > 0");
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->dmlp && hipMalloc(&c->dmlp, sizeof(float) * GM_MLP_NPARAM) != hipSuccess)
    return gm_fail(c, GM_ERR_NOMEM, "proposer parameter allocation failed");
  HIPCHK(c, hipMemcpy(c->dmlp, params, sizeof(float) * GM_MLP_NPARAM, hipMemcpyHostToDevice));
  memset(&c->prop, 0, sizeof c->prop);
  for (int i = 0; i < GM_MLP_IN; ++i) {
    c->prop.mu[i] = norm[i];
    c->prop.isd[i] = norm[GM_MLP_IN + i] > 0.0f ? 1.0f / norm[GM_MLP_IN + i] : 1.0f;
  }
  for (int o = 0; o < GM_MLP_OUT; ++o) { c->prop.omu[o] = norm[2 * GM_MLP_IN + o]; c->prop.osd[o] = norm[2 * GM_MLP_IN + 2 + o]; }
  c->prop.safety = safety;
  c->prop.kcheck = (enable == 1 && kcheck) ? 1 : 0;
  c->prop.first_step = (enable == 2 && kcheck) ? 1 : 0;
  c->prop.every = c->prop_every;
  c->prop_enabled = enable;
  return GM_OK;
}

extern "C" int gm_collect_steplog(gm_ctx* c, int nfreq, const double* freqs, const int* lmin, const int* nl,
                                  double dwn, double sigma, double eps, int attmode, int jsl, int jsr, int maxrec,
                                  float* records, int* nrec, long long* stats) {
  if (!c || maxrec <= 0 || !records || dwn < 0.0) return GM_ERR_ARG;
  c->log_max = maxrec;
  c->log_host = records;
  c->log_count = 0;
  int rc = run(c, nfreq, freqs, lmin, nl, dwn, sigma, eps, attmode, jsl, jsr, jsr, 1, false, 0, nullptr, nullptr, nullptr,
               nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, stats);
  c->log_max = 0;
  c->log_host = nullptr;
  if (nrec) *nrec = c->log_count;
  return rc;
}
