// One engine variant of the CPU build: compiled with -DSIM_NS=<namespace>,
// -DSIM_FAST=0|1 (the library's exact or fast arithmetic) and
// -DSIM_LEARNED_FUSED=0|1 (learned control on the sequential or the fused
// step engine), -DSIM_ENTRY=<C entry point name>.
#include <algorithm>
#include <vector>
#include "gm_common.h"
extern "C" int g_sim_adj_pair;
extern "C" int g_sim_stress;    // sim_api.cpp: GM_FLAG_STRESS_SPECTRA of the columns   // sim_api.cpp: GM_OPT_ADJOINT_PAIR of the CPU build
extern "C" int g_sim_nsnod;      // sim_api.cpp: source nodes of a multi-source run (-1: none)
extern "C" int g_sim_jsnod[];
#define GM_NS SIM_NS
namespace SIM_NS {
static constexpr bool kFast = SIM_FAST != 0;
static constexpr bool kLearnedFused = SIM_LEARNED_FUSED != 0;
}
#include "gm_device.inc"
#include "gm_kernels.inc"
#include "sim_driver.inc"

extern "C" int SIM_ENTRY(const GmModel* m, int nfreq, const double* freqs, const int* lmin, const int* nl, double dwn,
                         double sigma, double eps, int attmode, int jsl, int jsr, int jre, int istyp, int full,
                         int ctl_mode, const float* mlp, const GmProposer* prop, int logmax, int pol_net, int kfixed,
                         double* ut, double* dt, double* dett, double* us, double* ds, double* ulg, double* dlg,
                         double* dets, double* gf, int* status, unsigned long long* hist, float* logbuf, int* logcount,
                         float* policy) {
  SIM_NS::SimOut o;
  int rc = SIM_NS::sim_run(*m, nfreq, freqs, lmin, nl, dwn, sigma, eps, attmode, jsl, jsr, jre, istyp, full != 0,
                           ctl_mode, mlp, prop, logmax, pol_net, kfixed, o);
  if (rc) return rc;
  auto cp = [](double* dst, const std::vector<double>& v) { if (dst && !v.empty()) std::memcpy(dst, v.data(), sizeof(double) * v.size()); };
  cp(ut, o.ut); cp(dt, o.dt); cp(dett, o.dett); cp(us, o.us); cp(ds, o.ds); cp(ulg, o.ulg); cp(dlg, o.dlg);
  cp(dets, o.dets); cp(gf, o.gf);
  if (status) std::memcpy(status, o.status.data(), sizeof(int) * o.status.size());
  if (hist) std::memcpy(hist, o.hist, sizeof o.hist);
  if (logcount) *logcount = o.logcount;
  if (policy && !o.policy.empty()) std::memcpy(policy, o.policy.data(), sizeof(float) * o.policy.size());
  if (logbuf && ctl_mode == 1) {
    const int nc = std::min(o.logcount, logmax);
    std::memcpy(logbuf, o.logbuf.data(), sizeof(float) * GM_LOGREC * (size_t)nc);
  }
  return 0;
}

// node properties for frequency f (gm_node_properties): the spline kernel on
// one frequency, then the node evaluation kernel
#define SIM_CAT2(a, b) a##b
#define SIM_CAT(a, b) SIM_CAT2(a, b)
extern "C" int SIM_CAT(SIM_ENTRY, _node_props)(const GmModel* m, double f, int attmode, double* ro, double* elcon,
                                               double* qinv) {
  using namespace SIM_NS;
  const size_t lstride = gmImageBytes<GmKnotL>(m->nk, m->nlay), tstride = gmImageBytes<GmKnotT>(m->nk, m->nlay);
  std::vector<GmKnot> tables((size_t)(m->nk + 1));
  std::vector<unsigned char> ltab(lstride), ttab(tstride);
  grid(1, [&] { k_spline(m, &f, attmode, tables.data(), ltab.data(), (int)lstride, ttab.data(), (int)tstride); });
  grid(1, [&] { k_node_props(m, tables.data(), ro, elcon, qinv); });
  return 0;
}
