// Model and proposer set-up of the CPU build of the engine: the same host-side
// derivations as gm_model_set, gm_nodes_set and gm_set_proposer (gm_capi.cpp),
// into heap objects the variant entry points take.
#include <cstring>
#include "gm_common.h"

dim3 threadIdx, blockIdx, blockDim, gridDim;

extern "C" GmModel* sim_model_new(int nk, const double* rk, const double* ro, const double* vpv, const double* vph,
                                  const double* vsv, const double* vsh, const double* eta, const double* qkinv,
                                  const double* qminv, int nlay, const int* iktop, const double* rb, const int* iflso,
                                  int nic, int noc, int fsflag, int equidis, double rearth, double ric, double roc,
                                  double fref, int nnod, const double* rnod, const int* state, const int* layer) {
  if (nk < 2 || nk > GM_MAXK || nlay < 2 || nlay > GM_MAXL || nnod < 1 || nnod > GM_MAXN) return nullptr;
  GmModel* pm = new GmModel;
  GmModel& m = *pm;
  std::memset(&m, 0, sizeof m);
  m.nk = nk; m.nlay = nlay; m.nic = nic; m.noc = noc; m.fsflag = fsflag; m.equidis = equidis;
  m.rearth = rearth; m.ric = ric; m.roc = roc; m.fref = fref;
  for (int i = 0; i < nk; ++i) {
    m.rk[i + 1] = rk[i]; m.ro[i + 1] = ro[i]; m.vpv[i + 1] = vpv[i]; m.vph[i + 1] = vph[i]; m.vsv[i + 1] = vsv[i];
    m.vsh[i + 1] = vsh[i]; m.eta[i + 1] = eta[i]; m.qkinv[i + 1] = qkinv[i]; m.qminv[i + 1] = qminv[i];
  }
  for (int i = 0; i < nlay; ++i) { m.iktop[i + 1] = iktop[i]; m.rb[i + 1] = rb[i]; m.iflso[i + 1] = iflso[i]; }
  m.ikbot[1] = 0;
  for (int l = 2; l <= nlay; ++l) m.ikbot[l] = m.iktop[l - 1] + 1;
  for (int l = 1; l <= nlay; ++l) m.drk[l] = -1.0;
  if (equidis == 1)
    for (int l = 2; l <= nlay; ++l) m.drk[l] = (m.rk[m.iktop[l]] - m.rk[m.ikbot[l]]) / (m.iktop[l] - m.ikbot[l]);
  m.hascore = (noc > 0 && m.iflso[noc] == 1) ? 1 : 0;
  m.hasocean = m.iflso[nlay] == 1 ? 1 : 0;
  m.ruc = m.hasocean ? m.rb[nlay - 1] : rearth;
  m.rhs = m.rb[1];
  m.nnod = nnod;
  for (int i = 0; i < nnod; ++i) { m.rnod[i + 1] = rnod[i]; m.state[i + 1] = state[i]; m.layer[i + 1] = layer[i]; }
  return pm;
}

extern "C" void sim_model_free(GmModel* m) { delete m; }

extern "C" GmProposer* sim_proposer_new(const float* norm, float safety, int kcheck, int every) {
  GmProposer* p = new GmProposer;
  std::memset(p, 0, sizeof *p);
  for (int i = 0; i < GM_MLP_IN; ++i) {
    p->mu[i] = norm[i];
    p->isd[i] = norm[GM_MLP_IN + i] > 0.0f ? 1.0f / norm[GM_MLP_IN + i] : 1.0f;
  }
  for (int o = 0; o < GM_MLP_OUT; ++o) { p->omu[o] = norm[2 * GM_MLP_IN + o]; p->osd[o] = norm[2 * GM_MLP_IN + 2 + o]; }
  p->safety = safety;
  p->kcheck = kcheck ? 1 : 0;
  p->first_step = kcheck ? 1 : 0;
  p->every = every;
  return p;
}

extern "C" void sim_proposer_free(GmProposer* p) { delete p; }

extern "C" {
int g_sim_adj_pair = 0;
void sim_set_adjoint_pair(int v) { g_sim_adj_pair = v; }
int g_sim_stress = 0;
void sim_set_stress(int v) { g_sim_stress = v; }
// multi-source runs (gm_aski_batch): nsnod >= 1 source nodes, -1 for the receiver-only batch
int g_sim_nsnod = -1;
int g_sim_jsnod[GM_MAXN + 1];
void sim_set_sources(int nsnod, const int* jsnod) {
  g_sim_nsnod = nsnod;
  for (int i = 0; i < nsnod && i <= GM_MAXN; ++i) g_sim_jsnod[i] = jsnod[i];
}
}
