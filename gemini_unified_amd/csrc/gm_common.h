// Shared host/device definitions of the Green-FK engine for MI355X (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "gm_mlp.h"

#define GM_MAXK 512      // maximum earth-model knots
#define GM_MAXL 32       // maximum layers
#define GM_MAXN 128      // maximum external radial nodes

// error codes returned per (omega, l) instance and by the C ABI
enum {
  GM_OK = 0, GM_ERR_MINSTEP = 1, GM_ERR_MAXSTEP = 2, GM_ERR_NODEIDX = 3, GM_ERR_CF = 4,
  GM_ERR_RANGE = 5, GM_ERR_SPLINE = 6, GM_ERR_STATE = 7, GM_ERR_UNSUPPORTED = 8, GM_ERR_ARG = 9,
  GM_ERR_HIP = 10, GM_ERR_NOMEM = 11
};

// work counters kept by every kernel (same meaning as the oracle's)
enum {
  GM_ST_PROP = 0, GM_ST_ACC = 1, GM_ST_REJ = 2, GM_ST_MMP = 3, GM_ST_MATVEC = 4, GM_ST_SYSMAT = 5,
  GM_ST_NEVILLE = 6, GM_ST_FLOPS = 7, GM_ST_ZERODET = 8, GM_ST_SUBINT = 9, GM_NSTAT = 10
};

// systems of the radial problem
enum { GM_SYS_MINORS = 0, GM_SYS_TOROIDAL = 1, GM_SYS_FLUID = 2, GM_SYS_RADIAL = 3, GM_SYS_SOLIDT = 4 };

// Earth model + external nodes, 1-based arrays as in the reference
struct GmModel {
  int nk, nlay, nic, noc, fsflag, equidis, nnod, hascore;
  int hasocean, pad0, pad1, pad2;
  double rearth, ric, roc, fref, ruc, rhs;
  double rk[GM_MAXK + 1];
  double ro[GM_MAXK + 1], vpv[GM_MAXK + 1], vph[GM_MAXK + 1], vsv[GM_MAXK + 1], vsh[GM_MAXK + 1];
  double eta[GM_MAXK + 1], qkinv[GM_MAXK + 1], qminv[GM_MAXK + 1];
  double rb[GM_MAXL + 1], drk[GM_MAXL + 1];
  int iktop[GM_MAXL + 1], ikbot[GM_MAXL + 1], iflso[GM_MAXL + 1], pad3;
  double rnod[GM_MAXN + 1];
  int state[GM_MAXN + 1], layer[GM_MAXN + 1];
};

// spline table entry for one knot and one frequency: density, its spline
// coefficient, 7 complex moduli (A,C,F,L,N,kappa,mu) and their coefficients
struct GmKnot {
  double ro, ro2;
  double er[7], ei[7];
  double e2r[7], e2i[7];
  double rh;     // 1 / (rk(j+1) - rk(j)) inside the layer (0 at a layer top)
  // 272-byte records: a multiple of 16 bytes, and consecutive knots start 4
  // LDS banks apart, so lanes reading one field of different knots spread
  // over 8 bank groups instead of all hitting the same banks
  double pad[3];
};
static_assert(sizeof(GmKnot) == 272, "GmKnot record size");

// Compact record of one knot staged in LDS: density and the moduli 1..5
// (A, C, F, L, N) with their spline coefficients and 1/h, the set the minors
// systems evaluate at every grid point.  184 bytes: consecutive records start
// 14 banks apart (16 distinct bank offsets), and a 200-knot table fits four
// times into one CU's LDS.  Moduli 6 and 7 (kappa, mu: start radius and
// initial values only) are read from the full table in global memory.
struct GmKnotL {
  static constexpr int kFirst = 0, kCount = 5;   // moduli held: kFirst .. kFirst + kCount - 1 (0-based)
  double ro, ro2, rh;
  double er[5], ei[5], e2r[5], e2i[5];
};
static_assert(sizeof(GmKnotL) == 184, "GmKnotL record size");
// The toroidal kernel's record: density and the moduli 4, 5 (L, N) only.
// 88 bytes (22 banks apart): a 200-knot image of 19 KB lets eight one-wave
// blocks, two per SIMD, share a CU.
struct GmKnotT {
  static constexpr int kFirst = 3, kCount = 2;
  double ro, ro2, rh;
  double er[2], ei[2], e2r[2], e2i[2];
};
static_assert(sizeof(GmKnotT) == 88, "GmKnotT record size");

// LDS image of one frequency's table, staged by every block of the minors
// kernels (and laid out the same in global memory, one image per frequency):
//   R rec[nk+1] | double rk[nk+1] | double drk[nlay+1] | int iktop[nlay+1]
// so that the knot lookup of a spline evaluation touches LDS only.
template <class R>
__host__ __device__ inline size_t gmImageBytes(int nk, int nlay) {
  size_t b = (size_t)(nk + 1) * (sizeof(R) + sizeof(double)) + (size_t)(nlay + 1) * (sizeof(double) + sizeof(int));
  return (b + 15) & ~(size_t)15;
}
__host__ __device__ inline size_t gmLdsBytes(int nk, int nlay) { return gmImageBytes<GmKnotL>(nk, nlay); }

// one batch of (omega, l) instances and its device buffers
struct GmBatch {
  int n;                           // number of (omega, l) instances
  int nnod, jsl, jsr, jre, istyp;
  int stress;                      // 1: FK columns with stresses, U, R, V, S and W, T (STRESSFLAG)
  double omim, eps;
  const int* __restrict__ q;       // frequency slot per instance
  const int* __restrict__ l;       // harmonic degree (dwn == 0) or 1-based wavenumber index (dwn > 0)
  double dwn;                      // wavenumber spacing of the regional mode, 0 for harmonic degrees
  const double* __restrict__ omre; // per frequency slot
  int nkp1;                        // table stride (nk + 1)
  const GmKnot* __restrict__ tables;
  const unsigned char* __restrict__ ltab;  // LDS images (GmKnotL), one per frequency slot, lstride bytes apart
  int lstride;
  const unsigned char* __restrict__ ttab;  // toroidal LDS images (GmKnotT), tstride bytes apart
  int tstride;
  int tor_split;                           // blocks per work unit of the toroidal kernel
  int* __restrict__ status;
  double* __restrict__ rs;         // [2][n]: toroidal, spheroidal start radius
  double* __restrict__ ut;         // toroidal up   [nnod][2][n] complex (re,im interleaved)
  double* __restrict__ dt;         // toroidal down [nnod][2][n]
  double* __restrict__ dett;       // [nnod][n] complex
  double* __restrict__ us;         // spheroidal up   [nnod][6][n]
  double* __restrict__ ds;         // spheroidal down [nnod][6][n]
  double* __restrict__ ulg;        // [nnod][n]
  double* __restrict__ dlg;        // [nnod][n]
  double* __restrict__ dets;       // [nnod][n]
  double* __restrict__ ujs;        // spheroidal UJGF [nnod][4 jumps][4 comps][n] complex
  // third component of the adjoint at the end of each leg [2 legs][4 jumps][n]
  // complex (leg 0: at the ocean bottom or surface, leg 1: at the CMB, start
  // radius or lowest node; null when no node lies in a fluid layer or the inner core)
  double* __restrict__ adjend;
  int adj_jobs;                    // adjoint jobs per instance: adj_jumps, + 1 inner-core propagator
  int adj_jumps;                   // 4: one adjoint per unit jump (the reference's); 2: the two
                                   // adjoints whose combinations give the four (GM_OPT_ADJOINT_PAIR)
  double* __restrict__ gf;         // FK column, complex [nnod][nsp][n] (may be null)
  float* __restrict__ gfre;        // FP32 real part [nnod][nsp][n] (may be null)
  float* __restrict__ gfim;
  unsigned long long* __restrict__ hist;   // [6][8]: 5 systems + misc counters
  // work units of the persistent kernels: instance ranges of one frequency
  int nunits;
  const int* __restrict__ unit_q;
  const int* __restrict__ unit_start;
  const int* __restrict__ unit_end;
  int* __restrict__ unit_counter;          // GM_NCOUNTERS work counters, zeroed by k_setup
  // step control of the minors system: 0 plain (reference), 1 plain with step
  // logging, 2 learned proposer
  int sph_split;                           // blocks per work unit of the spheroidal kernel (1, 2 or 4)
  int ctl_mode;
  const float* __restrict__ mlp;           // proposer parameters (GM_MLP_NPARAM floats)
  float* __restrict__ logbuf;              // step log records [logmax][GM_LOGREC]
  int* __restrict__ logcount;
  int logmax;
  GmProposer prop;
  // per-instance step policy [n][2] (null: the reference's control): the
  // target try count of the x1.25 / x0.8 rule (complexBulirschStep.f90:346-359
  // targets 6) and the first step of the spheroidal up pass in metres (<= 0:
  // the start value's lambda/4); read by the toroidal and spheroidal kernels
  const float* __restrict__ policy;
  // multi-source assembly (k_aski_assemble): output block of this source, its
  // source type and the spectra per node of the output layout
  int aski_js, aski_styp, aski_nspmax;
  // unit-jump Green functions out instead of FK columns (k_assemble, gm_ujgf_batch):
  // complex [nnod][jump 4][component 4][n] and [nnod][jump 2][component 2][n]
  double* __restrict__ ujs_out;
  double* __restrict__ ujt_out;
};
#define GM_LOGREC 20
// work counters (GmBatch::unit_counter): persistent toroidal / spheroidal
// kernels, and the block claims of the one-lane-per-job kernels
enum { GM_CNT_TOR_PERSIST = 0, GM_CNT_SPH_PERSIST = 1, GM_CNT_TOR = 2, GM_CNT_SPH = 3, GM_CNT_ADJ = 4,
       GM_NCOUNTERS = 5 };
