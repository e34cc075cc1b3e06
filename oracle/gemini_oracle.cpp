// ============================================================================
//  CPU ORACLE — test infrastructure only, never part of the product path.
//
//  A scalar FP64 restatement of GEMINI_UNIFIED's Green-FK hot path, written
//  to follow the reference Fortran operation by operation so that the HIP
//  kernels of gemini_unified_amd can be checked against it.  Only tests/,
//  __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
//
//  Restated reference files (paths relative to the reference root):
//    episode/parametersBulirschStep.f90:28-48      step-control constants
//    episode/complexBulirschStep.f90:111-359       BS step, MMP, Neville, adjust
//    episode/integrationStep.f90:98-190            step loop per sub-interval
//    episode/propagateOde.f90:132-331              node/disco stops, lg-norm
//    f90/geminiIntegrationEnvironment.f90:59-203   nodes / discontinuities
//    f90/locatePoint.f90:80-110                    bisection locate
//    f90/nodeEarthmodel.f90:287-330,599-694        moduli at knots, layer queries
//    f90/complexElasticConstants.f90:48-106        complex VTI moduli
//    f90/cubicSpline.f90:74-109,176-220            natural cubic spline (+dgtsv)
//    f90/splineEarthmodelCoefficients.f90:46-188   spline tables and evaluation
//    f90/{minors,toroidal,fluid,radial,solidt}OdeSystem.f90   system matrices
//    f90/initialValues.f90:71-574                  start radius, start values
//    f90/toroidalMinors.f90:71-258                 toroidal minors driver
//    f90/spheroidalMinors.f90:72-585               spheroidal minors driver
//    f90/spheroidalAdjoint.f90:53-159              adjoint propagation
//    f90/unitJumpGreenFunctions.f90:62-380         unit-jump Green functions
//    f90/reciprocity.f90:82-180                    reciprocity relations
//    f90/sourceTerms.f90:59-118                    source jump vectors
//    f90/computeGreenFKSpectraForSynthetics.f90:186-236,371-494  driver
//    f90/externalRadialNodes.f90:337-418,538-547   node splitting / lookup
//
//  Arithmetic conventions follow gfortran on x86-64 without FMA:
//   * complex * complex  = (ac-bd, ad+bc); complex * real is componentwise;
//   * complex / complex uses Smith's range reduction (gfortran's default
//     "wide" complex division); complex / real is componentwise;
//   * abs(complex) = hypot; sqrt(complex) = libm csqrt;
//   * literals written without a kind suffix in the Fortran are FP32 values
//     (pi = 3.1415927410125732, 0.8 = 0.800000011920929, 0.1 = 0.100000001490116).
//  The file is compiled with -ffp-contract=off so no multiply-add is fused.
//
//  Indices are kept 1-based (element 0 unused) so that every loop reads like
//  the Fortran it restates.
// ============================================================================
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <string>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ----------------------------------------------------------------------------
//  Constants (mathConstants.f90:26-35, parametersBulirschStep.f90:28-48)
// ----------------------------------------------------------------------------
static const double MC_PID = (double)3.141592653589793f;      // FP32 literal in a double parameter
static const double MC_TWO_PID = 2.0 * MC_PID;
static const double GROW = 1.25;
static const double SHRINK = (double)0.8f;
static const int MAXTRY = 7, MINTRY = 3, OPTTRY = 6, KGV = 48;
static const int SEQDIV[8] = {0, 2, 4, 6, 8, 12, 16, 24};
static const int NXDIV[8] = {0, 3, 2, 4, 4, 4, 8, 8};
static const int IXDIV[8][9] = {
    {0},
    {0, 0, 24, 48, 0, 0, 0, 0, 0},
    {0, 12, 36, 0, 0, 0, 0, 0, 0},
    {0, 8, 16, 32, 40, 0, 0, 0, 0},
    {0, 6, 18, 30, 42, 0, 0, 0, 0},
    {0, 4, 20, 28, 44, 0, 0, 0, 0},
    {0, 3, 9, 15, 21, 27, 33, 39, 45},
    {0, 2, 10, 14, 22, 26, 34, 38, 46}};
static const double MINSIZE = 1.e-12;          // integrationStep.f90:35
static const int MAXNUM = 200000;              // integrationStep.f90:36
static const double MAXY = 1.e10;              // propagateOde.f90:55
static const double DEPS = 2.220446049250313e-16;   // epsilon(1.d0)
static const double FEPS = 1.1920928955078125e-07;  // epsilon(1.0)
static const double TINYD = 2.2250738585072014e-308; // tiny(1.d0)

// algorithmic flop constants per system-matrix evaluation (spline + assembly)
enum { SYS_MINORS = 0, SYS_TOROIDAL = 1, SYS_FLUID = 2, SYS_RADIAL = 3, SYS_SOLIDT = 4 };
static const int NNZ[5] = {19, 4, 3, 4, 14};
static const int ESYS[5] = {219, 148, 156, 196, 206};

// error codes (level-2 error_message conditions)
enum {
  OK = 0, ERR_MINSTEP = 1, ERR_MAXSTEP = 2, ERR_NODEIDX = 3, ERR_CF = 4,
  ERR_RANGE = 5, ERR_SPLINE = 6, ERR_STATE = 7, ERR_UNSUPPORTED = 8, ERR_ARG = 9
};

// ----------------------------------------------------------------------------
//  Complex arithmetic with gfortran semantics
// ----------------------------------------------------------------------------
struct cx { double re, im; };
static inline cx C(double r, double i = 0.0) { cx z; z.re = r; z.im = i; return z; }
static inline cx add(cx a, cx b) { return C(a.re + b.re, a.im + b.im); }
static inline cx sub(cx a, cx b) { return C(a.re - b.re, a.im - b.im); }
static inline cx neg(cx a) { return C(-a.re, -a.im); }
static inline cx mul(cx a, cx b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cx rmul(double s, cx a) { return C(s * a.re, s * a.im); }
static inline cx mulr(cx a, double s) { return C(a.re * s, a.im * s); }
static inline cx divr(cx a, double s) { return C(a.re / s, a.im / s); }
static inline cx addr(cx a, double s) { return C(a.re + s, a.im); }      // complex + real
static inline cx rsub(double s, cx a) { return C(s - a.re, -a.im); }     // real - complex
static inline cx subr(cx a, double s) { return C(a.re - s, a.im); }      // complex - real
static inline cx cdiv(cx a, cx b) {                                      // Smith's method
  double ratio, den, tr, ti;
  if (std::fabs(b.re) < std::fabs(b.im)) {
    ratio = b.re / b.im; den = (b.re * ratio) + b.im;
    tr = (a.re * ratio) + a.im; ti = (a.im * ratio) - a.re;
  } else {
    ratio = b.im / b.re; den = (b.im * ratio) + b.re;
    tr = (a.im * ratio) + a.re; ti = a.im - (a.re * ratio);
  }
  return C(tr / den, ti / den);
}
static inline cx rdiv(double a, cx b) { return cdiv(C(a, 0.0), b); }     // real / complex
static inline double cabs_(cx a) { return std::hypot(a.re, a.im); }
// 10^x of the minors renormalisation (10**lgn, spheroidalMinors.f90:535-552).
// gfortran evaluates it with the C library's pow; here it is computed from
// basic operations only (Cody-Waite reduction x = k log10(2) + r, k the nearest
// integer of x log2(10), log10(2) as a 32-bit head and a tail, the Taylor series
// of exp(r ln 10) to degree 14 by Horner, exact scaling by 2^k), within a few
// ulp of 10^x, so that every IEEE platform gives the same bits.
static double exp10Det(double x) {
  if (!(std::fabs(x) < 4000.0)) return x != x ? x : (x > 0.0 ? HUGE_VAL : 0.0);
  static const double c[14] = {1.0, 2.302585092994046, 2.650949055239199, 2.034678592293476, 1.171255148912267,
                               0.5393829291955814, 0.2069958486968681, 0.06808936507443707, 0.019597694626478524,
                               0.00501392883377544, 0.0011544997789984348, 0.00024166672554424694,
                               4.6371516642572196e-05, 8.213412535439387e-06};
  static const double c13 = 1.3508629476223687e-06;
  const double k = std::floor(x * 3.321928094887362 + 0.5);
  const double r = (x - k * 0.3010299955494702) - k * 1.1451100898021838e-10;
  double p = c13;
  for (int n = 13; n >= 0; --n) p = p * r + c[n];
  return std::ldexp(p, (int)k);
}
static inline cx csqrt_(cx a) {
  std::complex<double> z(a.re, a.im);
  std::complex<double> s = std::sqrt(z);
  return C(s.real(), s.imag());
}

// ----------------------------------------------------------------------------
//  locate (locatePoint.f90:80-110), 1-based array x[1..n]
// ----------------------------------------------------------------------------
static int locate(double xx, int n, const double* x) {
  int lidx = 0, ridx = n + 1;
  while (ridx - lidx > 1) {
    int midx = (lidx + ridx) / 2;
    if (x[n] > x[1]) {
      if (xx > x[midx]) lidx = midx; else ridx = midx;
    } else {
      if (xx > x[midx]) ridx = midx; else lidx = midx;
    }
  }
  return lidx;
}

// ----------------------------------------------------------------------------
//  Node earth model (nodeEarthmodel.f90), SI units, 1-based arrays
// ----------------------------------------------------------------------------
struct Model {
  int nk = 0, nlay = 0, nic = 0, noc = 0, fsflag = 0, equidis = 0;
  double rearth = 0, ric = 0, roc = 0, fref = 1;
  std::vector<double> rk, ro, vpv, vph, vsv, vsh, eta, qkinv, qminv;   // [1..nk]
  std::vector<double> rb, drk;                                          // [1..nlay]
  std::vector<int> iktop, ikbot, iflso;                                 // [1..nlay]
};
struct Nodes {
  int nnod = 0;
  std::vector<double> rnod;     // [1..nnod]
  std::vector<int> state, layer;
};
// spline tables for one frequency (splineEarthmodelCoefficients.f90:28-41)
struct Spline {
  std::vector<double> ro, ro2;             // [1..nk]
  std::vector<double> qinv[3], qinv2[3];   // [1..2][1..nk]
  std::vector<cx> el[8], el2[8];           // [1..7][1..nk]
};

static void getLayerIndex(const Model& m, double r, int& nl, int& top) {     // nodeEarthmodel.f90:599-610
  top = 0;
  nl = locate(r, m.nlay, m.rb.data()) + 1;
  if (std::fabs(r / m.rb[nl] - 1.0) < DEPS) top = 1;
}
static bool isFluid(const Model& m, int nl) { return m.iflso[nl] == 1; }
static bool hasCore(const Model& m) { return m.noc > 0 && m.iflso[m.noc] == 1; }
static bool hasOcean(const Model& m) { return m.iflso[m.nlay] == 1; }
static double rhs_(const Model& m) { return m.rb[1]; }
static double ruc_(const Model& m) { return hasOcean(m) ? m.rb[m.nlay - 1] : m.rearth; }
static double findAbove(const Model& m, int nl) {                            // nodeEarthmodel.f90:619-641
  int npc;
  if (m.iflso[nl] == 1) npc = nl;
  else if (nl < m.noc) npc = m.noc - 1;
  else if (nl > m.noc) npc = (m.iflso[m.nlay] == 1) ? m.nlay - 1 : m.nlay;
  else npc = m.nlay;
  return m.rb[npc];
}
static double findBelow(const Model& m, int nl) {                            // nodeEarthmodel.f90:649-666
  int npc;
  if (m.iflso[nl] == 1) npc = nl;
  else if (nl < m.noc) npc = 1;
  else if (nl > m.noc) npc = m.noc + 1;
  else npc = nl;
  return m.rb[npc - 1];
}

// complexElasticVTIConstants (complexElasticConstants.f90:48-106)
static void vtiConstants(double fratio, double ro, double alfv, double alfh, double betv, double beth,
                         double eta, double qkinv, double qminv, cx z[8]) {
  double phik = std::atan(qkinv);
  double phim = std::atan(qminv);
  cx zqk = mulr(C(std::cos(phik), std::sin(phik)), std::pow(fratio, 2.0 * phik / MC_PID));
  cx zqm = mulr(C(std::cos(phim), std::sin(phim)), std::pow(fratio, 2.0 * phim / MC_PID));
  cx zpa, zpc, zpn, zpl, zpf, zkap, zmue;
  if (betv > (double)1.e-6f) {
    double epa = ro * alfh * alfh;
    double epc = ro * alfv * alfv;
    double epn = ro * beth * beth;
    double epl = ro * betv * betv;
    double epf = eta * (epa - 2. * epl);
    double kappa = (epc + 4. * epa - 4. * epn + 4. * epf) / 9.0;
    double mue = (epc + epa + 6. * epl + 5. * epn - 2. * epf) / 15.0;
    double lamp2mue = kappa + 4. * mue / 3.;
    double lambda = kappa - 2. * mue / 3.;
    double da = epa - lamp2mue, dc = epc - lamp2mue, df = epf - lambda, dn = epn - mue, dl = epl - mue;
    zkap = rmul(kappa, zqk);
    zmue = rmul(mue, zqm);
    cx zlamp2mue = add(zkap, divr(rmul(4., zmue), 3.));
    cx zlambda = sub(zkap, divr(rmul(2., zmue), 3.));
    zpa = addr(zlamp2mue, da); zpc = addr(zlamp2mue, dc); zpn = addr(zmue, dn);
    zpl = addr(zmue, dl); zpf = addr(zlambda, df);
  } else {
    double kappa = ro * alfv * alfv;
    zkap = rmul(kappa, zqk);
    zmue = C(0.0); zpa = zkap; zpc = zkap; zpf = zkap; zpl = C(0.0); zpn = C(0.0);
  }
  z[1] = zpa; z[2] = zpc; z[3] = zpf; z[4] = zpl; z[5] = zpn; z[6] = zkap; z[7] = zmue;
}

// natural cubic spline second derivatives (cubicSpline.f90:74-109) with the
// tridiagonal solve of LAPACK dgtsv (published reference algorithm; the spline
// matrix is strictly diagonally dominant so no row interchange is taken)
static int cubicSpline(int n, const double* x, const double* y, double* y2) {  // all 1-based
  std::vector<double> dl(n + 2, 0.0), d(n + 2, 0.0), du(n + 2, 0.0);
  for (int j = 2; j <= n - 1; ++j) {
    dl[j - 1] = (x[j] - x[j - 1]) / 6.;
    d[j] = (x[j + 1] - x[j - 1]) / 3.;
    du[j] = (x[j + 1] - x[j]) / 6.;
    y2[j] = (y[j + 1] - y[j]) / (x[j + 1] - x[j]) - (y[j] - y[j - 1]) / (x[j] - x[j - 1]);
  }
  d[1] = x[2] - x[1];
  d[n] = x[n] - x[n - 1];
  du[1] = 0.0;
  dl[n - 1] = 0.0;
  y2[1] = 0.0;
  y2[n] = 0.0;
  // dgtsv forward elimination (NRHS = 1)
  for (int i = 1; i <= n - 1; ++i) {
    if (std::fabs(d[i]) >= std::fabs(dl[i])) {
      if (d[i] == 0.0) return ERR_SPLINE;
      double fact = dl[i] / d[i];
      d[i + 1] = d[i + 1] - fact * du[i];
      y2[i + 1] = y2[i + 1] - fact * y2[i];
      if (i <= n - 2) dl[i] = 0.0;
    } else {
      return ERR_SPLINE;   // row interchange: never reached for a spline matrix
    }
  }
  if (d[n] == 0.0) return ERR_SPLINE;
  y2[n] = y2[n] / d[n];
  if (n > 1) y2[n - 1] = (y2[n - 1] - du[n - 1] * y2[n]) / d[n - 1];
  for (int i = n - 2; i >= 1; --i) y2[i] = (y2[i] - du[i] * y2[i + 1] - dl[i] * y2[i + 2]) / d[i];
  return OK;
}

// computeSplineEarthmodelCoefficients (splineEarthmodelCoefficients.f90:46-94)
// attmode: 0 ELASTIC, 1 ATTENUATION_ONLY, 2 DISPERSION_ONLY, 3 ATTENUATION_AND_DISPERSION
static int buildSpline(const Model& m, double f, int attmode, Spline& s) {
  int nk = m.nk;
  s.ro.assign(nk + 1, 0.0); s.ro2.assign(nk + 1, 0.0);
  for (int k = 1; k <= 2; ++k) { s.qinv[k].assign(nk + 1, 0.0); s.qinv2[k].assign(nk + 1, 0.0); }
  for (int k = 1; k <= 7; ++k) { s.el[k].assign(nk + 1, C(0.0)); s.el2[k].assign(nk + 1, C(0.0)); }
  for (int j = 1; j <= nk; ++j) {                     // nodeEarthmodel.f90:287-330
    cx z[8];
    double fratio = 1.0, qk = 0.0, qm = 0.0;
    if (attmode == 1) { qk = m.qkinv[j]; qm = m.qminv[j]; }
    else if (attmode == 2) { fratio = f / m.fref; }
    else if (attmode == 3) { fratio = f / m.fref; qk = m.qkinv[j]; qm = m.qminv[j]; }
    vtiConstants(fratio, m.ro[j], m.vpv[j], m.vph[j], m.vsv[j], m.vsh[j], m.eta[j], qk, qm, z);
    for (int k = 1; k <= 7; ++k) s.el[k][j] = z[k];
    s.ro[j] = m.ro[j];
    s.qinv[1][j] = m.qkinv[j];
    s.qinv[2][j] = m.qminv[j];
  }
  std::vector<double> y(nk + 2), y2(nk + 2);
  for (int nl = 2; nl <= m.nlay; ++nl) {
    int j1 = m.ikbot[nl], j2 = m.iktop[nl];
    int n = j2 - j1 + 1;
    const double* x = &m.rk[j1 - 1];                  // x[1] = rk[j1]
    auto spl = [&](const double* yy, double* out) -> int {
      return cubicSpline(n, x, yy, out);
    };
    if (spl(&s.ro[j1 - 1], &s.ro2[j1 - 1]) != OK) return ERR_SPLINE;
    for (int k = 1; k <= 2; ++k)
      if (spl(&s.qinv[k][j1 - 1], &s.qinv2[k][j1 - 1]) != OK) return ERR_SPLINE;
    for (int k = 1; k <= 7; ++k) {                    // doubleComplexCubicSpline: re and im apart
      for (int i = 1; i <= n; ++i) y[i] = s.el[k][j1 - 1 + i].re;
      if (spl(y.data(), y2.data()) != OK) return ERR_SPLINE;
      for (int i = 1; i <= n; ++i) s.el2[k][j1 - 1 + i] = C(y2[i], 0.0);
      for (int i = 1; i <= n; ++i) y[i] = s.el[k][j1 - 1 + i].im;
      if (spl(y.data(), y2.data()) != OK) return ERR_SPLINE;
      for (int i = 1; i <= n; ++i) {
        cx t = s.el2[k][j1 - 1 + i];
        s.el2[k][j1 - 1 + i] = C(t.re + 0.0 * y2[i], t.im + 1.0 * y2[i]);
      }
    }
  }
  return OK;
}

// evalComplexElasticConstantsSplineEarthmodel (splineEarthmodelCoefficients.f90:118-188)
static void evalElcon(const Model& m, const Spline& s, double r, int nl, double& ro, cx z[8],
                      double* qinv = nullptr) {
  if (nl == 1) {
    int j = m.iktop[1];
    ro = s.ro[j];
    for (int k = 1; k <= 7; ++k) z[k] = s.el[k][j];
    if (qinv) { qinv[0] = s.qinv[1][j]; qinv[1] = s.qinv[2][j]; }
    return;
  }
  int j1 = m.iktop[nl - 1] + 1, j2 = m.iktop[nl], j;
  if (j2 - j1 == 1) {
    j = j1;
  } else if (m.equidis == 1) {
    j = (int)((r - m.rk[j1]) / m.drk[nl]) + j1;
    if (j == j2) j = j2 - 1;
  } else {
    j = j1 - 1;
    int jr = j2 + 1;
    while (jr - j > 1) {
      int mm = (j + jr) / 2;
      if (r > m.rk[mm]) j = mm; else jr = mm;
    }
    if (j == j1 - 1) j = j1;
  }
  double x1 = m.rk[j], x2 = m.rk[j + 1];
  double h = x2 - x1;
  double a = (x2 - r) / h;
  double b = 1.0 - a;
  if (j2 - j1 == 1) {
    ro = a * s.ro[j] + b * s.ro[j + 1];
    for (int k = 1; k <= 7; ++k) z[k] = add(rmul(a, s.el[k][j]), rmul(b, s.el[k][j + 1]));
    if (qinv) for (int k = 1; k <= 2; ++k) qinv[k - 1] = a * s.qinv[k][j] + b * s.qinv[k][j + 1];
  } else {
    double c = h * h * a * (a * a - 1.0) / 6.0;
    double d = h * h * b * (b * b - 1.0) / 6.0;
    ro = a * s.ro[j] + b * s.ro[j + 1] + c * s.ro2[j] + d * s.ro2[j + 1];
    for (int k = 1; k <= 7; ++k)
      z[k] = add(add(add(rmul(a, s.el[k][j]), rmul(b, s.el[k][j + 1])), rmul(c, s.el2[k][j])),
                 rmul(d, s.el2[k][j + 1]));
    if (qinv)
      for (int k = 1; k <= 2; ++k)
        qinv[k - 1] = a * s.qinv[k][j] + b * s.qinv[k][j + 1] + c * s.qinv2[k][j] + d * s.qinv2[k][j + 1];
  }
}

// ----------------------------------------------------------------------------
//  Work counters (the same counters are kept by the HIP kernels)
// ----------------------------------------------------------------------------
struct Stats {
  int64_t v[10] = {0};
  // 0 propagations, 1 accepted steps, 2 rejected steps, 3 MMP sequences,
  // 4 matvecs, 5 sysmat evaluations (reference caching), 6 Neville vector cells,
  // 7 algorithmic flops, 8 flagged zero determinants, 9 integration sub-intervals
  void merge(const Stats& o) { for (int i = 0; i < 10; ++i) v[i] += o.v[i]; }
};

// ----------------------------------------------------------------------------
//  Integration environment (geminiIntegrationEnvironment.f90)
// ----------------------------------------------------------------------------
struct Env {
  const Model* m;
  const Nodes* nod;
  const Spline* s;
  double omre, omim, dll1;
  int nl;
};

// ODE systems: sysmat A(x) with nonzero pattern; entries stored dense nvar x nvar
struct SysMat { cx a[6][6]; };   // 1-based

static void sysmatMinors(const Env& e, double x, SysMat& S) {      // minorsOdeSystem.f90:77-126
  double ro; cx z[8];
  evalElcon(*e.m, *e.s, x, e.nl, ro, z);
  cx zom = C(e.omre, e.omim);
  double dll1 = e.dll1;
  double rr = 1.0 / x;
  double rr2 = rr * rr;
  cx zomro = mulr(mul(zom, zom), ro);
  cx zrpc = rdiv(1.0, z[2]);
  cx z1 = sub(sub(z[1], mul(mul(z[3], z[3]), zrpc)), z[5]);
  cx z2 = mul(rmul(2.0, z[3]), zrpc);
  cx z3 = mulr(rmul(4.0, z1), rr2);
  for (int i = 1; i <= 5; ++i) for (int j = 1; j <= 5; ++j) S.a[i][j] = C(0.0);
  S.a[1][1] = C(-2.0 * rr);
  S.a[1][2] = neg(mulr(rmul(2.0 * dll1, z1), rr2));
  S.a[1][3] = C(dll1 * rr);
  S.a[1][4] = neg(mulr(mul(rmul(dll1, z[3]), zrpc), rr));
  S.a[2][2] = mulr(rsub(1.0, z2), rr);
  S.a[2][3] = rdiv(1.0, z[4]);
  S.a[2][4] = zrpc;
  S.a[3][1] = neg(mulr(z2, rr));
  S.a[3][2] = add(neg(zomro), mulr(sub(rmul(dll1, add(z1, z[5])), rmul(2.0, z[5])), rr2));
  S.a[3][3] = neg(mulr(C(3.0 + z2.re, z2.im), rr));
  S.a[3][5] = S.a[2][4];
  S.a[4][1] = neg(S.a[1][1]);
  S.a[4][2] = add(neg(zomro), z3);
  S.a[4][4] = neg(S.a[2][2]);
  S.a[4][5] = S.a[2][3];
  S.a[5][1] = z3;
  S.a[5][3] = S.a[4][2];
  S.a[5][4] = S.a[3][2];
  S.a[5][5] = mulr(subr(z2, 5.0), rr);
}
static void sysmatToroidal(const Env& e, double x, SysMat& S) {    // toroidalOdeSystem.f90:81-102
  double ro; cx z[8];
  evalElcon(*e.m, *e.s, x, e.nl, ro, z);
  cx zom = C(e.omre, e.omim);
  double dll1 = e.dll1;
  cx zomro = mulr(mul(zom, zom), ro);
  double rr = 1.0 / x;
  double rr2 = rr * rr;
  S.a[1][1] = C(rr);
  S.a[1][2] = rdiv(1.0, z[4]);
  S.a[2][1] = sub(neg(zomro), mulr(mulr(z[5], rr2), 2.0 - dll1));
  S.a[2][2] = C(-3. * rr);
}
static void sysmatFluid(const Env& e, double x, SysMat& S) {       // fluidOdeSystem.f90:78-95
  double ro; cx z[8];
  evalElcon(*e.m, *e.s, x, e.nl, ro, z);
  cx zom = C(e.omre, e.omim);
  double dll1 = e.dll1;
  cx zomro = mulr(mul(zom, zom), ro);
  S.a[1][1] = C(-2.0 / x);
  S.a[1][2] = sub(rdiv(1.0, z[2]), rdiv(dll1, mulr(mulr(zomro, x), x)));
  S.a[2][1] = neg(zomro);
  S.a[2][2] = C(0.0);
}
static void sysmatRadial(const Env& e, double x, SysMat& S) {      // radialOdeSystem.f90:78-99
  double ro; cx z[8];
  evalElcon(*e.m, *e.s, x, e.nl, ro, z);
  cx zom = C(e.omre, e.omim);
  cx zomro = mulr(mul(zom, zom), ro);
  double rr = 1.0 / x;
  double rr2 = rr * rr;
  S.a[1][1] = cdiv(mulr(rmul(-2., z[3]), rr), z[2]);
  S.a[1][2] = rdiv(1.0, z[2]);
  S.a[2][1] = add(neg(zomro), rmul(4 * rr2, sub(sub(z[1], cdiv(mul(z[3], z[3]), z[2])), z[5])));
  S.a[2][2] = rmul(+2. * rr, subr(cdiv(z[3], z[2]), 1.0));
}
static void sysmatSolidt(const Env& e, double x, SysMat& S) {      // solidtOdeSystem.f90:77-114
  double ro; cx z[8];
  evalElcon(*e.m, *e.s, x, e.nl, ro, z);
  cx zom = C(e.omre, e.omim);
  double dll1 = e.dll1;
  double rr = 1.0 / x;
  double rr2 = rr * rr;
  cx zrpc = rdiv(1.0, z[2]);
  cx z1 = sub(sub(z[1], mul(mul(z[3], z[3]), zrpc)), z[5]);
  cx z2 = mul(mulr(z[3], rr), zrpc);
  cx zdlz1 = rmul(dll1, z1);
  cx zomro = mulr(mul(zom, zom), ro);
  S.a[1][1] = rmul(2.0, z2);
  S.a[2][1] = neg(zrpc);
  S.a[3][1] = neg(rmul(dll1, z2));
  S.a[4][1] = C(0.0);
  S.a[1][2] = sub(zomro, mulr(rmul(4.0, z1), rr2));
  S.a[2][2] = addr(neg(S.a[1][1]), 2.0 * rr);
  S.a[3][2] = mul(C(2.0 * rr2), zdlz1);
  S.a[4][2] = C(-dll1 * rr);
  S.a[1][3] = C(rr);
  S.a[2][3] = C(0.0);
  S.a[3][3] = neg(S.a[1][3]);
  S.a[4][3] = neg(rdiv(1.0, z[4]));
  S.a[1][4] = mulr(rmul(2.0, z1), rr2);
  S.a[2][4] = z2;
  S.a[3][4] = sub(zomro, mulr(add(zdlz1, rmul(dll1 - 2.0, z[5])), rr2));
  S.a[4][4] = C(3.0 * rr);
}

static int nvarOf(int sys) { return sys == SYS_MINORS ? 5 : (sys == SYS_SOLIDT ? 4 : 2); }
static void sysmat(int sys, const Env& e, double x, SysMat& S) {
  switch (sys) {
    case SYS_MINORS: sysmatMinors(e, x, S); break;
    case SYS_TOROIDAL: sysmatToroidal(e, x, S); break;
    case SYS_FLUID: sysmatFluid(e, x, S); break;
    case SYS_RADIAL: sysmatRadial(e, x, S); break;
    default: sysmatSolidt(e, x, S); break;
  }
}

// optional constant-matrix system for engine tests (closed-form exp(A h))
struct ConstSys { int nvar; cx a[6][6]; };

// y_out = A * y (matmul order k = 1..n; exact zero entries skipped: adding a
// signed zero product does not change a sum)
static inline void matvec(int n, const SysMat& S, const cx* y, cx* out) {
  for (int i = 1; i <= n; ++i) {
    cx acc = C(0.0);
    bool first = true;
    for (int k = 1; k <= n; ++k) {
      const cx& a = S.a[i][k];
      if (a.re == 0.0 && a.im == 0.0) continue;
      cx p = mul(a, y[k]);
      acc = first ? p : add(acc, p);
      first = false;
    }
    out[i] = acc;
  }
}

// ----------------------------------------------------------------------------
//  Bulirsch-Stoer step engine (complexBulirschStep.f90)
// ----------------------------------------------------------------------------
struct Engine {
  int sys;            // system id, or -1 for the constant test system
  int nvar;
  const Env* env;
  const ConstSys* cs;
  double eps;
  int donetry;
  bool success;
  cx ysol[8][6];      // [try][var]
  cx yext[6], yerr[6];
  SysMat sm[KGV + 1];
  Stats* st;

  void getSysmat(double x, SysMat& S) {
    if (sys < 0) { for (int i = 1; i <= nvar; ++i) for (int k = 1; k <= nvar; ++k) S.a[i][k] = cs->a[i][k]; }
    else sysmat(sys, *env, x, S);
  }
  // doLinearModifiedMidpointComplexBulirsch (complexBulirschStep.f90:180-238)
  void mmp(int jtry, double h, double xstart, const cx* y0) {
    for (int j = 1; j <= NXDIV[jtry]; ++j) {
      int ix = IXDIV[jtry][j];
      double x = xstart + (ix * h) / KGV;
      getSysmat(x, sm[ix]);
    }
    st->v[5] += NXDIV[jtry];
    int nstep = SEQDIV[jtry];
    double hdiv = h / nstep;
    cx z0[6], z1[6], z2[6], w[6];
    for (int i = 1; i <= nvar; ++i) z0[i] = y0[i];
    matvec(nvar, sm[0], z0, w);
    for (int i = 1; i <= nvar; ++i) z1[i] = add(z0[i], rmul(hdiv, w[i]));
    for (int m = 1; m <= nstep - 1; ++m) {
      int ix = m * KGV / nstep;
      matvec(nvar, sm[ix], z1, w);
      for (int i = 1; i <= nvar; ++i) z2[i] = add(z0[i], rmul(2.0 * hdiv, w[i]));
      for (int i = 1; i <= nvar; ++i) { z0[i] = z1[i]; z1[i] = z2[i]; }
    }
    matvec(nvar, sm[KGV], z1, w);
    for (int i = 1; i <= nvar; ++i) ysol[jtry][i] = rmul(0.5, add(add(z1[i], z0[i]), rmul(hdiv, w[i])));
    donetry = jtry;
    st->v[3] += 1;
    st->v[4] += nstep + 1;
    if (sys >= 0)
      st->v[7] += (int64_t)NXDIV[jtry] * ESYS[sys] + (int64_t)(nstep + 1) * 8 * NNZ[sys] + 4LL * nvar * (nstep + 1);
  }
  // extrapolateComplexBulirsch (complexBulirschStep.f90:297-340)
  void extrapolate(double hs, cx* yy, cx* dy) {
    int n = donetry;
    double x[8];
    cx c[8][6], d[8][6], u[6];
    for (int i = 1; i <= n; ++i) x[i] = hs / SEQDIV[i];
    int js = n;
    for (int j = 1; j <= n; ++j) for (int v = 1; v <= nvar; ++v) { c[j][v] = ysol[j][v]; d[j][v] = c[j][v]; }
    for (int v = 1; v <= nvar; ++v) yy[v] = ysol[js][v];
    for (int m = 1; m <= n - 1; ++m) {
      for (int i = 1; i <= n - m; ++i) {
        double h = x[i] - x[i + m];
        for (int v = 1; v <= nvar; ++v) {
          u[v] = divr(sub(c[i + 1][v], d[i][v]), h);
          c[i][v] = rmul(x[i], u[v]);
          d[i][v] = rmul(x[i + m], u[v]);
        }
        st->v[6] += 1;
        st->v[7] += 8 * nvar;
      }
      for (int v = 1; v <= nvar; ++v) { yy[v] = add(yy[v], d[js - 1][v]); dy[v] = d[js - 1][v]; }
      js = js - 1;
    }
  }
  // doFullStepComplexBulirsch (complexBulirschStep.f90:111-170)
  void fullStep(double h, double xstart, const cx* y) {
    success = false;
    donetry = 0;
    int ntry = MINTRY, nfirst = 1;
    cx zyy[6], zdy[6];
    for (;;) {
      for (int j = nfirst; j <= ntry; ++j) mmp(j, h, xstart, y);
      extrapolate(h, zyy, zdy);
      double sdy = 0.0, syy = 0.0;
      for (int v = 1; v <= nvar; ++v) { double a = cabs_(zdy[v]); sdy = sdy + a * a; }
      for (int v = 1; v <= nvar; ++v) { double a = cabs_(zyy[v]); syy = syy + a * a; }
      double errmax = std::sqrt(sdy / syy);
      if (errmax / eps < 1.0) {
        for (int v = 1; v <= nvar; ++v) { yext[v] = zyy[v]; yerr[v] = zdy[v]; }
        success = true;
        return;
      }
      if (ntry < MAXTRY) { ntry = ntry + 1; nfirst = ntry; }
      else { success = false; return; }
    }
  }
  // adjustStepSizeComplexBulirsch (complexBulirschStep.f90:346-359)
  void adjust(double& h, bool succ) {
    if (succ) {
      if (donetry > OPTTRY) h = h * SHRINK;
      else if (donetry < OPTTRY) h = h * GROW;
    } else {
      h = 0.5 * h / (double)(1 << ((MAXTRY - OPTTRY) / 2));
    }
  }
};

// doIntegrationStep (integrationStep.f90:98-190). hdid carries over between
// sub-intervals of one propagation (integration_step object lifetime).
static int integrationStep(Engine& eng, double& hdid, double x1, double x2, cx* y) {
  if (std::fabs(x2 - x1) < MINSIZE) return OK;
  Stats& st = *eng.st;
  double x = x1;
  double h = std::copysign(std::fabs(hdid), x2 - x1);
  cx ys[6];
  for (int v = 1; v <= eng.nvar; ++v) ys[v] = y[v];
  int nstep = 0;
  while (nstep < MAXNUM) {
    if ((x + h - x2) * (x + h - x1) > 0.0) h = x2 - x;
    eng.fullStep(h, x, ys);
    nstep = nstep + 1;
    if (eng.success) {
      st.v[1] += 1;
      x = x + h;
      for (int v = 1; v <= eng.nvar; ++v) ys[v] = eng.yext[v];
      if (std::fabs(x - x2) < MINSIZE) {
        for (int v = 1; v <= eng.nvar; ++v) y[v] = ys[v];
        return OK;
      } else {
        eng.adjust(h, true);
        hdid = h;
      }
    } else {
      st.v[2] += 1;
      eng.adjust(h, false);
      if (std::fabs(h) < MINSIZE) return ERR_MINSTEP;
    }
  }
  return ERR_MAXSTEP;
}

// propagation result (propagateOde.f90 members ynr/yni/lgynorm/jl/jr)
struct PropResult {
  int jl = 1, jr = 0;              // node index limits (jr < jl: none)
  std::vector<std::vector<cx>> yn; // [k][var], k = 1..nn in integration order
  std::vector<double> lgyn;
  cx yend[6];
  double lgnorm = 0.0;
};

// getNodesGemini (geminiIntegrationEnvironment.f90:150-199)
static void getNodes(const Env& e, double x1, double x2, int& jl, int& jr, std::vector<double>& xn) {
  const Nodes& N = *e.nod;
  double xl = x1 > x2 ? x2 : x1, xr = x1 > x2 ? x1 : x2;
  xn.clear();
  const double* rn = N.rnod.data();
  jl = locate(xl, N.nnod, rn) + 1;
  if (jl > N.nnod) { jr = jl - 1; return; }
  jr = locate(xr, N.nnod, rn);
  if (jr < N.nnod) { if (std::fabs(1.0 - xr / rn[jr + 1]) < DEPS) jr = jr + 1; }
  if (jr == 0) { jr = 0; jl = 1; return; }
  int nn = jr - jl + 1;
  if (nn > 0) {
    if (x2 >= x1) for (int j = jl; j <= jr; ++j) xn.push_back(rn[j]);
    else for (int j = jr; j >= jl; --j) xn.push_back(rn[j]);
  }
}
// getDiscontinuitiesGemini (geminiIntegrationEnvironment.f90:92-135)
static void getDiscos(const Env& e, double x1, double x2, std::vector<double>& xd) {
  const Model& M = *e.m;
  double xl = x1 > x2 ? x2 : x1, xr = x1 > x2 ? x1 : x2;
  xd.clear();
  const double* rb = M.rb.data();
  int jl = locate(xl, M.nlay, rb) + 1, jr;
  if (jl > M.nlay) return;
  if (std::fabs(1.0 - xl / rb[jl]) < DEPS) jl = jl + 1;
  if (jl >= M.nlay) return;
  jr = locate(xr, M.nlay, rb);
  if (jr == 0) return;
  int nd = jr - jl + 1;
  if (nd > 0) {
    if (x2 >= x1) for (int j = jl; j <= jr; ++j) xd.push_back(rb[j]);
    else for (int j = jr; j >= jl; --j) xd.push_back(rb[j]);
  }
}

// doPropagateOde (propagateOde.f90:132-331)
static int propagate(Engine& eng, Env& env, bool do_norm, double lgnorm0, double x1, double x2, double htry,
                     const cx* ystart, PropResult& P) {
  int nvar = eng.nvar;
  cx y[6];
  for (int v = 1; v <= nvar; ++v) y[v] = ystart[v];
  P.lgnorm = lgnorm0;
  std::vector<double> xn, xd;
  getNodes(env, x1, x2, P.jl, P.jr, xn);
  int nnod = (int)xn.size();
  P.yn.assign(nnod + 1, std::vector<cx>(6, C(0.0)));
  P.lgyn.assign(nnod + 1, 0.0);
  getDiscos(env, x1, x2, xd);
  int nd = (int)xd.size();
  int idir = (x2 - x1) >= 0 ? 1 : -1;
  if (nd > 0) {
    if ((x1 - xd[0]) * idir > 0 || (xd[nd - 1] - x2) * idir > 0) return ERR_RANGE;
  }
  if (nnod > 0) {
    if ((x1 - xn[0]) * idir > 0 || (xn[nnod - 1] - x2) * idir > 0) return ERR_RANGE;
  }
  int jd = 1, jnod = 1;
  eng.st->v[0] += 1;
  if (std::fabs(x2 - x1) < MINSIZE && nnod == 0) {
    for (int v = 1; v <= nvar; ++v) P.yend[v] = y[v];
    return OK;
  } else if (std::fabs(x2 - x1) < MINSIZE && nnod > 1) {
    return ERR_RANGE;
  } else if (std::fabs(x2 - x1) < MINSIZE && nnod == 1) {
    for (int v = 1; v <= nvar; ++v) P.yn[1][v] = y[v];
    P.lgyn[1] = P.lgnorm;          // untouched in the reference (zero-initialised there)
    P.lgyn[1] = 0.0;
    jnod = 2; jd = 1;
  } else {
    double hdid = htry;
    double xa = x1;
    while (std::fabs(x2 - xa) > MINSIZE) {
      double steps[4];
      steps[3] = std::fabs(x2 - xa);
      steps[1] = 2. * std::fabs(x2 - x1);
      steps[2] = 3. * std::fabs(x2 - x1);
      if (jnod <= nnod) steps[1] = std::fabs(xn[jnod - 1] - xa);
      if (jd <= nd) steps[2] = std::fabs(xd[jd - 1] - xa);
      int imin = 1;
      double smin = steps[1];
      if (steps[2] < smin) { smin = steps[2]; imin = 2; }
      if (steps[3] < smin) { smin = steps[3]; imin = 3; }
      double xe = xa + idir * smin;
      eng.st->v[9] += 1;
      int rc = integrationStep(eng, hdid, xa, xe, y);
      if (rc != OK) return rc;
      if (do_norm) {
        double ynorm = 0.0;
        for (int v = 1; v <= nvar; ++v) ynorm = ynorm + (std::fabs(y[v].re) + std::fabs(y[v].im));
        ynorm = ynorm / nvar;
        if (ynorm > MAXY) {
          for (int v = 1; v <= nvar; ++v) y[v] = C(y[v].re / ynorm, y[v].im / ynorm);
          P.lgnorm = P.lgnorm + std::log10(ynorm);
        }
      }
      if (imin == 1) {
        for (int v = 1; v <= nvar; ++v) P.yn[jnod][v] = y[v];
        P.lgyn[jnod] = P.lgnorm;
        jnod = jnod + 1;
        if (std::fabs(steps[1] - steps[2]) < MINSIZE) { env.nl += idir; jd = jd + 1; }
      } else if (imin == 2) {
        jd = jd + 1;
        env.nl += idir;
        if (std::fabs(steps[1] - steps[2]) < MINSIZE) {
          for (int v = 1; v <= nvar; ++v) P.yn[jnod][v] = y[v];
          P.lgyn[jnod] = P.lgnorm;
          jnod = jnod + 1;
        }
      }
      xa = xe;
    }
  }
  if (jnod - 1 != nnod) return ERR_RANGE;
  if (jd - 1 != nd) return ERR_RANGE;
  for (int v = 1; v <= nvar; ++v) P.yend[v] = y[v];
  return OK;
}

// ----------------------------------------------------------------------------
//  initialValues.f90
// ----------------------------------------------------------------------------
static double vslow(double om, double dll1, double r, double v) { return dll1 / (r * r) - (om / v) * (om / v); }

static void velocityAt(const Env& e, double r, int nl, double& v) {
  double ro; cx z[8];
  evalElcon(*e.m, *e.s, r, nl, ro, z);
  v = std::sqrt(z[7].re / ro);
  if (isFluid(*e.m, nl)) v = std::sqrt(z[6].re / ro);
}

static void integrateSlowness(const Env& e, double om, double dll1, double rturn, double dr, bool tor, double& rs) {
  const Model& M = *e.m;
  double q = 0.0, r = rturn, rbot;
  if (tor && hasCore(M)) rbot = M.roc; else rbot = rhs_(M) - (double)0.1f * dr;
  double qmax = -std::log(1.e-5);
  while (q < qmax) {
    if (r - dr < rbot) { r = rbot; break; }
    r = r - dr;
    int nl, top;
    getLayerIndex(M, r, nl, top);
    double v;
    velocityAt(e, r, nl, v);
    q = q + dr * std::sqrt(vslow(om, dll1, r, v));
  }
  rs = r;
}

// startRadiusInitialValues (initialValues.f90:71-149)
static double startRadius(const Env& e, bool tor) {
  const Model& M = *e.m;
  double dll1 = e.dll1, om = e.omre, rearth = M.rearth, rhs = rhs_(M), roc = M.roc;
  int noc = M.noc, nl, top;
  double r;
  if (tor && hasCore(M)) { r = roc; nl = noc + 1; } else { r = rhs; nl = 1; }
  double dr = (rearth - r) / 400.0;
  double v;
  velocityAt(e, r, nl, v);
  double p = vslow(om, dll1, r, v), pold = p;
  if (p <= 0.0) {
    if (tor && hasCore(M)) return roc;
    return rhs - (double)0.1f * dr;
  }
  double rmax = (tor && hasOcean(M)) ? ruc_(M) : rearth;
  double res;
  while (p > 0.0) {
    pold = p;
    if (std::fabs(1.0 - r / rmax) < DEPS) {
      integrateSlowness(e, om, dll1, rmax, dr, tor, res);
      return res;
    }
    r = std::fmin(r + dr, rmax);
    getLayerIndex(M, r, nl, top);
    velocityAt(e, r, nl, v);
    p = vslow(om, dll1, r, v);
  }
  double rturn = r - dr * p / (p - pold);
  integrateSlowness(e, om, dll1, rturn, dr, tor, res);
  return res;
}

// sphancfInitialValues (initialValues.f90:487-526)
static int sphancf(cx zx, double el, cx& zf, int kind) {
  const double tiny = 1.e-30, eps = 1.e-6;
  const int maxit = 10000;
  zf = C(tiny);
  cx zc = zf, zd = C(0.0);
  double elph2 = (el + .5) * (el + .5);
  int j;
  for (j = 1; j <= maxit; ++j) {
    cx zb = rmul(2.0, C(zx.re + 0.0, zx.im + (double)j));
    double za = ((double)j - .5) * ((double)j - .5) - elph2;
    zd = add(zb, rmul(za, zd));
    if (cabs_(zd) < tiny) zd = C(tiny);
    zc = add(zb, rdiv(za, zc));
    if (cabs_(zc) < tiny) zc = C(tiny);
    zd = rdiv(1.0, zd);
    cx zdelta = mul(zc, zd);
    zf = mul(zf, zdelta);
    if (cabs_(subr(zdelta, 1.0)) < eps) break;
  }
  if (j > maxit) return ERR_CF;
  cx izx = C(-zx.im, zx.re), izf = C(-zf.im, zf.re);
  if (kind == 1) zf = sub(sub(C(el + 1.0), izx), izf);
  else zf = add(add(C(el + 1.0), izx), izf);
  return OK;
}
// zetlInitialValues (initialValues.f90:544-574)
static int zetl(cx zx2, double el, cx& z) {
  const double tiny = 1.e-30, eps = 1.e-6;
  const int maxit = 10000;
  z = C(tiny);
  cx zd = C(2.0 * el + 3.0);
  cx zc = add(zd, cdiv(zx2, z));
  zd = rdiv(1.0, zd);
  cx zdelta = mul(zc, zd);
  z = mul(z, zdelta);
  for (int i = 2; i <= maxit; ++i) {
    zd = rsub(2.0 * (el + (double)i) + 1.0, mul(zx2, zd));
    if (zd.re == 0.0 && zd.im == 0.0) zd = C(tiny);
    zc = rsub(2.0 * (el + (double)i) + 1.0, cdiv(zx2, zc));
    if (zc.re == 0.0 && zc.im == 0.0) zc = C(tiny);
    zd = rdiv(1.0, zd);
    zdelta = mul(zc, zd);
    z = mul(z, zdelta);
    if (std::fabs(cabs_(zdelta) - 1.0) < eps) return OK;
  }
  return ERR_CF;
}

// minorsBottomInitialValues (initialValues.f90:202-275)
static int minorsBottom(const Env& e, double rstart, cx* ys, double& htry) {
  int nl = e.nl;
  double ro; cx z[8];
  evalElcon(*e.m, *e.s, rstart, nl, ro, z);
  bool liquid = isFluid(*e.m, nl);
  double el = -0.5 + std::sqrt(0.25 + e.dll1);
  double elp1 = e.dll1;
  cx zom = C(e.omre, e.omim);
  double rsq = rstart * rstart;
  double v = std::sqrt(z[7].re / ro);
  if (liquid) v = std::sqrt(z[6].re / ro);
  htry = 0.25 * v * MC_TWO_PID / e.omre;
  cx zxa2 = cdiv(mulr(mul(rmul(ro, zom), zom), rsq), add(z[6], divr(rmul(4.0, z[7]), 3.0)));
  cx zeta = C(0.0), zetb = C(0.0), zxb2 = C(0.0);
  int rc;
  if (el < DEPS) {
    if ((rc = zetl(zxa2, el, zeta)) != OK) return rc;
  } else {
    if (cabs_(divr(zxa2, elp1)) <= 1.0) { if ((rc = zetl(zxa2, el, zeta)) != OK) return rc; }
    else { if ((rc = sphancf(csqrt_(zxa2), el, zeta, 1)) != OK) return rc; }
    if (!liquid) {
      zxb2 = cdiv(mulr(mul(rmul(ro, zom), zom), rsq), z[7]);
      if (cabs_(divr(zxb2, elp1)) <= 1.0) { if ((rc = zetl(zxb2, el, zetb)) != OK) return rc; }
      else { if ((rc = sphancf(csqrt_(zxb2), el, zetb, 1)) != OK) return rc; }
    }
  }
  if (el < DEPS) {
    ys[1] = neg(zeta);
    ys[2] = add(neg(mul(rmul(ro * rstart, zom), zom)), mul(divr(rmul(4.0, z[7]), rstart), zeta));
  } else if (liquid) {
    ys[1] = neg(cdiv(rsub(1.0, divr(zeta, el)), mulr(mulr(mul(zom, zom), ro), rstart)));
    ys[2] = C(1.0 / el);
  } else {
    cx za = divr(zeta, el);
    cx zb = divr(zetb, el + 1.0);
    cx mr = divr(z[7], rstart);
    ys[2] = divr(add(neg(za), mul(zb, subr(za, 1.0))), el);
    ys[1] = mul(mr, add(divr(zxb2, el), rmul(2.0 * elp1, ys[2])));
    ys[3] = mul(mr, add(neg(rmul(2.0, ys[2])), mul(divr(zxb2, elp1), subr(za, 1.0))));
    ys[4] = neg(mul(mr, add(mul(divr(zxb2, el * el), rsub(1.0, zb)), rmul(4.0, ys[2]))));
    cx t1 = neg(rmul(4.0 * (el - 1.0) * (el + 2.0), ys[2]));
    cx inner = sub(sub(subr(divr(zxb2, elp1), 2.0 * (el - 1.0) * (2.0 * el + 1.0) / elp1),
                       rmul(4.0 / (el + 1.0), za)),
                   rmul(2.0 / el, zb));
    ys[5] = mul(divr(mul(z[7], z[7]), rsq), add(t1, mul(divr(zxb2, el), inner)));
  }
  return OK;
}
// The full-space start vector of minorsSurfaceInitialValues
// (initialValues.f90:312-343) at radius rs in layer nl: minors of the outgoing
// (second-kind spherical Hankel) P and S solutions of the homogeneous space.
static int minorsOutgoing(const Env& e, double rs, int nl, cx* ys) {
  const Model& M = *e.m;
  double elp1 = e.dll1;
  double el = -0.5 + std::sqrt(0.25 + elp1);
  bool liquid = isFluid(M, nl);
  double ro; cx z[8];
  evalElcon(M, *e.s, rs, nl, ro, z);
  cx zom = C(e.omre, e.omim);
  double rsq = rs * rs;
  cx zxa2 = cdiv(mulr(mul(rmul(ro, zom), zom), rsq), add(z[6], divr(rmul(4.0, z[7]), 3.0)));
  cx zeta, zetb = C(0.0), zxb2 = C(0.0);
  int rc;
  if ((rc = sphancf(csqrt_(zxa2), el, zeta, 2)) != OK) return rc;
  if (!liquid) {
    zxb2 = cdiv(mulr(mul(rmul(ro, zom), zom), rsq), z[7]);
    if ((rc = sphancf(csqrt_(zxb2), el, zetb, 2)) != OK) return rc;
  }
  if (el < DEPS) {
    ys[1] = neg(zeta);
    ys[2] = add(neg(mul(rmul(ro * rs, zom), zom)), mul(divr(rmul(4.0, z[7]), rs), zeta));
  } else if (liquid) {
    ys[1] = neg(cdiv(rsub(1.0, divr(zeta, el)), mulr(mulr(mul(zom, zom), ro), rs)));
    ys[2] = C(1.0 / el);
  } else {
    cx za = divr(zeta, el);
    cx zb = divr(zetb, el + 1.0);
    cx mr = divr(z[7], rs);
    ys[2] = divr(add(neg(za), mul(zb, subr(za, 1.0))), el);
    ys[1] = mul(mr, add(divr(zxb2, el), rmul(2.0 * elp1, ys[2])));
    ys[3] = mul(mr, add(neg(rmul(2.0, ys[2])), mul(divr(zxb2, elp1), subr(za, 1.0))));
    ys[4] = neg(mul(mr, add(mul(divr(zxb2, el * el), rsub(1.0, zb)), rmul(4.0, ys[2]))));
    cx t1 = neg(rmul(4.0 * (el - 1.0) * (el + 2.0), ys[2]));
    cx inner = sub(sub(subr(divr(zxb2, elp1), 2.0 * (el - 1.0) * (2.0 * el + 1.0) / elp1),
                       rmul(4.0 / (el + 1.0), za)),
                   rmul(2.0 / el, zb));
    ys[5] = mul(divr(mul(z[7], z[7]), rsq), add(t1, mul(divr(zxb2, el), inner)));
  }
  return OK;
}
// minorsSurfaceInitialValues (initialValues.f90:287-367): free surface, or the
// outgoing solution of a full space above the model (full_space = 1)
static int minorsSurface(const Env& e, cx* ys, double& htry) {
  const Model& M = *e.m;
  int nlay = M.nlay;
  double elp1 = e.dll1;
  double el = -0.5 + std::sqrt(0.25 + elp1);
  bool liquid = isFluid(M, nlay);
  double ro; cx z[8];
  evalElcon(M, *e.s, M.rearth, nlay, ro, z);
  double v = std::sqrt(z[7].re / ro);
  if (liquid) v = std::sqrt(z[6].re / ro);
  htry = 0.25 * v * MC_TWO_PID / e.omre;
  if (M.fsflag == 1) return minorsOutgoing(e, M.rearth, nlay, ys);
  if (el < FEPS || liquid) { ys[1] = C(1.0); ys[2] = C(0.0); }
  else { ys[1] = C(0.0); ys[2] = C(1.0); ys[3] = C(0.0); ys[4] = C(0.0); ys[5] = C(0.0); }
  return OK;
}
// toroidalBottomInitialValues (initialValues.f90:376-413)
static int toroidalBottom(const Env& e, double rstart, cx* ys, double& htry) {
  int nl = e.nl;
  double ro; cx z[8];
  evalElcon(*e.m, *e.s, rstart, nl, ro, z);
  double el = -0.5 + std::sqrt(0.25 + e.dll1);
  double elp1 = e.dll1;
  cx zom = C(e.omre, e.omim);
  double rsq = rstart * rstart;
  double v = std::sqrt(z[7].re / ro);
  htry = 0.25 * v * MC_TWO_PID / e.omre;
  if (rstart > e.m->roc) {
    cx zxb2 = cdiv(mulr(mul(rmul(ro, zom), zom), rsq), z[7]);
    cx zetb;
    int rc;
    if (cabs_(divr(zxb2, elp1)) < 1.0) { if ((rc = zetl(zxb2, el, zetb)) != OK) return rc; }
    else { if ((rc = sphancf(csqrt_(zxb2), el, zetb, 1)) != OK) return rc; }
    ys[1] = C(1.0 / el);
    ys[2] = mul(divr(z[7], rstart * el), rsub(el - 1.0, zetb));
  } else {
    ys[1] = C(1.0 / el);
    ys[2] = C(0.0);
  }
  return OK;
}
// The full-space start vector of toroidalSurfaceInitialValues
// (initialValues.f90:458-463) at radius rs in layer nl: the outgoing S solution.
static int toroidalOutgoing(const Env& e, double rs, int nl, cx* ys) {
  double ro; cx z[8];
  evalElcon(*e.m, *e.s, rs, nl, ro, z);
  double el = -0.5 + std::sqrt(0.25 + e.dll1);
  cx zom = C(e.omre, e.omim);
  cx zxb2 = cdiv(mulr(mul(rmul(ro, zom), zom), rs * rs), z[7]);
  cx zetb;
  int rc = sphancf(csqrt_(zxb2), el, zetb, 2);
  if (rc != OK) return rc;
  ys[1] = C(1.0 / el);
  ys[2] = mul(divr(z[7], rs * el), rsub(el - 1.0, zetb));
  return OK;
}
// toroidalSurfaceInitialValues (initialValues.f90:427-469)
static int toroidalSurface(const Env& e, cx* ys, double& htry, double& rsurf, int& nl) {
  const Model& M = *e.m;
  if (hasOcean(M)) { nl = M.nlay - 1; rsurf = ruc_(M); } else { nl = M.nlay; rsurf = M.rearth; }
  double ro; cx z[8];
  evalElcon(M, *e.s, rsurf, nl, ro, z);
  double v = std::sqrt(z[7].re / ro);
  htry = 0.25 * v * MC_TWO_PID / e.omre;
  if (M.fsflag == 1) return toroidalOutgoing(e, rsurf, nl, ys);
  ys[1] = C(1.0); ys[2] = C(0.0);
  return OK;
}

// ----------------------------------------------------------------------------
//  Minors drivers
// ----------------------------------------------------------------------------
struct ToroidalMinors {
  std::vector<cx> u, d;   // [2*(n-1)+i] i=1..2 -> stored as (n)(i) 1-based flattened
  std::vector<cx> det;    // [1..nnod]
  cx& U(int i, int n) { return u[(n - 1) * 2 + (i - 1)]; }
  cx& D(int i, int n) { return d[(n - 1) * 2 + (i - 1)]; }
};
struct SpheroidalMinors {
  std::vector<cx> u, d;                 // (6, nnod)
  std::vector<double> ulg, dlg;         // [1..nnod]
  std::vector<cx> det;                  // [1..nnod]
  cx& U(int i, int n) { return u[(n - 1) * 6 + (i - 1)]; }
  cx& D(int i, int n) { return d[(n - 1) * 6 + (i - 1)]; }
};

// computeToroidalMinors (toroidalMinors.f90:71-258)
static int toroidalMinors(Env env, int jsl, int jsr, double eps, ToroidalMinors& T, Stats& st, double* rstart_out) {
  const Nodes& N = *env.nod;
  int nnod = N.nnod;
  T.u.assign(2 * nnod, C(0.0)); T.d.assign(2 * nnod, C(0.0)); T.det.assign(nnod + 1, C(0.0));
  if (jsl < 1 || jsl > nnod || jsr < 1 || jsr > nnod) return ERR_NODEIDX;
  double rstart = startRadius(env, true);
  if (rstart_out) *rstart_out = rstart;
  if (env.dll1 < DEPS) return OK;
  double rsl = N.rnod[jsl];
  double rmin = std::fmax(rstart, rsl);
  double rmax = N.rnod[jsr];
  if (rstart >= rmax) return OK;
  Engine eng; eng.sys = SYS_TOROIDAL; eng.nvar = 2; eng.env = &env; eng.cs = nullptr; eng.eps = eps; eng.st = &st;
  cx ys[6]; double htry, rsurf; int nl;
  int rc = toroidalSurface(env, ys, htry, rsurf, nl);
  if (rc != OK) return rc;
  if (rsurf < rmin) return OK;
  env.nl = nl;
  PropResult P;
  rc = propagate(eng, env, false, 0.0, rsurf, rmin, htry, ys, P);
  if (rc != OK) return rc;
  if (P.jr - P.jl < 0) return ERR_RANGE;
  for (int k = 1; k <= P.jr - P.jl + 1; ++k) {          // descending order in P
    int n = P.jr - (k - 1);
    T.D(1, n) = P.yn[k][1]; T.D(2, n) = P.yn[k][2];
  }
  rmax = std::fmin(rmax, rsurf);
  int top;
  getLayerIndex(*env.m, rstart, nl, top);
  if (top == 1) nl = nl + 1;
  env.nl = nl;
  rc = toroidalBottom(env, rstart, ys, htry);
  if (rc != OK) return rc;
  PropResult Q;
  rc = propagate(eng, env, false, 0.0, rstart, rmax, htry, ys, Q);
  if (rc != OK) return rc;
  if (Q.jr - Q.jl < 0) return ERR_RANGE;
  for (int k = 1; k <= Q.jr - Q.jl + 1; ++k) {
    int n = Q.jl + (k - 1);
    T.U(1, n) = Q.yn[k][1]; T.U(2, n) = Q.yn[k][2];
  }
  for (int js = jsl; js <= jsr; ++js) T.det[js] = sub(mul(T.U(1, js), T.D(2, js)), mul(T.D(1, js), T.U(2, js)));
  return OK;
}

// computeSpheroidalMinors (spheroidalMinors.f90:72-585)
static int spheroidalMinors(Env env, int jsl, int jsr, double eps, SpheroidalMinors& S, Stats& st, double* rstart_out) {
  const Nodes& N = *env.nod;
  const Model& M = *env.m;
  int nnod = N.nnod;
  S.u.assign(6 * nnod, C(0.0)); S.d.assign(6 * nnod, C(0.0));
  S.ulg.assign(nnod + 1, 0.0); S.dlg.assign(nnod + 1, 0.0); S.det.assign(nnod + 1, C(0.0));
  if (jsl < 1 || jsl > nnod || jsr < 1 || jsr > nnod) return ERR_NODEIDX;
  double rstart = startRadius(env, false);
  if (rstart_out) *rstart_out = rstart;
  double rsl = N.rnod[jsl];
  double rmin = std::fmax(rstart, rsl);
  double rmax = N.rnod[jsr];
  if (rstart >= rmax) return OK;
  double rearth = M.rearth;
  bool l0 = env.dll1 < DEPS;
  Engine eng; eng.env = &env; eng.cs = nullptr; eng.eps = eps; eng.st = &st;
  for (int pass = 0; pass < 2; ++pass) {
    bool up = (pass == 0);
    double rbeg = up ? rstart : rearth, rend = rbeg;
    bool liq_old = false;
    double lgnorm = 0.0, lgnormend = 0.0;
    cx yend[6];
    double htry = 0.0;     // set by the start-value routine, reused by later segments
    while (up ? (rend < rmax) : (rend > rmin)) {
      int nl, top;
      getLayerIndex(M, rbeg, nl, top);
      if (up && top == 1) nl = nl + 1;
      bool liq = isFluid(M, nl);
      env.nl = nl;
      rend = up ? std::fmin(findAbove(M, nl), rmax) : std::fmax(findBelow(M, nl), rmin);
      int sys = liq ? SYS_FLUID : (l0 ? SYS_RADIAL : SYS_MINORS);
      int nvar = nvarOf(sys);
      eng.sys = sys; eng.nvar = nvar;
      cx ys[6];
      for (int v = 0; v < 6; ++v) ys[v] = C(0.0);
      if (rbeg == (up ? rstart : rearth)) {
        int rc = up ? minorsBottom(env, rbeg, ys, htry) : minorsSurface(env, ys, htry);
        if (rc != OK) return rc;
      } else {
        if (liq && !liq_old) {
          if (l0) { ys[1] = yend[1]; ys[2] = yend[2]; } else { ys[1] = yend[3]; ys[2] = yend[5]; }
        } else if (!liq && liq_old) {
          if (l0) { ys[1] = yend[1]; ys[2] = yend[2]; }
          else { ys[4] = yend[2]; ys[2] = yend[1]; ys[1] = C(0.0); ys[3] = C(0.0); ys[5] = C(0.0); }
        } else {
          return ERR_STATE;
        }
      }
      PropResult P;
      int rc = propagate(eng, env, true, lgnorm, rbeg, rend, htry, ys, P);
      if (rc != OK) return rc;
      int jl = P.jl, jr = P.jr;
      if (up) {
        if (jr - jl >= 0) {
          for (int k = 1; k <= jr - jl + 1; ++k) {
            int n = jl + (k - 1);
            S.ulg[n] = P.lgyn[k];
            for (int v = 1; v <= nvar; ++v) S.U(v, n) = P.yn[k][v];
            S.U(6, n) = divr(neg(S.U(1, n)), env.dll1);
          }
        }
      } else {
        if (jr - jl < 0) return ERR_RANGE;
        for (int k = 1; k <= jr - jl + 1; ++k) {
          int n = jr - (k - 1);
          S.dlg[n] = P.lgyn[k];
          for (int v = 1; v <= nvar; ++v) S.D(v, n) = P.yn[k][v];
        }
        for (int n = jl; n <= jr; ++n) S.D(6, n) = divr(neg(S.D(1, n)), env.dll1);
      }
      liq_old = liq;
      rbeg = rend;
      for (int v = 1; v <= nvar; ++v) yend[v] = P.yend[v];
      lgnormend = P.lgnorm;
      lgnorm = lgnormend;
    }
    for (int n = 1; n <= nnod; ++n) {
      if (up) {
        double lgn = S.ulg[n] - lgnormend;
        double f = exp10Det(lgn);
        for (int i = 1; i <= 6; ++i) S.U(i, n) = mulr(S.U(i, n), f);
      } else {
        double lgn = S.dlg[n] - lgnormend;
        double f = exp10Det(lgn);
        for (int i = 1; i <= 6; ++i) S.D(i, n) = mulr(S.D(i, n), f);
      }
    }
  }
  for (int j = jsl; j <= jsr; ++j) {
    if (N.state[j] == 1 || l0) {
      S.det[j] = sub(mul(S.U(1, j), S.D(2, j)), mul(S.D(1, j), S.U(2, j)));
    } else {
      cx t = add(mul(S.U(1, j), S.D(6, j)), mul(S.U(6, j), S.D(1, j)));
      t = sub(t, mul(S.U(2, j), S.D(5, j)));
      t = add(t, mul(S.U(3, j), S.D(4, j)));
      t = add(t, mul(S.U(4, j), S.D(3, j)));
      t = sub(t, mul(S.U(5, j), S.D(2, j)));
      S.det[j] = t;
    }
  }
  return OK;
}

}  // namespace orc

namespace orc {

// computeSpheroidalAdjoint (spheroidalAdjoint.f90:53-159): returns adj at nodes
// (4 x nnod, zero where not reached) and overwrites ys with the end value.
static int spheroidalAdjoint(Env env, cx* ys, double ra, double re, double eps, std::vector<cx>& adj, Stats& st) {
  const Model& M = *env.m;
  const Nodes& N = *env.nod;
  int nnod = N.nnod;
  adj.assign(4 * nnod, C(0.0));
  int nl, top;
  getLayerIndex(M, ra, nl, top);
  if (top == 1 && ra < re) nl = std::min(nl + 1, M.nlay);
  env.nl = nl;
  double ro; cx z[8];
  evalElcon(M, *env.s, ra, nl, ro, z);
  double v = std::sqrt(z[7].re / ro);
  if (v < DEPS) v = std::sqrt(add(z[6], rmul(4.0 / 3.0, z[7])).re / ro);
  double htry = 0.25 * v * MC_TWO_PID / env.omre;
  Engine eng; eng.sys = SYS_SOLIDT; eng.nvar = 4; eng.env = &env; eng.cs = nullptr; eng.eps = eps; eng.st = &st;
  PropResult P;
  int rc = propagate(eng, env, false, 0.0, ra, re, htry, ys, P);
  if (rc != OK) return rc;
  if (P.jr - P.jl < 0) return ERR_RANGE;
  for (int j = P.jl; j <= P.jr; ++j) {
    int k = (re > ra) ? (j - P.jl + 1) : (P.jr - j + 1);
    for (int i = 1; i <= 4; ++i) adj[(j - 1) * 4 + (i - 1)] = P.yn[k][i];
  }
  for (int i = 1; i <= 4; ++i) ys[i] = P.yend[i];
  return OK;
}

// minors matrix M~ and eps*M~ (spheroidalMinors.f90:306-323) from the six minors
// stored for node n in v[(n-1)*6 .. (n-1)*6+5]
static void mtilFromMinors(const std::vector<cx>& v, int n, cx mt[5][5], cx me[5][5]) {
  cx m6[7];
  for (int i = 1; i <= 6; ++i) m6[i] = v[(n - 1) * 6 + (i - 1)];
  static const int idx[5][5] = {{0}, {0, 0, 1, 2, 3}, {0, 0, 0, 4, 5}, {0, 0, 0, 0, 6}, {0}};
  for (int i = 1; i <= 4; ++i) for (int j = 1; j <= 4; ++j) { mt[i][j] = C(0.0); me[i][j] = C(0.0); }
  for (int i = 1; i <= 4; ++i)
    for (int j = i + 1; j <= 4; ++j) {
      double vf = -1.0 + 2.0 * ((i + j) % 2);
      mt[i][j] = m6[idx[i][j]];
      mt[j][i] = neg(mt[i][j]);
      me[i][j] = rmul(vf, m6[7 - idx[i][j]]);
      me[j][i] = neg(me[i][j]);
    }
}

struct Column {                // one (omega, l) Green-FK column
  std::vector<cx> ujs;         // ujgfsph (4, 4, nnod): [(n-1)*16 + (j-1)*4 + (i-1)]
  std::vector<cx> ujt;         // ujgftor (2, 2, nnod): [(n-1)*4 + (j-1)*2 + (i-1)]
  std::vector<cx> gf;          // (nsp, nnod): [(n-1)*nsp + (isp-1)]
  int nsp = 0;
  int detflag = 0;             // 1: toroidal det below tiny, 2: spheroidal det below tiny
};

// computeUnitJumpGreenFunctions (unitJumpGreenFunctions.f90:62-380), source at node jn
static int unitJumpGreen(const Env& env, ToroidalMinors& T, SpheroidalMinors& S, int jn, double eps, Column& col,
                         Stats& st) {
  const Model& M = *env.m;
  const Nodes& N = *env.nod;
  int nnod = N.nnod;
  if (jn < 1 || jn > nnod) return ERR_NODEIDX;
  double ruc = ruc_(M), roc = M.roc, ric = M.ric;
  col.ujs.assign(16 * nnod, C(0.0));
  col.ujt.assign(4 * nnod, C(0.0));
  // a jump node in a fluid layer: the ocean (unitJumpGreenFunctions.f90:119-123, 242) only
  const bool ocean = N.state[jn] == 1;
  if (ocean && N.layer[jn] != M.nlay) return ERR_UNSUPPORTED;
  auto US = [&](int i, int j, int n) -> cx& { return col.ujs[(n - 1) * 16 + (j - 1) * 4 + (i - 1)]; };
  auto UT = [&](int i, int j, int n) -> cx& { return col.ujt[(n - 1) * 4 + (j - 1) * 2 + (i - 1)]; };
  double ujv[5][5] = {{0}};
  for (int i = 1; i <= 4; ++i) ujv[i][i] = 1.0;
  // toroidal (none for a jump in the ocean: njt = 0)
  if (ocean) {
  } else if (cabs_(T.det[jn]) > TINYD) {
    double rstart = startRadius(env, true);
    for (int j = 1; j <= 2; ++j) {
      cx d = cdiv(sub(rmul(ujv[2][j], T.D(1, jn)), rmul(ujv[1][j], T.D(2, jn))), T.det[jn]);
      for (int n = 1; n <= jn - 1; ++n)
        if (N.state[n] == 0 && N.rnod[n] > rstart) { UT(1, j, n) = mul(d, T.U(1, n)); UT(2, j, n) = mul(d, T.U(2, n)); }
      cx e = cdiv(sub(rmul(ujv[2][j], T.U(1, jn)), rmul(ujv[1][j], T.U(2, jn))), T.det[jn]);
      for (int n = jn; n <= nnod; ++n)
        if (N.state[n] == 0) { UT(1, j, n) = mul(e, T.D(1, n)); UT(2, j, n) = mul(e, T.D(2, n)); }
    }
  } else {
    col.detflag |= 1;
  }
  // spheroidal
  if (!(cabs_(S.det[jn]) > TINYD)) { col.detflag |= 2; st.v[8] += 1; return OK; }
  if (env.dll1 < DEPS) {
    for (int j = 1; j <= 2; ++j) {
      cx d = cdiv(sub(rmul(ujv[2][j], S.D(1, jn)), rmul(ujv[1][j], S.D(2, jn))), S.det[jn]);
      cx e = cdiv(sub(rmul(ujv[2][j], S.U(1, jn)), rmul(ujv[1][j], S.U(2, jn))), S.det[jn]);
      for (int n = jn; n <= nnod; ++n) { US(1, j, n) = mul(e, S.D(1, n)); US(2, j, n) = mul(e, S.D(2, n)); }
      for (int n = jn - 1; n >= 1; --n) { US(1, j, n) = mul(d, S.U(1, n)); US(2, j, n) = mul(d, S.U(2, n)); }
    }
    return OK;
  }
  double rstart = startRadius(env, false);
  std::vector<cx> adjic;
  {
    double ra = ric, re = std::fmax(rstart, N.rnod[1]);
    if (re < ra) {
      cx ys[6] = {C(0), C(0), C(0), C(0), C(1.0), C(0)};
      int rc = spheroidalAdjoint(env, ys, ra, re, eps, adjic, st);
      if (rc != OK) return rc;
    } else {
      adjic.assign(4 * nnod, C(0.0));
    }
  }
  cx mtn[5][5], men[5][5];
  if (ocean) {
    // jump in the ocean (:242-298): the mantle propagator P_4m(r) from the ocean
    // bottom down to the start radius, the CMB or the lowest node, once for both
    // jumps; e D(r) at and above the jump, d U(r) in the ocean below it, d u_mtil
    // P_4m in the mantle, d P_4m(roc)_3 U(r) in the outer core and
    // d P_4m(roc)_3 u_mtil P_4m^ic in the inner core
    std::vector<cx> adjma(4 * nnod, C(0.0));
    cx adjma_roc3 = C(0.0);
    const double ra = ruc, re = std::fmax(std::fmax(rstart, roc), N.rnod[1]);
    if (re < ra) {
      cx ys[6] = {C(0), C(0), C(0), C(0), C(1.0), C(0)};
      int rc = spheroidalAdjoint(env, ys, ra, re, eps, adjma, st);
      if (rc != OK) return rc;
      adjma_roc3 = ys[3];
    }
    for (int j = 1; j <= 2; ++j) {
      cx d = cdiv(sub(rmul(ujv[2][j], S.D(1, jn)), rmul(ujv[1][j], S.D(2, jn))), S.det[jn]);
      cx e = cdiv(sub(rmul(ujv[2][j], S.U(1, jn)), rmul(ujv[1][j], S.U(2, jn))), S.det[jn]);
      for (int n = jn; n <= nnod; ++n) { US(1, j, n) = mul(e, S.D(1, n)); US(2, j, n) = mul(e, S.D(2, n)); }
      for (int n = jn - 1; n >= 1; --n) {
        if (N.rnod[n] < rstart) continue;
        if (N.state[n] == 1) {
          if (N.rnod[n] > ruc) {
            US(1, j, n) = mul(d, S.U(1, n)); US(2, j, n) = mul(d, S.U(2, n));
          } else if (N.rnod[n] < roc) {
            const cx s = mul(d, adjma_roc3);
            US(1, j, n) = mul(s, S.U(1, n)); US(2, j, n) = mul(s, S.U(2, n));
          }
        } else {
          const std::vector<cx>* a = nullptr;
          cx s = d;
          if (N.rnod[n] > roc) a = &adjma;
          else if (N.rnod[n] < ric) { a = &adjic; s = mul(d, adjma_roc3); }
          if (a) {
            mtilFromMinors(S.u, n, mtn, men);
            for (int i = 1; i <= 4; ++i) {
              cx acc = C(0.0);
              for (int k = 1; k <= 4; ++k) acc = add(acc, mul(mtn[i][k], (*a)[(n - 1) * 4 + (k - 1)]));
              US(i, j, n) = mul(s, acc);
            }
          }
        }
      }
    }
    return OK;
  }
  for (int j = 1; j <= 4; ++j) {
    // adjoint above the source: f(+)(r_s) = -s_i/det * u_meps(:,i)
    cx ys[6];
    for (int v = 0; v < 6; ++v) ys[v] = C(0.0);
    mtilFromMinors(S.u, jn, mtn, men);
    for (int i = 1; i <= 4; ++i)
      for (int v = 1; v <= 4; ++v) ys[v] = sub(ys[v], cdiv(rmul(ujv[i][j], men[v][i]), S.det[jn]));
    std::vector<cx> adjp, adjm;
    int rc = spheroidalAdjoint(env, ys, N.rnod[jn], ruc, eps, adjp, st);
    if (rc != OK) return rc;
    cx adj_ruc3 = ys[3];
    // adjoint below the source: f(-)(r_s) = +s_i/det * d_meps(:,i)
    for (int v = 0; v < 6; ++v) ys[v] = C(0.0);
    mtilFromMinors(S.d, jn, mtn, men);
    for (int i = 1; i <= 4; ++i)
      for (int v = 1; v <= 4; ++v) ys[v] = add(ys[v], cdiv(rmul(ujv[i][j], men[v][i]), S.det[jn]));
    double re = std::fmax(std::fmax(rstart, roc), N.rnod[1]);
    rc = spheroidalAdjoint(env, ys, N.rnod[jn], re, eps, adjm, st);
    if (rc != OK) return rc;
    cx adj_roc3 = ys[3];
    for (int n = jn; n <= nnod; ++n) {
      if (N.state[n] == 0 && N.rnod[n] < ruc) {
        mtilFromMinors(S.d, n, mtn, men);
        for (int i = 1; i <= 4; ++i) {
          cx acc = C(0.0);
          for (int k = 1; k <= 4; ++k) acc = add(acc, mul(mtn[i][k], adjp[(n - 1) * 4 + (k - 1)]));
          US(i, j, n) = acc;
        }
      } else if (N.state[n] == 1 && N.rnod[n] > ruc) {
        US(1, j, n) = mul(adj_ruc3, S.D(1, n)); US(2, j, n) = mul(adj_ruc3, S.D(2, n));
      }
    }
    for (int n = jn - 1; n >= 1; --n) {
      if (N.state[n] == 0 && N.rnod[n] >= rstart) {
        const std::vector<cx>* a = nullptr;
        bool scale = false;
        if (N.rnod[n] > roc) a = &adjm;
        else if (N.rnod[n] < ric) { a = &adjic; scale = true; }
        if (a) {
          mtilFromMinors(S.u, n, mtn, men);
          for (int i = 1; i <= 4; ++i) {
            cx acc = C(0.0);
            for (int k = 1; k <= 4; ++k) acc = add(acc, mul(mtn[i][k], (*a)[(n - 1) * 4 + (k - 1)]));
            US(i, j, n) = scale ? mul(adj_roc3, acc) : acc;
          }
        }
      } else if (N.state[n] == 1 && N.rnod[n] >= rstart) {
        US(1, j, n) = mul(adj_roc3, S.U(1, n)); US(2, j, n) = mul(adj_roc3, S.U(2, n));
      }
    }
  }
  return OK;
}

// spheroidalReciprocity for a solid receiver (reciprocity.f90:82-153):
// bg(i,k) for i = component at the receiver, k = jump; out 4x4 (solid node) or 4x2 (fluid node)
static void reciprocitySph(double elp1, double re, double r, int state_node, const cx* bgin /*(4,4) col-major*/,
                           cx bgout[5][5]) {
  auto B = [&](int i, int k) { return bgin[(k - 1) * 4 + (i - 1)]; };
  double rlp1 = 1.0 / elp1;
  double f = (r / re) * (r / re);
  for (int i = 1; i <= 4; ++i) for (int k = 1; k <= 4; ++k) bgout[i][k] = C(0.0);
  if (state_node == 1) {      // node in fluid, source (receiver position) in solid: bgout(4,2)
    bgout[1][1] = neg(rmul(f, B(2, 2)));
    bgout[2][1] = rmul(f, B(2, 1));
    if (elp1 > DEPS) {
      bgout[3][1] = neg(mulr(rmul(f, B(2, 4)), rlp1));
      bgout[4][1] = mulr(rmul(f, B(2, 3)), rlp1);
    }
    bgout[1][2] = rmul(f, B(1, 2));
    bgout[2][2] = neg(rmul(f, B(1, 1)));
    if (elp1 > DEPS) {
      bgout[3][2] = mulr(rmul(f, B(1, 4)), rlp1);
      bgout[4][2] = neg(mulr(rmul(f, B(1, 3)), rlp1));
    }
  } else {
    bgout[1][1] = neg(rmul(f, B(2, 2)));
    bgout[2][1] = rmul(f, B(2, 1));
    bgout[3][1] = neg(mulr(rmul(f, B(2, 4)), rlp1));
    bgout[4][1] = mulr(rmul(f, B(2, 3)), rlp1);
    bgout[1][2] = rmul(f, B(1, 2));
    bgout[2][2] = neg(rmul(f, B(1, 1)));
    bgout[3][2] = mulr(rmul(f, B(1, 4)), rlp1);
    bgout[4][2] = neg(mulr(rmul(f, B(1, 3)), rlp1));
    bgout[1][3] = neg(mulr(rmul(f, B(4, 2)), elp1));
    bgout[2][3] = mulr(rmul(f, B(4, 1)), elp1);
    bgout[3][3] = neg(rmul(f, B(4, 4)));
    bgout[4][3] = rmul(f, B(4, 3));
    bgout[1][4] = mulr(rmul(f, B(3, 2)), elp1);
    bgout[2][4] = neg(mulr(rmul(f, B(3, 1)), elp1));
    bgout[3][4] = rmul(f, B(3, 4));
    bgout[4][4] = neg(rmul(f, B(3, 3)));
  }
}
static void reciprocityTor(double re, double r, const cx* bgin /*(2,2)*/, cx bgout[3][3]) {  // reciprocity.f90:169-180
  auto B = [&](int i, int k) { return bgin[(k - 1) * 2 + (i - 1)]; };
  double f = (r / re) * (r / re);
  bgout[1][1] = neg(rmul(f, B(2, 2)));
  bgout[2][1] = rmul(f, B(2, 1));
  bgout[1][2] = rmul(f, B(1, 2));
  bgout[2][2] = neg(rmul(f, B(1, 1)));
}

// oneNodeSourceTerms (sourceTerms.f90:59-118); istyp 1 = MOMENT, 0 = FORCE
static void sourceTerms(const Env& env, int istyp, double rs, int state, int nl, cx zs[5][5], cx zt[3][3]) {
  double elp1 = env.dll1, om = env.omre;
  double rr2 = 1. / (rs * rs);
  double rr3 = rr2 / rs;
  for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) zs[i][j] = C(0.0);
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) zt[i][j] = C(0.0);
  double ro; cx z[8];
  if (istyp == 1) {
    evalElcon(*env.m, *env.s, rs, nl, ro, z);
    if (state == 1) {
      zs[1][1] = rmul(rr2, sub(rdiv(1.0, z[2]), C(elp1 * rr2 / (om * om * ro))));
      zs[2][1] = C(rr3 * 2.0);
      zs[1][2] = C(0.5 * elp1 * rr2 * rr2 / (om * om * ro));
      zs[2][2] = C(-rr3);
    } else {
      zs[1][1] = rdiv(rr2, z[2]);
      zs[2][1] = cdiv(rmul(rr3 * 2.0, z[3]), z[2]);
      zs[4][1] = cdiv(rmul(-rr3, z[3]), z[2]);
      zs[2][2] = C(-rr3);
      zs[4][2] = C(0.5 * rr3);
      zs[3][3] = rdiv(0.5 * rr2, z[4]);
      zs[4][4] = C(0.25 * rr3);
      zt[1][1] = rdiv(0.5 * rr2, z[4]);
      zt[2][2] = C(0.25 * rr3);
    }
  } else {
    if (state == 1) {
      evalElcon(*env.m, *env.s, rs, nl, ro, z);
      zs[1][1] = C(0.0);
      zs[2][1] = C(-rr2);
      zs[1][2] = C(elp1 * rr3 / (om * om * ro));
      zs[2][2] = C(0.0);
    } else {
      zs[2][1] = C(-rr2);
      zs[4][2] = C(0.5 * rr2);
      zt[2][1] = C(0.5 * rr2);
    }
  }
}

struct Ctx {
  Model m;
  Nodes nod;
  Spline s;
  double f = -1.0;
  int attmode = 3;
};

static void horcompFluid(const Env& env, double r, int nj, cx gf[5][5], cx v[5], cx vr[5]);

// full column for one (omega, l): computeGreenFKSpectraForSynthetics.f90:413-494.
// For a receiver in the ocean only rows 1-2 of the reciprocal Green functions
// enter, and spheroidalReciprocity's rows 1-2 for a source in the ocean
// (reciprocity.f90:99-103, 119-128) are those for a source in the solid.
// dsvstep 2: U, V (and W) per node; 1 (STRESSFLAG, :209-212): U, R, V, S and W, T, with the
// spheroidal mask (1, 1, 1, 1) also for a receiver in the ocean, whose S slots stay zero (:313)
static int gfkColumn(const Ctx& c, double omre, double omim, double dll1, int jre, int istyp, double eps,
                     Column& col, Stats& st, ToroidalMinors* Tout = nullptr, SpheroidalMinors* Sout = nullptr,
                     int dsvstep = 2) {
  Env env; env.m = &c.m; env.nod = &c.nod; env.s = &c.s; env.omre = omre; env.omim = omim; env.dll1 = dll1; env.nl = 1;
  const Nodes& N = c.nod;
  int nnod = N.nnod;
  if (jre < 1 || jre > nnod) return ERR_NODEIDX;
  // a receiver in a fluid layer: the ocean only (ncsph = 2, nctor = 0, :293-300)
  const bool ocean = N.state[jre] == 1;
  if (ocean && N.layer[jre] != c.m.nlay) return ERR_UNSUPPORTED;
  ToroidalMinors T; SpheroidalMinors S;
  int rc = toroidalMinors(env, jre, jre, eps, T, st, nullptr);
  if (rc != OK) return rc;
  rc = spheroidalMinors(env, jre, jre, eps, S, st, nullptr);
  if (rc != OK) return rc;
  rc = unitJumpGreen(env, T, S, jre, eps, col, st);
  if (rc != OK) return rc;
  int njsph = istyp == 1 ? 4 : 2, njtor = istyp == 1 ? 2 : 1;
  const int ncsph = ocean ? 2 : 4, nctor = ocean ? 0 : 2;
  // U and V of the spheroidal jumps (V from (U, R) for a receiver in the ocean), W of the toroidal
  // ones; with stresses also R, S and T
  const int nmask = dsvstep == 1 ? 4 : 2;
  col.nsp = nmask * njsph + (ocean ? 0 : (nmask / 2) * njtor);
  col.gf.assign(col.nsp * nnod, C(0.0));
  double re = N.rnod[jre];
  for (int n = 1; n <= nnod; ++n) {
    cx bgs[5][5], bgt[3][3], zs[5][5], zt[3][3], gfs[5][5], gft[3][3];
    for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) gfs[i][j] = C(0.0);
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) gft[i][j] = C(0.0);
    reciprocitySph(dll1, re, N.rnod[n], N.state[n], &col.ujs[(n - 1) * 16], bgs);
    reciprocityTor(re, N.rnod[n], &col.ujt[(n - 1) * 4], bgt);
    sourceTerms(env, istyp, N.rnod[n], N.state[n], N.layer[n], zs, zt);
    if (N.state[n] == 1) {
      for (int i = 1; i <= ncsph; ++i)
        for (int j = 1; j <= 2; ++j) {
          cx acc = C(0.0);
          for (int k = 1; k <= 2; ++k) acc = add(acc, mul(bgs[i][k], zs[k][j]));
          gfs[i][j] = acc;
        }
    } else {
      for (int i = 1; i <= ncsph; ++i)
        for (int j = 1; j <= njsph; ++j) {
          cx acc = C(0.0);
          for (int k = 1; k <= 4; ++k) acc = add(acc, mul(bgs[i][k], zs[k][j]));
          gfs[i][j] = acc;
        }
      for (int i = 1; i <= nctor; ++i)
        for (int j = 1; j <= njtor; ++j) {
          cx acc = C(0.0);
          for (int k = 1; k <= 2; ++k) acc = add(acc, mul(bgt[i][k], zt[k][j]));
          gft[i][j] = acc;
        }
    }
    int isp = 0;
    for (int i = 1; i <= ncsph; i += dsvstep)
      for (int j = 1; j <= njsph; ++j) { isp++; col.gf[(n - 1) * col.nsp + (isp - 1)] = gfs[i][j]; }
    if (ocean) {
      // V = -R / (r om^2 ro) of the receiver's fluid, evaluated by the reference at the
      // node radius rnod(n) (:465, getHorcompComplexFluid)
      cx v[5], vr[5];
      horcompFluid(env, N.rnod[n], njsph, gfs, v, vr);
      for (int j = 1; j <= njsph; ++j) { isp++; col.gf[(n - 1) * col.nsp + (isp - 1)] = v[j]; }
    }
    for (int i = 1; i <= nctor; i += dsvstep)
      for (int j = 1; j <= njtor; ++j) { isp++; col.gf[(n - 1) * col.nsp + (isp - 1)] = gft[i][j]; }
  }
  if (Tout) *Tout = T;
  if (Sout) *Sout = S;
  return OK;
}


// getSysmatComplexSolid (solidOdeSystem.f90:78-111): the 4 x 4 system of the
// displacement-stress vector (U, R, V, S), used only for the radial derivatives
// of the direct Green functions of the multi-source driver
static void sysmatSolid(const Env& e, double x, SysMat& S) {
  double ro; cx z[8];
  evalElcon(*e.m, *e.s, x, e.nl, ro, z);
  cx zom = C(e.omre, e.omim);
  double dll1 = e.dll1;
  double rr = 1.0 / x;
  double rr2 = rr * rr;
  cx zrpc = rdiv(1.0, z[2]);
  cx z1 = sub(sub(z[1], mul(mul(z[3], z[3]), zrpc)), z[5]);
  cx z2 = mul(mulr(z[3], rr), zrpc);
  cx zdlz1 = rmul(dll1, z1);
  cx zomro = mulr(mul(zom, zom), ro);
  S.a[1][1] = rmul(-2.0, z2);
  S.a[1][2] = zrpc;
  S.a[1][3] = rmul(dll1, z2);
  S.a[1][4] = C(-0.0);
  S.a[2][1] = add(neg(zomro), mulr(rmul(4.0, z1), rr2));
  S.a[2][2] = subr(rmul(2.0, z2), 2.0 * rr);
  S.a[2][3] = mul(C(-2.0 * rr2), zdlz1);
  S.a[2][4] = C(dll1 * rr);
  S.a[3][1] = C(-rr);
  S.a[3][2] = C(-0.0);
  S.a[3][3] = C(rr);
  S.a[3][4] = rdiv(1.0, z[4]);
  S.a[4][1] = mulr(rmul(-2.0, z1), rr2);
  S.a[4][2] = neg(z2);
  S.a[4][3] = add(neg(zomro), mulr(add(zdlz1, rmul(dll1 - 2.0, z[5])), rr2));
  S.a[4][4] = C(-3.0 * rr);
}

// computeDsvDerivatives (dsvDerivatives.f90:38-68): gfd = A(r) gf for nj jump
// columns of nvar components, A of the system `sys` (0 solid, 1 toroidal, 2
// fluid) in the layer that holds r (getLayerIndexNodeEarthmodel); gf and gfd
// are [comp][jump], 1-based
static void dsvDerivatives(const Env& env0, int sys, double r, int nj, cx gf[5][5], cx gfd[5][5]) {
  Env env = env0;
  int nl, top;
  getLayerIndex(*env.m, r, nl, top);
  env.nl = nl;
  SysMat A;
  for (int i = 0; i < 6; ++i) for (int j = 0; j < 6; ++j) A.a[i][j] = C(0.0);
  int nvar = 2;
  if (sys == 0) { sysmatSolid(env, r, A); nvar = 4; }
  else if (sys == 1) sysmatToroidal(env, r, A);
  else sysmatFluid(env, r, A);
  for (int i = 1; i <= nvar; ++i)
    for (int j = 1; j <= nj; ++j) {
      cx acc = C(0.0);
      for (int k = 1; k <= nvar; ++k) acc = add(acc, mul(A.a[i][k], gf[k][j]));
      gfd[i][j] = acc;
    }
}

// getHorcompComplexFluid (fluidOdeSystem.f90:171-188): V = -R / (r om^2 ro) and
// dV/dr = -(U - V) / r at a fluid node, for nj jump columns of (U, R)
static void horcompFluid(const Env& env, double r, int nj, cx gf[5][5], cx v[5], cx vr[5]) {
  int nl, top;
  getLayerIndex(*env.m, r, nl, top);
  double ro; cx z[8];
  evalElcon(*env.m, *env.s, r, nl, ro, z);
  cx zom = C(env.omre, env.omim);
  cx zomro = mulr(mul(zom, zom), ro);
  for (int j = 1; j <= nj; ++j) {
    v[j] = neg(cdiv(gf[2][j], mulr(zomro, r)));
    vr[j] = mulr(sub(gf[1][j], v[j]), -1.0 / r);
  }
}

// One (omega, l) of the multi-source driver (computeGreenFKSpectraForASKI.f90:
// 458-590): minors over the source node range, then for each source node js
// (nsnod moment or force sources at jsnod[js], and a FORCE at the receiver jre
// as source nsnod+1) the unit-jump Green functions of a jump at that node, the
// source terms there, and at every node n the direct Green functions (U, V) and
// (W) with their radial derivatives.  out: [js][n][nspmax] with
// isp = U(j) V(j) dU(j) dV(j) W(j) dW(j), nsp(js) = 4 njsph + 2 njtor, the rest zero.
static int askiColumn(const Ctx& c, double omre, double omim, double dll1, int nsnod, const int* jsnod, int jre,
                      int istyp, double eps, int nspmax, std::vector<cx>& out, Stats& st) {
  Env env; env.m = &c.m; env.nod = &c.nod; env.s = &c.s; env.omre = omre; env.omim = omim; env.dll1 = dll1; env.nl = 1;
  const Nodes& N = c.nod;
  const int nnod = N.nnod;
  if (jre < 1 || jre > nnod) return ERR_NODEIDX;
  int jsl = nnod + 1, jsr = 0;
  bool jre_in = false;
  for (int js = 0; js < nsnod; ++js) {
    if (jsnod[js] < 1 || jsnod[js] > nnod) return ERR_NODEIDX;
    jsl = std::min(jsl, jsnod[js]);
    jsr = std::max(jsr, jsnod[js]);
    if (jsnod[js] == jre) jre_in = true;
  }
  if (!jre_in) return ERR_NODEIDX;          // the receiver node must also be a source node (:287-290)
  out.assign((size_t)(nsnod + 1) * nnod * nspmax, C(0.0));
  ToroidalMinors T; SpheroidalMinors S;
  int rc = toroidalMinors(env, jsl, jsr, eps, T, st, nullptr);
  if (rc != OK) return rc;
  rc = spheroidalMinors(env, jsl, jsr, eps, S, st, nullptr);
  if (rc != OK) return rc;
  for (int js = 1; js <= nsnod + 1; ++js) {
    const int styp = js == nsnod + 1 ? 0 : istyp;
    const int jsh = js == nsnod + 1 ? jre : jsnod[js - 1];
    // a source in the ocean (:359-376): two (U, R) jumps for a force or a moment, no toroidal
    // motion; unitJumpGreen refuses a source in the fluid core
    const bool osrc = N.state[jsh] == 1;
    const int njsph = osrc ? 2 : (styp == 1 ? 4 : 2), njtor = osrc ? 0 : (styp == 1 ? 2 : 1);
    const int nk = osrc ? 2 : 4;
    Column col;
    rc = unitJumpGreen(env, T, S, jsh, eps, col, st);
    if (rc != OK) return rc;
    cx zs[5][5], zt[3][3];
    sourceTerms(env, styp, N.rnod[jsh], N.state[jsh], N.layer[jsh], zs, zt);
    for (int n = 1; n <= nnod; ++n) {
      cx gfs[5][5], gfd[5][5], gft[5][5], gfdt[5][5];
      for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) gfs[i][j] = gfd[i][j] = gft[i][j] = gfdt[i][j] = C(0.0);
      auto U = [&](int i, int k) { return col.ujs[(n - 1) * 16 + (k - 1) * 4 + (i - 1)]; };
      auto W = [&](int i, int k) { return col.ujt[(n - 1) * 4 + (k - 1) * 2 + (i - 1)]; };
      if (N.state[n] == 1) {                 // receiver node in the ocean: (U, R), then V and dV/dr
        for (int i = 1; i <= 2; ++i)
          for (int j = 1; j <= njsph; ++j) {
            cx acc = C(0.0);
            for (int k = 1; k <= nk; ++k) acc = add(acc, mul(U(i, k), zs[k][j]));
            gfs[i][j] = acc;
          }
        dsvDerivatives(env, 2, N.rnod[n], njsph, gfs, gfd);
        cx v[5], vr[5];
        horcompFluid(env, N.rnod[n], njsph, gfs, v, vr);
        for (int j = 1; j <= njsph; ++j) { gfs[3][j] = v[j]; gfd[3][j] = vr[j]; }
      } else {
        for (int i = 1; i <= 4; ++i)
          for (int j = 1; j <= njsph; ++j) {
            cx acc = C(0.0);
            for (int k = 1; k <= nk; ++k) acc = add(acc, mul(U(i, k), zs[k][j]));
            gfs[i][j] = acc;
          }
        for (int i = 1; i <= 2; ++i)
          for (int j = 1; j <= njtor; ++j) {
            cx acc = C(0.0);
            for (int k = 1; k <= 2; ++k) acc = add(acc, mul(W(i, k), zt[k][j]));
            gft[i][j] = acc;
          }
        dsvDerivatives(env, 0, N.rnod[n], njsph, gfs, gfd);
        dsvDerivatives(env, 1, N.rnod[n], njtor, gft, gfdt);
      }
      cx* o = &out[((size_t)(js - 1) * nnod + (n - 1)) * nspmax];
      int isp = 0;
      for (int i = 1; i <= 4; i += 2)
        for (int j = 1; j <= njsph; ++j) o[isp++] = gfs[i][j];
      for (int i = 1; i <= 4; i += 2)
        for (int j = 1; j <= njsph; ++j) o[isp++] = gfd[i][j];
      for (int j = 1; j <= njtor; ++j) o[isp++] = gft[1][j];
      for (int j = 1; j <= njtor; ++j) o[isp++] = gfdt[1][j];
    }
  }
  return OK;
}

}  // namespace orc

using namespace orc;

// ============================================================================
//  C ABI (loaded by tests/ and bench.py through ctypes)
// ============================================================================
extern "C" {

void* orc_create(int nk, const double* rk, const double* ro, const double* vpv, const double* vph,
                 const double* vsv, const double* vsh, const double* eta, const double* qkinv,
                 const double* qminv, int nlay, const int* iktop, const double* rb, const int* iflso, int nic,
                 int noc, int fsflag, int equidis, double rearth, double ric, double roc, double fref, int nnod,
                 const double* rnod, const int* state, const int* layer) {
  Ctx* c = new Ctx();
  Model& m = c->m;
  m.nk = nk; m.nlay = nlay; m.nic = nic; m.noc = noc; m.fsflag = fsflag; m.equidis = equidis;
  m.rearth = rearth; m.ric = ric; m.roc = roc; m.fref = fref;
  auto cp = [&](std::vector<double>& v, const double* src, int n) { v.assign(n + 1, 0.0); for (int i = 0; i < n; ++i) v[i + 1] = src[i]; };
  cp(m.rk, rk, nk); cp(m.ro, ro, nk); cp(m.vpv, vpv, nk); cp(m.vph, vph, nk); cp(m.vsv, vsv, nk);
  cp(m.vsh, vsh, nk); cp(m.eta, eta, nk); cp(m.qkinv, qkinv, nk); cp(m.qminv, qminv, nk); cp(m.rb, rb, nlay);
  m.iktop.assign(nlay + 1, 0); m.ikbot.assign(nlay + 1, 0); m.iflso.assign(nlay + 1, 0);
  for (int i = 0; i < nlay; ++i) { m.iktop[i + 1] = iktop[i]; m.iflso[i + 1] = iflso[i]; }
  m.ikbot[1] = 0;
  for (int nl = 2; nl <= nlay; ++nl) m.ikbot[nl] = m.iktop[nl - 1] + 1;
  m.drk.assign(nlay + 1, -1.0);
  if (equidis == 1)
    for (int nl = 2; nl <= nlay; ++nl)
      m.drk[nl] = (m.rk[m.iktop[nl]] - m.rk[m.ikbot[nl]]) / (m.iktop[nl] - m.ikbot[nl]);
  Nodes& N = c->nod;
  N.nnod = nnod;
  cp(N.rnod, rnod, nnod);
  N.state.assign(nnod + 1, 0); N.layer.assign(nnod + 1, 0);
  for (int i = 0; i < nnod; ++i) { N.state[i + 1] = state[i]; N.layer[i + 1] = layer[i]; }
  return c;
}

void orc_destroy(void* h) { delete (Ctx*)h; }

int orc_set_frequency(void* h, double f, int attmode) {
  Ctx* c = (Ctx*)h;
  c->f = f; c->attmode = attmode;
  return buildSpline(c->m, f, attmode, c->s);
}

// spline tables as flat arrays: ro[nk], ro2[nk], el[7][nk] complex, el2[7][nk] complex
int orc_get_spline(void* h, double* ro, double* ro2, double* el, double* el2) {
  Ctx* c = (Ctx*)h;
  int nk = c->m.nk;
  for (int j = 1; j <= nk; ++j) { ro[j - 1] = c->s.ro[j]; ro2[j - 1] = c->s.ro2[j]; }
  for (int k = 1; k <= 7; ++k)
    for (int j = 1; j <= nk; ++j) {
      el[2 * ((k - 1) * nk + j - 1)] = c->s.el[k][j].re; el[2 * ((k - 1) * nk + j - 1) + 1] = c->s.el[k][j].im;
      el2[2 * ((k - 1) * nk + j - 1)] = c->s.el2[k][j].re; el2[2 * ((k - 1) * nk + j - 1) + 1] = c->s.el2[k][j].im;
    }
  return 0;
}

int orc_eval_elcon(void* h, double r, int nl, double* ro, double* zel, double* qinv) {
  Ctx* c = (Ctx*)h;
  cx z[8];
  evalElcon(c->m, c->s, r, nl, *ro, z, qinv);
  for (int k = 1; k <= 7; ++k) { zel[2 * (k - 1)] = z[k].re; zel[2 * (k - 1) + 1] = z[k].im; }
  return 0;
}

int orc_layer_index(void* h, double r, int* nl, int* top) {
  Ctx* c = (Ctx*)h;
  getLayerIndex(c->m, r, *nl, *top);
  return 0;
}

int orc_start_radius(void* h, double omre, double dll1, int tor, double* rstart) {
  Ctx* c = (Ctx*)h;
  Env env; env.m = &c->m; env.nod = &c->nod; env.s = &c->s; env.omre = omre; env.omim = 0; env.dll1 = dll1; env.nl = 1;
  *rstart = startRadius(env, tor != 0);
  return 0;
}

int orc_initial_values(void* h, double omre, double omim, double dll1, double r, int nl, int kind, double* y,
                       double* htry) {
  Ctx* c = (Ctx*)h;
  Env env; env.m = &c->m; env.nod = &c->nod; env.s = &c->s; env.omre = omre; env.omim = omim; env.dll1 = dll1; env.nl = nl;
  cx ys[6];
  for (int v = 0; v < 6; ++v) ys[v] = C(0.0);
  int rc;
  double rs;
  int nlo;
  if (kind == 0) rc = minorsBottom(env, r, ys, *htry);
  else if (kind == 1) rc = minorsSurface(env, ys, *htry);
  else if (kind == 2) rc = toroidalBottom(env, r, ys, *htry);
  else if (kind == 3) rc = toroidalSurface(env, ys, *htry, rs, nlo);
  else if (kind == 4) rc = minorsOutgoing(env, r, nl, ys);      // full-space start vector at (r, nl)
  else rc = toroidalOutgoing(env, r, nl, ys);
  for (int v = 1; v <= 5; ++v) { y[2 * (v - 1)] = ys[v].re; y[2 * (v - 1) + 1] = ys[v].im; }
  return rc;
}

int orc_sphancf(double zxr, double zxi, double el, int kind, double* out) {
  cx z;
  int rc = sphancf(C(zxr, zxi), el, z, kind);
  out[0] = z.re; out[1] = z.im;
  return rc;
}
int orc_zetl(double zr, double zi, double el, double* out) {
  cx z;
  int rc = zetl(C(zr, zi), el, z);
  out[0] = z.re; out[1] = z.im;
  return rc;
}

static void put(double* dst, const std::vector<cx>& src) {
  for (size_t i = 0; i < src.size(); ++i) { dst[2 * i] = src[i].re; dst[2 * i + 1] = src[i].im; }
}

int orc_toroidal_minors(void* h, double omre, double omim, double dll1, int jsl, int jsr, double eps, double* u,
                        double* d, double* det, int64_t* stats) {
  Ctx* c = (Ctx*)h;
  Env env; env.m = &c->m; env.nod = &c->nod; env.s = &c->s; env.omre = omre; env.omim = omim; env.dll1 = dll1; env.nl = 1;
  ToroidalMinors T; Stats st;
  int rc = toroidalMinors(env, jsl, jsr, eps, T, st, nullptr);
  if (rc == OK) {
    put(u, T.u); put(d, T.d);
    std::vector<cx> dd(T.det.begin() + 1, T.det.end());
    put(det, dd);
  }
  if (stats) for (int i = 0; i < 10; ++i) stats[i] += st.v[i];
  return rc;
}

int orc_spheroidal_minors(void* h, double omre, double omim, double dll1, int jsl, int jsr, double eps, double* u,
                          double* d, double* ulg, double* dlg, double* det, int64_t* stats) {
  Ctx* c = (Ctx*)h;
  Env env; env.m = &c->m; env.nod = &c->nod; env.s = &c->s; env.omre = omre; env.omim = omim; env.dll1 = dll1; env.nl = 1;
  SpheroidalMinors S; Stats st;
  int rc = spheroidalMinors(env, jsl, jsr, eps, S, st, nullptr);
  if (rc == OK) {
    put(u, S.u); put(d, S.d);
    for (int n = 1; n <= c->nod.nnod; ++n) { ulg[n - 1] = S.ulg[n]; dlg[n - 1] = S.dlg[n]; }
    std::vector<cx> dd(S.det.begin() + 1, S.det.end());
    put(det, dd);
  }
  if (stats) for (int i = 0; i < 10; ++i) stats[i] += st.v[i];
  return rc;
}

// computeSpheroidalAdjoint on its own: the solidt system integrated from ra to
// re, starting from ys[4] (complex); adj[nnod][4] complex at the nodes passed
// (zero elsewhere), yend[4] the end value.  Exposed for the pins of the adjoint
// system.
int orc_spheroidal_adjoint(void* h, double omre, double omim, double dll1, const double* ys, double ra, double re,
                           double eps, double* adj, double* yend, int64_t* stats) {
  Ctx* c = (Ctx*)h;
  Env env; env.m = &c->m; env.nod = &c->nod; env.s = &c->s; env.omre = omre; env.omim = omim; env.dll1 = dll1; env.nl = 1;
  cx y[6];
  for (int i = 0; i < 6; ++i) y[i] = C(0.0);
  for (int i = 1; i <= 4; ++i) y[i] = C(ys[2 * (i - 1)], ys[2 * (i - 1) + 1]);
  std::vector<cx> a;
  Stats st;
  int rc = spheroidalAdjoint(env, y, ra, re, eps, a, st);
  if (rc == OK) {
    put(adj, a);
    for (int i = 1; i <= 4; ++i) { yend[2 * (i - 1)] = y[i].re; yend[2 * (i - 1) + 1] = y[i].im; }
  }
  if (stats) for (int i = 0; i < 10; ++i) stats[i] += st.v[i];
  return rc;
}

// one column: gf (nsp, nnod) complex, ujgfsph (4,4,nnod), ujgftor (2,2,nnod); returns detflag in *flag
int orc_gfk_column(void* h, double omre, double omim, double dll1, int jre, int istyp, double eps, double* gf,
                   double* ujs, double* ujt, int* flag, int64_t* stats) {
  Ctx* c = (Ctx*)h;
  Column col; Stats st;
  int rc = gfkColumn(*c, omre, omim, dll1, jre, istyp, eps, col, st);
  if (rc == OK) {
    if (gf) put(gf, col.gf);
    if (ujs) put(ujs, col.ujs);
    if (ujt) put(ujt, col.ujt);
    if (flag) *flag = col.detflag;
  }
  if (stats) for (int i = 0; i < 10; ++i) stats[i] += st.v[i];
  return rc;
}

// All degrees l = lmin .. lmin+nl-1 of one frequency (the driver's iwn loop,
// computeGreenFKSpectraForSynthetics.f90:410-494).  gf layout: complex
// [n][isp][il] (Fortran gf(il, isp, n)).  status[il] = error code per degree.
// The spline tables must have been built for this frequency.
// dll1 = (k R)^2 of degree l (dwn = 0) or of the 1-based wavenumber index l (dwn > 0):
// computeGreenFKSpectraForSynthetics.f90:411-412 with geminiIntegrationEnvironment.f90:71
static double instanceDll1(int l, double dwn, double rearth) {
  double wn = dwn > 0.0 ? (double)(l - 1) * dwn : std::sqrt((double)(l * (l + 1))) / rearth;
  return (wn * rearth) * (wn * rearth);
}

static int gfkInstances(void* h, double omre, double omim, double rearth, int lmin, int nl, double dwn, int jre,
                        int istyp, double eps, int nthreads, double* gf, int* status, int64_t* stats,
                        int dsvstep = 2) {
  Ctx* c = (Ctx*)h;
  int nnod = c->nod.nnod;
  if (jre < 1 || jre > nnod) return ERR_NODEIDX;
  const int njsph = istyp == 1 ? 4 : 2, njtor = istyp == 1 ? 2 : 1, nmask = dsvstep == 1 ? 4 : 2;
  const int nsp = nmask * njsph + (c->nod.state[jre] == 1 ? 0 : (nmask / 2) * njtor);
  Stats total;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
  {
    Stats st;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int il = 0; il < nl; ++il) {
      int l = lmin + il;
      double dll1 = instanceDll1(l, dwn, rearth);
      Column col;
      int rc = gfkColumn(*c, omre, omim, dll1, jre, istyp, eps, col, st, nullptr, nullptr, dsvstep);
      status[il] = rc;
      for (int n = 0; n < nnod; ++n)
        for (int isp = 0; isp < nsp; ++isp) {
          size_t o = ((size_t)n * nsp + isp) * nl + il;
          cx v = (rc == OK) ? col.gf[n * nsp + isp] : C(0.0);
          gf[2 * o] = v.re; gf[2 * o + 1] = v.im;
        }
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    total.merge(st);
  }
  if (stats) for (int i = 0; i < 10; ++i) stats[i] += total.v[i];
  int bad = 0;
  for (int il = 0; il < nl; ++il) if (status[il] != OK) bad = status[il];
  return bad;
}

int orc_gfk_degrees(void* h, double omre, double omim, double rearth, int lmin, int nl, int jre, int istyp,
                    double eps, int nthreads, double* gf, int* status, int64_t* stats) {
  return gfkInstances(h, omre, omim, rearth, lmin, nl, 0.0, jre, istyp, eps, nthreads, gf, status, stats);
}

// the same columns with stresses (STRESSFLAG: dsvstep = 1)
int orc_gfk_degrees_stress(void* h, double omre, double omim, double rearth, int lmin, int nl, int jre, int istyp,
                           double eps, int nthreads, double* gf, int* status, int64_t* stats) {
  return gfkInstances(h, omre, omim, rearth, lmin, nl, 0.0, jre, istyp, eps, nthreads, gf, status, stats, 1);
}

// regional mode: 1-based wavenumber indices iwn = iwnmin .. iwnmin+nwn-1, k = (iwn-1) dwn
int orc_gfk_wavenumbers(void* h, double omre, double omim, double rearth, int iwnmin, int nwn, double dwn, int jre,
                        int istyp, double eps, int nthreads, double* gf, int* status, int64_t* stats) {
  return gfkInstances(h, omre, omim, rearth, iwnmin, nwn, dwn, jre, istyp, eps, nthreads, gf, status, stats);
}

// the regional columns with stresses (STRESSFLAG: dsvstep = 1)
int orc_gfk_wavenumbers_stress(void* h, double omre, double omim, double rearth, int iwnmin, int nwn, double dwn,
                               int jre, int istyp, double eps, int nthreads, double* gf, int* status, int64_t* stats) {
  return gfkInstances(h, omre, omim, rearth, iwnmin, nwn, dwn, jre, istyp, eps, nthreads, gf, status, stats, 1);
}

// Engine test: integrate y' = A y (A constant, nvar <= 5, row-major complex) from x1 to x2.
// All degrees l = lmin .. lmin+nl-1 of one frequency through the multi-source
// driver (computeGreenFKSpectraForASKI.f90:458-590).  gf: complex
// [nsnod+1][nnod][nspmax][nl] (Fortran gfre(nwn, nspmax, nnod, nsnod+1)),
// nspmax = 20 (moment) or 10 (force), zero beyond each source's nsp.
int orc_aski_degrees(void* h, double omre, double omim, double rearth, int lmin, int nl, int nsnod,
                     const int* jsnod, int jre, int istyp, double eps, int nthreads, double* gf, int* status,
                     int64_t* stats) {
  Ctx* c = (Ctx*)h;
  const int nnod = c->nod.nnod;
  const int nspmax = istyp == 1 ? 20 : 10;
  Stats total;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
  {
    Stats st;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int il = 0; il < nl; ++il) {
      const double dll1 = instanceDll1(lmin + il, 0.0, rearth);
      std::vector<cx> col;
      const int rc = askiColumn(*c, omre, omim, dll1, nsnod, jsnod, jre, istyp, eps, nspmax, col, st);
      status[il] = rc;
      for (size_t b = 0; b < (size_t)(nsnod + 1) * nnod * nspmax; ++b) {
        const size_t o = b * nl + il;
        const cx v = rc == OK ? col[b] : C(0.0);
        gf[2 * o] = v.re; gf[2 * o + 1] = v.im;
      }
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    total.merge(st);
  }
  if (stats) for (int i = 0; i < 10; ++i) stats[i] += total.v[i];
  int bad = 0;
  for (int il = 0; il < nl; ++il) if (status[il] != OK) bad = status[il];
  return bad;
}

// system matrix A(r) of the solid (0), toroidal (1) or fluid (2) system in the
// layer holding r, complex [nvar][nvar] row-major (a check of the restatements)
int orc_sysmat_at(void* h, int sys, double omre, double omim, double dll1, double r, double* a) {
  Ctx* c = (Ctx*)h;
  Env env; env.m = &c->m; env.nod = &c->nod; env.s = &c->s; env.omre = omre; env.omim = omim; env.dll1 = dll1;
  int nl, top;
  getLayerIndex(c->m, r, nl, top);
  env.nl = nl;
  SysMat A;
  for (int i = 0; i < 6; ++i) for (int j = 0; j < 6; ++j) A.a[i][j] = C(0.0);
  int nv = 2;
  if (sys == 0) { sysmatSolid(env, r, A); nv = 4; }
  else if (sys == 1) sysmatToroidal(env, r, A);
  else if (sys == 2) sysmatFluid(env, r, A);
  else { sysmatSolidt(env, r, A); nv = 4; }
  for (int i = 1; i <= nv; ++i)
    for (int j = 1; j <= nv; ++j) { a[2 * ((i - 1) * nv + j - 1)] = A.a[i][j].re; a[2 * ((i - 1) * nv + j - 1) + 1] = A.a[i][j].im; }
  return OK;
}

int orc_propagate_const(int nvar, const double* A, double x1, double x2, double htry, double eps, const double* y0,
                        double* y1, int64_t* stats) {
  ConstSys cs; cs.nvar = nvar;
  for (int i = 1; i <= nvar; ++i)
    for (int k = 1; k <= nvar; ++k) cs.a[i][k] = C(A[2 * ((i - 1) * nvar + k - 1)], A[2 * ((i - 1) * nvar + k - 1) + 1]);
  Model m; Nodes N; Spline s;
  m.nlay = 1; m.rb.assign(2, 1e300); N.nnod = 0; N.rnod.assign(1, 0.0);
  Env env; env.m = &m; env.nod = &N; env.s = &s; env.omre = 1; env.omim = 0; env.dll1 = 0; env.nl = 1;
  Stats st;
  Engine eng; eng.sys = -1; eng.nvar = nvar; eng.env = &env; eng.cs = &cs; eng.eps = eps; eng.st = &st;
  cx ys[6];
  for (int v = 1; v <= nvar; ++v) ys[v] = C(y0[2 * (v - 1)], y0[2 * (v - 1) + 1]);
  double hdid = htry;
  int rc = integrationStep(eng, hdid, x1, x2, ys);
  for (int v = 1; v <= nvar; ++v) { y1[2 * (v - 1)] = ys[v].re; y1[2 * (v - 1) + 1] = ys[v].im; }
  if (stats) for (int i = 0; i < 10; ++i) stats[i] += st.v[i];
  return rc;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"

extern "C" {
double orc_exp10(double x) { return exp10Det(x); }

// reciprocity relations for tests: bgin (4,4) column-major complex -> bgout (4,4)
int orc_reciprocity_sph(double elp1, double re, double r, int state_node, const double* bgin, double* bgout) {
  std::vector<orc::cx> in(16);
  for (int i = 0; i < 16; ++i) in[i] = orc::C(bgin[2 * i], bgin[2 * i + 1]);
  orc::cx out[5][5];
  orc::reciprocitySph(elp1, re, r, state_node, in.data(), out);
  for (int i = 1; i <= 4; ++i)
    for (int k = 1; k <= 4; ++k) { bgout[2 * ((k - 1) * 4 + i - 1)] = out[i][k].re; bgout[2 * ((k - 1) * 4 + i - 1) + 1] = out[i][k].im; }
  return 0;
}
// complexElasticVTIConstants for tests
int orc_vti(double fratio, double ro, double alfv, double alfh, double betv, double beth, double eta, double qkinv,
            double qminv, double* z14) {
  orc::cx z[8];
  orc::vtiConstants(fratio, ro, alfv, alfh, betv, beth, eta, qkinv, qminv, z);
  for (int k = 1; k <= 7; ++k) { z14[2 * (k - 1)] = z[k].re; z14[2 * (k - 1) + 1] = z[k].im; }
  return 0;
}
}

extern "C" {
// minors only (toroidal + spheroidal) for l = lmin .. lmin+nl-1 at the current
// frequency: the unit of the solves/s metric.  Results are discarded.
static int minorsInstances(void* h, double omre, double omim, double rearth, int lmin, int nl, double dwn, int jsl,
                           int jsr, double eps, int nthreads, int* status, int64_t* stats) {
  orc::Ctx* c = (orc::Ctx*)h;
  orc::Stats total;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
  {
    orc::Stats st;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int il = 0; il < nl; ++il) {
      int l = lmin + il;
      double dll1 = instanceDll1(l, dwn, rearth);
      orc::Env env; env.m = &c->m; env.nod = &c->nod; env.s = &c->s; env.omre = omre; env.omim = omim; env.dll1 = dll1; env.nl = 1;
      orc::ToroidalMinors T; orc::SpheroidalMinors S;
      int rc = orc::toroidalMinors(env, jsl, jsr, eps, T, st, nullptr);
      if (rc == orc::OK) rc = orc::spheroidalMinors(env, jsl, jsr, eps, S, st, nullptr);
      status[il] = rc;
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    total.merge(st);
  }
  if (stats) for (int i = 0; i < 10; ++i) stats[i] += total.v[i];
  return 0;
}

int orc_minors_degrees(void* h, double omre, double omim, double rearth, int lmin, int nl, int jsl, int jsr,
                       double eps, int nthreads, int* status, int64_t* stats) {
  return minorsInstances(h, omre, omim, rearth, lmin, nl, 0.0, jsl, jsr, eps, nthreads, status, stats);
}

int orc_minors_wavenumbers(void* h, double omre, double omim, double rearth, int iwnmin, int nwn, double dwn, int jsl,
                           int jsr, double eps, int nthreads, int* status, int64_t* stats) {
  return minorsInstances(h, omre, omim, rearth, iwnmin, nwn, dwn, jsl, jsr, eps, nthreads, status, stats);
}
}
