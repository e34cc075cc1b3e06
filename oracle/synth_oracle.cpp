// CPU restatement of the synthetics half of the Green-FK path: Legendre
// functions, moment-tensor harmonic degree sums, the single-frequency
// displacement combination, time shift / velocity / ZNE rotation and the
// frequency-to-time transform.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU kernels are
// compared against; no product code path may call it.
//
// Each function cites the reference routine it restates.  Arithmetic is FP64
// throughout, in the reference's operation order where that is cheap to keep.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace syn {

using cd = std::complex<double>;
// mc_pid (f90/mathConstants.f90:34): the literal has no d0 exponent, so the
// reference's "double precision" pi is the FP32 value 3.1415927410125732
static const double kPi = 3.1415927410125732;

// Unnormalised associated Legendre functions P_l^m(cos theta), m = 0..2, without
// the Condon-Shortley sign, and their theta derivatives / ratios to sin(theta).
// Restates computeLegendreFunctions, f90/legendreFunctions.f90:48-157.
// Layout: plm, dpdt, d2pdt2: [3][lmax+1]; pdst, dtpdst: [3][lmax+1] (m = 0 row unused).
void legendre(int lmax, double theta, double* plm, double* dpdt, double* pdst, double* d2pdt2, double* dtpdst) {
  const int L1 = lmax + 1;
  for (int i = 0; i < 3 * L1; ++i) plm[i] = dpdt[i] = pdst[i] = d2pdt2[i] = dtpdst[i] = 0.0;
  auto P = [&](int l, int m) -> double& { return plm[m * L1 + l]; };
  auto D = [&](int l, int m) -> double& { return dpdt[m * L1 + l]; };
  auto S = [&](int l, int m) -> double& { return pdst[m * L1 + l]; };
  auto D2 = [&](int l, int m) -> double& { return d2pdt2[m * L1 + l]; };
  auto DS = [&](int l, int m) -> double& { return dtpdst[m * L1 + l]; };
  const double ct = std::cos(theta), st = std::sin(theta);
  const bool poles = !(st > 0.0);
  // legendreFunctions.f90:84-95: P_m^m = (2m-1)!! sin^m, upward recurrence in l
  for (int m = 0; m <= 2 && m <= lmax; ++m) {
    double pmm = 1.0, odd = 1.0;
    for (int j = 1; j <= m; ++j) { pmm = pmm * odd * st; odd = odd + 2.0; }
    P(m, m) = pmm;
    if (m + 1 <= lmax) P(m + 1, m) = ct * (2 * m + 1) * P(m, m);
    for (int l = m + 2; l <= lmax; ++l)
      P(l, m) = (ct * double(2 * l - 1) * P(l - 1, m) - (l + m - 1) * P(l - 2, m)) / double(l - m);
  }
  if (!poles) {
    // legendreFunctions.f90:99-112
    const double rst = 1.0 / st;
    for (int m = 0; m <= 2 && m <= lmax; ++m) {
      D(m, m) = m * ct * P(m, m) * rst;
      for (int l = m + 1; l <= lmax; ++l) D(l, m) = (l * ct * P(l, m) - (l + m) * P(l - 1, m)) * rst;
    }
    for (int m = 1; m <= 2; ++m)
      for (int l = m; l <= lmax; ++l) S(l, m) = P(l, m) * rst;
    // legendreFunctions.f90:128-141
    for (int m = 0; m <= 2 && m <= lmax; ++m)
      for (int l = m; l <= lmax; ++l) D2(l, m) = -ct * rst * D(l, m) + (m * m * (rst * rst) - double(l * (l + 1))) * P(l, m);
    for (int m = 1; m <= 2; ++m)
      for (int l = m; l <= lmax; ++l) DS(l, m) = rst * (D(l, m) - ct * rst * P(l, m));
  } else {
    // legendreFunctions.f90:113-127, 142-155 (pole limits; sign(ct)^l)
    const double s = ct < 0.0 ? -1.0 : 1.0;
    for (int l = 0; l <= lmax; ++l) {
      const double sl = (l % 2) ? s : 1.0, sl1 = ((l + 1) % 2) ? s : 1.0;
      const double ll1 = double(l) * (l + 1);
      const double q = double(l + 2) * (l + 1) * l * (l - 1);   // 64-bit: the reference's default integer overflows for l > 215
      D(l, 0) = 0.0; D(l, 1) = 0.5 * ll1 * sl; D(l, 2) = 0.0;
      D2(l, 0) = 0.5 * ll1 * sl; D2(l, 1) = 0.0; D2(l, 2) = 0.25 * q * sl;
      if (l >= 1) {
        S(l, 1) = 0.5 * ll1 * sl1; S(l, 2) = 0.0;
        DS(l, 1) = 0.0; DS(l, 2) = 0.125 * q * sl1;
      }
    }
  }
}

// cos^2 taper over the last tapfrac of the degrees 0..lmax.
// Restates taperHarmonicDegreeSums, f90/harmonicDegreeSums.f90:356-369.
void taper(int lmax, double tapfrac, double* w) {
  int jt = (int)std::lround(lmax * (1.0 - tapfrac));
  if (jt > lmax) jt = lmax;
  if (jt < 1) jt = 1;
  const double c1 = lmax * tapfrac * 2.0 / kPi;
  const double c2 = (1.0 - tapfrac) / tapfrac * 0.5 * kPi;
  for (int j = 0; j <= lmax; ++j) {
    if (j < jt) { w[j] = 1.0; continue; }
    const double c = std::cos(double(j) / c1 - c2);
    w[j] = c * c;
  }
}

// (basis row, Green spectrum index, factor) of the 14 moment-tensor sums.
// Restates the case table of momentHarmonicDegreeSums
// (f90/harmonicDegreeSums.f90:251-307) and the (component, jump) -> spectrum
// offset map (:44-48 with positionDataGreenFKSpectra, f90/greenFKSpectra.f90:204-212
// for dsvmask (U, V, W) and 4 spheroidal / 2 toroidal jumps).
// basis: 0 P_l0, 1 P_l1, 2 P_l2, 3 dP_l0, 4 dP_l1, 5 dP_l2, 6 P_l1/sin, 7 P_l2/sin
struct SumDef { int basis, isp, lmin; double fac; bool over_ll1; };
static const SumDef kSums[14] = {
    {0, 0, 0, 1.0, false}, {0, 1, 0, 1.0, false}, {1, 2, 1, 2.0, true}, {2, 3, 2, 2.0, true},
    {3, 4, 1, 1.0, false}, {3, 5, 1, 1.0, false}, {4, 6, 1, 2.0, true}, {5, 7, 2, 2.0, true},
    {6, 8, 1, 2.0, true},  {7, 9, 2, 4.0, true},  {4, 8, 1, 2.0, true}, {5, 9, 2, -2.0, true},
    {6, 6, 1, 2.0, true},  {7, 7, 2, -4.0, true}};

// (basis row, Green spectrum index, factor) of the 7 force sums (kint 1..7:
// displacement only).  Restates the case table of forceHarmonicDegreeSums
// (f90/harmonicDegreeSums.f90:115-143) and the (component, jump) -> spectrum
// map (:34-36 with positionDataGreenFKSpectra, f90/greenFKSpectra.f90:204-212,
// for dsvmask (U, V, W) and 2 spheroidal / 1 toroidal jumps: U1 U2 V1 V2 W1).
static const SumDef kForceSums[7] = {
    {0, 0, 0, 1.0, false},  {1, 1, 1, -2.0, true}, {3, 2, 1, 1.0, false}, {4, 3, 1, -2.0, true},
    {6, 4, 1, -2.0, true},  {4, 4, 1, 2.0, true},  {6, 3, 1, 2.0, true}};

// Harmonic degree sums of one source type at one receiver for each frequency.
// gr, gi: [nsp][nkf] FP32 spectra of one node, k over frequencies nf1..nf2 with
// nwn[jf-1] degrees each (l fastest) — the layout readDataNodeGreenFKSpectra
// returns; hdsum: [nf][nsum] complex.
// Restates the frequency loop of sphericalHarmonicsDisplacementSpectra
// (f90/displacementSpectra.f90:49-106) with momentHarmonicDegreeSums or
// forceHarmonicDegreeSums.
static void degree_sums(const SumDef* defs, int nsum, int nf1, int nf2, const int* nwn, long nkf, const float* gr,
                        const float* gi, double theta, double tapfrac, cd* hdsum) {
  int lmax = 0;
  for (int jf = nf1; jf <= nf2; ++jf) lmax = std::max(lmax, nwn[jf - 1] - 1);
  const int L1 = lmax + 1;
  std::vector<double> plm(3 * L1), dpdt(3 * L1), pdst(3 * L1), d2(3 * L1), ds(3 * L1), w(L1);
  legendre(lmax, theta, plm.data(), dpdt.data(), pdst.data(), d2.data(), ds.data());
  const double* rows[8] = {&plm[0], &plm[L1], &plm[2 * L1], &dpdt[0], &dpdt[L1], &dpdt[2 * L1], &pdst[L1], &pdst[2 * L1]};
  const double fourpi = 4.0 * kPi;
  long off = 0;
  for (int jf = nf1; jf <= nf2; ++jf) {
    const int n = nwn[jf - 1];
    const int lm = n - 1;
    taper(lm, tapfrac, w.data());
    for (int s = 0; s < nsum; ++s) {
      const SumDef& d = defs[s];
      cd z = 0.0;
      for (int l = d.lmin; l <= lm; ++l) {
        double b = rows[d.basis][l];
        if (d.over_ll1) b = b * d.fac / double(l * (l + 1));
        const cd g(gr[(long)d.isp * nkf + off + l], gi[(long)d.isp * nkf + off + l]);
        z += b * g * (2.0 * l + 1.0) / fourpi * w[l];
      }
      hdsum[(long)(jf - nf1) * nsum + s] = z;
    }
    off += n;
  }
}

// The 14 moment harmonic degree sums (kint 1..14); hdsum: [nf][14] complex.
void moment_sums(int nf1, int nf2, const int* nwn, long nkf, const float* gr, const float* gi, double theta,
                 double tapfrac, cd* hdsum) {
  degree_sums(kSums, 14, nf1, nf2, nwn, nkf, gr, gi, theta, tapfrac, hdsum);
}

// The 7 force harmonic degree sums (kint 1..7); hdsum: [nf][7] complex.
void force_sums(int nf1, int nf2, const int* nwn, long nkf, const float* gr, const float* gi, double theta,
                double tapfrac, cd* hdsum) {
  degree_sums(kForceSums, 7, nf1, nf2, nwn, nkf, gr, gi, theta, tapfrac, hdsum);
}

// Displacement (r, theta, phi) at one frequency from the 14 sums and the moment
// tensor mt = (Mrr, Mtt, Mff, Mrt, Mrf, Mtf).
// Restates momentSingleFrequencyDisplacement, f90/singleFrequencyDisplacement.f90:108-125.
void moment_displacement(double phi, const double* mt, const cd* w, cd* zsp) {
  const double sp = std::sin(phi), cp = std::cos(phi), s2 = std::sin(2.0 * phi), c2 = std::cos(2.0 * phi);
  const double a = mt[3] * cp + mt[4] * sp, b = (mt[2] - mt[1]) * c2 - 2.0 * mt[5] * s2;
  const double c = mt[4] * cp - mt[3] * sp, d = (mt[2] - mt[1]) * s2 + 2.0 * mt[5] * c2;
  zsp[0] = w[0] * mt[0] + w[1] * (mt[1] + mt[2]) + w[2] * a + w[3] * b;
  zsp[1] = w[4] * mt[0] + w[5] * (mt[1] + mt[2]) + (w[6] + w[8]) * a + (w[7] + w[9]) * b;
  zsp[2] = (w[10] + w[12]) * c + (w[11] + w[13]) * d;
}

// Displacement (r, theta, phi) at one frequency from the 7 force sums and the
// force f = (Fr, Ft, Ff).
// Restates forceSingleFrequencyDisplacement, f90/singleFrequencyDisplacement.f90:35-46.
void force_displacement(double phi, const double* f, const cd* w, cd* zsp) {
  const double sp = std::sin(phi), cp = std::cos(phi);
  const double a = f[1] * cp + f[2] * sp, b = f[1] * sp - f[2] * cp;
  zsp[0] = w[0] * f[0] + w[1] * a;
  zsp[1] = w[2] * f[0] + (w[3] + w[4]) * a;
  zsp[2] = (w[5] + w[6]) * b;
}

// ---------------------------------------------------------------------------
// Strain spectra (spectra with radial derivatives: derivflag = 1, the files of
// the multi-source driver).  Restates momentHarmonicDegreeSums /
// forceHarmonicDegreeSums for every kint (f90/harmonicDegreeSums.f90:95-351),
// getIndexArrayHarmonicDegreeSums (:375-420), the sum -> (component, jump)
// maps (:34-46) with positionDataGreenFKSpectra (f90/greenFKSpectra.f90:
// 204-212), the loop of sphericalHarmonicsDisplacementSpectra with strainsp
// (f90/displacementSpectra.f90:49-106, 200-228) and the routines of
// f90/singleFrequencyDisplacement.f90:35-278.
// ---------------------------------------------------------------------------
// the reference's (component, jump) of each sum: moment kint 1..46, force 1..23
static const int kMomentMap[46][2] = {
    {1, 1}, {1, 2}, {1, 3}, {1, 4}, {3, 1}, {3, 2}, {3, 3}, {3, 4}, {7, 1}, {7, 2}, {7, 1}, {7, 2}, {3, 3}, {3, 4},
    {5, 1}, {5, 2}, {5, 3}, {5, 4}, {6, 1}, {6, 2}, {6, 3}, {6, 4}, {9, 1}, {9, 2}, {9, 1}, {9, 2}, {6, 3}, {6, 4},
    {1, 1}, {1, 2}, {1, 3}, {1, 4}, {3, 1}, {3, 2}, {3, 3}, {3, 4}, {7, 1}, {7, 2}, {7, 1}, {7, 2}, {3, 3}, {3, 4},
    {2, 1}, {2, 2}, {2, 3}, {2, 4}};
static const int kForceMap[23][2] = {
    {1, 1}, {1, 2}, {3, 1}, {3, 2}, {7, 1}, {7, 1}, {3, 2}, {5, 1}, {5, 2}, {6, 1}, {6, 2}, {9, 1}, {9, 1}, {6, 2},
    {1, 1}, {1, 2}, {3, 1}, {3, 2}, {7, 1}, {7, 1}, {3, 2}, {2, 1}, {2, 2}};
// Legendre rows: 0 P_l0, 1 P_l1, 2 P_l2, 3 dP_l0, 4 dP_l1, 5 dP_l2, 6 P_l1/sin, 7 P_l2/sin,
// 8 d2P_l0, 9 d2P_l1, 10 d2P_l2, 11 d(P_l1/sin), 12 d(P_l2/sin); -1: the sum is not in the case table (0)
struct KintDef { int basis, lmin; double fac; bool over_ll1; };
static KintDef momentKint(int k) {
  switch (k) {
    case 1: case 2: case 15: case 16: return {0, 0, 1.0, false};
    case 3: case 17: return {1, 1, 2.0, true};
    case 4: case 18: return {2, 2, 2.0, true};
    case 5: case 19: case 29: case 6: case 20: case 30: return {3, 1, 1.0, false};
    case 7: case 21: case 31: return {4, 1, 2.0, true};
    case 8: case 22: case 32: return {5, 2, 2.0, true};
    case 9: case 23: return {6, 1, 2.0, true};
    case 10: case 24: return {7, 2, 4.0, true};
    case 11: case 25: return {4, 1, 2.0, true};
    case 12: case 26: return {5, 2, -2.0, true};
    case 13: case 27: return {6, 1, 2.0, true};
    case 14: case 28: return {7, 2, -4.0, true};
    case 33: case 34: return {8, 1, 1.0, false};
    case 35: return {9, 1, 2.0, true};
    case 36: return {10, 2, 2.0, true};
    case 37: return {11, 1, 2.0, true};
    case 38: return {12, 2, 4.0, true};
    case 39: return {9, 1, 2.0, true};
    case 40: return {10, 2, -2.0, true};
    case 41: return {11, 1, 2.0, true};
    case 42: return {12, 2, -4.0, true};
    default: return {-1, 0, 0.0, false};
  }
}
static KintDef forceKint(int k) {
  switch (k) {
    case 1: case 8: return {0, 0, 1.0, false};
    case 2: case 9: return {1, 1, -2.0, true};
    case 3: case 10: case 15: return {3, 1, 1.0, false};
    case 4: case 11: case 16: return {4, 1, -2.0, true};
    case 5: case 12: return {6, 1, -2.0, true};
    case 6: case 13: return {4, 1, 2.0, true};
    case 7: case 14: return {6, 1, 2.0, true};
    case 17: return {8, 1, 1.0, false};
    case 18: return {9, 1, -2.0, true};
    case 19: return {11, 1, -2.0, true};
    case 20: return {9, 1, 2.0, true};
    case 21: return {11, 1, 2.0, true};
    default: return {-1, 0, 0.0, false};
  }
}
// 0-based spectrum offset of (component isp, jump jsp) for dsvmask (U, V, UP, VP, W, WP)
// = (1,0,1,0,1,1,1,0,1) (greenFKSpectra.f90:109-111 with derivflag = 1, dsvstep = 2)
static int positionData(int isp, int jsp, int njsph, int njtor) {
  static const int mask[10] = {0, 1, 0, 1, 0, 1, 1, 1, 0, 1};
  int off = 0;
  if (isp <= 6) {
    for (int k = 1; k < isp; ++k) off += mask[k] * njsph;
  } else {
    for (int k = 1; k <= 6; ++k) off += mask[k] * njsph;
    for (int k = 7; k < isp; ++k) off += mask[k] * njtor;
  }
  return off + jsp - 1;
}

// All degree sums of one receiver (kint 1..42 moment, 1..21 force: derivatives
// without stresses, getIndexArrayHarmonicDegreeSums with dsvmask(2) = 0,
// dsvmask(5) = 1); hdsum: [nf][nsum] complex
void strain_sums(int istyp, int nf1, int nf2, const int* nwn, long nkf, const float* gr, const float* gi,
                 double theta, double tapfrac, cd* hdsum) {
  const int nsum = istyp == 1 ? 42 : 21, njsph = istyp == 1 ? 4 : 2, njtor = istyp == 1 ? 2 : 1;
  int lmax = 0;
  for (int jf = nf1; jf <= nf2; ++jf) lmax = std::max(lmax, nwn[jf - 1] - 1);
  const int L1 = lmax + 1;
  std::vector<double> plm(3 * L1), dpdt(3 * L1), pdst(3 * L1), d2(3 * L1), ds(3 * L1), w(L1);
  legendre(lmax, theta, plm.data(), dpdt.data(), pdst.data(), d2.data(), ds.data());
  const double* rows[13] = {&plm[0], &plm[L1], &plm[2 * L1], &dpdt[0], &dpdt[L1], &dpdt[2 * L1], &pdst[L1],
                            &pdst[2 * L1], &d2[0], &d2[L1], &d2[2 * L1], &ds[L1], &ds[2 * L1]};
  const double fourpi = 4.0 * kPi;
  long off = 0;
  for (int jf = nf1; jf <= nf2; ++jf) {
    const int n = nwn[jf - 1];
    const int lm = n - 1;
    taper(lm, tapfrac, w.data());
    for (int k = 1; k <= nsum; ++k) {
      const KintDef d = istyp == 1 ? momentKint(k) : forceKint(k);
      const int* cj = istyp == 1 ? kMomentMap[k - 1] : kForceMap[k - 1];
      const int isp = positionData(cj[0], cj[1], njsph, njtor);
      cd z = 0.0;
      if (d.basis >= 0)
        for (int l = d.lmin; l <= lm; ++l) {
          double b = rows[d.basis][l];
          if (d.over_ll1) b = b * d.fac / double(l * (l + 1));
          const cd g(gr[(long)isp * nkf + off + l], gi[(long)isp * nkf + off + l]);
          z += b * g * (2.0 * l + 1.0) / fourpi * w[l];
        }
      hdsum[(long)(jf - nf1) * nsum + k - 1] = z;
    }
    off += n;
  }
}

// displacement and its derivatives from the sums: d = 0 the displacement
// (moment/forceSingleFrequencyDisplacement), 1 d/dr (RadDeriv), 2 1/r d/dtheta
// (ThetaDeriv), 3 1/(r sin theta) d/dphi (PhiDeriv); f90/singleFrequencyDisplacement.f90:35-194
void strain_parts(int istyp, double re, double delta, double phi, const double* fm, const cd* w, cd u[4][3]) {
  const double sp = std::sin(phi), cp = std::cos(phi);
  if (istyp == 1) {
    const double s2 = std::sin(2. * phi), c2 = std::cos(2. * phi);
    const double a = fm[3] * cp + fm[4] * sp, b = (fm[2] - fm[1]) * c2 - 2. * fm[5] * s2;
    const double c = fm[4] * cp - fm[3] * sp, d = (fm[2] - fm[1]) * s2 + 2. * fm[5] * c2;
    for (int part = 0; part < 3; ++part) {
      const cd* h = w + 14 * part;
      u[part][0] = h[0] * fm[0] + h[1] * (fm[1] + fm[2]) + h[2] * a + h[3] * b;
      u[part][1] = h[4] * fm[0] + h[5] * (fm[1] + fm[2]) + (h[6] + h[8]) * a + (h[7] + h[9]) * b;
      u[part][2] = (h[10] + h[12]) * c + (h[11] + h[13]) * d;
    }
    // kint 1..14, then 15..28 (d/dr), 29..42 (d/dtheta) in the loop above: parts 0, 1, 2
    for (int c3 = 0; c3 < 3; ++c3) u[2][c3] = u[2][c3] / re;
    const double rst = re * std::sin(delta);
    const double ap = -fm[3] * sp + fm[4] * cp, bp = -2. * (fm[2] - fm[1]) * s2 - 4. * fm[5] * c2;
    const double cq = -fm[4] * sp - fm[3] * cp, dq = +2. * (fm[2] - fm[1]) * c2 - 4. * fm[5] * s2;
    u[3][0] = w[2] * ap + w[3] * bp;
    u[3][1] = (w[6] + w[8]) * ap + (w[7] + w[9]) * bp;
    u[3][2] = (w[10] + w[12]) * cq + (w[11] + w[13]) * dq;
    for (int c3 = 0; c3 < 3; ++c3) u[3][c3] = u[3][c3] / rst;
  } else {
    const double a = fm[1] * cp + fm[2] * sp, b = fm[1] * sp - fm[2] * cp;
    for (int part = 0; part < 3; ++part) {
      const cd* h = w + 7 * part;
      u[part][0] = h[0] * fm[0] + h[1] * a;
      u[part][1] = h[2] * fm[0] + (h[3] + h[4]) * a;
      u[part][2] = (h[5] + h[6]) * b;
    }
    for (int c3 = 0; c3 < 3; ++c3) u[2][c3] = u[2][c3] / re;
    const double rst = re * std::sin(delta);
    const double ap = -fm[1] * sp + fm[2] * cp, bp = fm[1] * cp + fm[2] * sp;
    u[3][0] = w[1] * ap;
    u[3][1] = (w[3] + w[4]) * ap;
    u[3][2] = (w[5] + w[6]) * bp;
    for (int c3 = 0; c3 < 3; ++c3) u[3][c3] = u[3][c3] / rst;
  }
}

// sphericalStrainSingleFrequencyDisplacement (:223-241): strain[3][3] from
// u, d/dr u, 1/r d/dtheta u, 1/(r sin) d/dphi u
void spherical_strain(const cd u[4][3], double r, double delta, cd e[3][3]) {
  const double rtandel = r * std::tan(delta);
  const cd *z = u[0], *zr = u[1], *zt = u[2], *zf = u[3];
  e[0][0] = zr[0];
  e[1][1] = zt[1] + z[0] / r;
  e[2][2] = zf[2] + z[0] / r + z[1] / rtandel;
  e[0][1] = 0.5 * (zt[0] + zr[1] - z[1] / r);
  e[0][2] = 0.5 * (zf[0] + zr[2] - z[2] / r);
  e[1][2] = 0.5 * (zf[1] + zt[2] - z[2] / rtandel);
  e[2][1] = e[1][2];
  e[2][0] = e[0][2];
  e[1][0] = e[0][1];
}

// sphericalStressSingleFrequencyDisplacement (:253-266) with elcon = (A, C, F, L, N, ...)
void spherical_stress(const cd e[3][3], const cd* el, cd t[3][3]) {
  t[0][0] = el[1] * e[0][0] + el[2] * (e[1][1] + e[2][2]);
  t[1][1] = el[2] * e[0][0] + el[0] * (e[1][1] + e[2][2]) - 2.0 * el[4] * e[2][2];
  t[2][2] = el[2] * e[0][0] + el[0] * (e[1][1] + e[2][2]) - 2.0 * el[4] * e[1][1];
  t[0][1] = 2.0 * el[3] * e[0][1];
  t[0][2] = 2.0 * el[3] * e[0][2];
  t[1][2] = 2.0 * el[4] * e[1][2];
  t[1][0] = t[0][1];
  t[2][0] = t[0][2];
  t[2][1] = t[1][2];
}

// In-place radix-2 complex FFT, e^{+i...} for is > 0, no normalisation.
// Restates doubleFastFourierTransform (f90/fourierTransform.f90:118-192):
// decimation in frequency with bit-reversed reordering at the end.
void fft(std::vector<double>& x, std::vector<double>& y, int n, int is) {
  const int ng = 1 << n;
  const double piz = 2.0 * kPi;
  for (int m = 1; m <= n; ++m) {
    const int nar = 1 << (m - 1);
    const int lar = ng / nar, larh = lar / 2;
    const double alpha = piz / double(is > 0 ? lar : -lar);
    for (int jr = 1; jr <= larh; ++jr) {
      const double beta = alpha * double(jr - 1);
      const double ec = std::cos(beta), es = std::sin(beta);
      int ja = jr - lar;
      for (int nr = 1; nr <= nar; ++nr) {
        ja += lar;
        const int jb = ja + larh;
        const double zx = x[ja - 1] - x[jb - 1], zy = y[ja - 1] - y[jb - 1];
        x[ja - 1] += x[jb - 1];
        y[ja - 1] += y[jb - 1];
        x[jb - 1] = zx * ec - zy * es;
        y[jb - 1] = zx * es + zy * ec;
      }
    }
  }
  for (int j = 0; j < ng; ++j) {
    int k = 0;
    for (int b = 0; b < n; ++b) if (j & (1 << b)) k |= 1 << (n - 1 - b);
    if (j < k) { std::swap(x[j], x[k]); std::swap(y[j], y[k]); }
  }
}

// Spectrum sp[jf-1], jf = 1..nf2 (entries below nf1 ignored) -> nsamp real
// samples.  Restates transformFrequencyTime (f90/frequencyTime.f90:55-113),
// including its decimation branch: when nf2 needs more than nsamp samples the
// transform runs at ns = 2^(ceil(log2 nf2)+1) samples of dts = 1/(df ns) and
// every ndec-th sample is kept, ndec = nint(dt/dts); the damping factor then
// advances by exp(sigma dts) per kept sample, as written there (:104-109).
// Returns 0, or 2 for a non-integer ratio dt/dts, 3 if dt does not match nsamp,
// 4 if the kept samples run past the transform.
int freq_time(int nf1, int nf2, double df, double sigma, const cd* sp, int nsamp, double dt, float* ts) {
  // the reference takes this log in single precision (log of real(nf2))
  int nslog = (int)std::ceil((double)std::log((float)nf2) / std::log(2.0)) + 1;
  int ns = 1 << nslog;
  double dts = dt;
  int ndec = 1;
  if (ns <= nsamp) {
    ns = nsamp;
    nslog = (int)std::ceil(-std::log(df * dt) / std::log(2.0));
    if ((1 << nslog) != ns) return 3;
  } else {
    dts = 1.0 / (df * ns);
    ndec = (int)std::lround(dt / dts);
    if (std::fabs(dt - ndec * dts) > 1.e-5) return 2;
    // the reference indexes x(ndec*(i-1)+1) up to i = nsamp without a check
    if (ndec < 1 || (size_t)ndec * (nsamp - 1) >= (size_t)ns) return 4;
  }
  std::vector<double> x(ns, 0.0), y(ns, 0.0);
  for (int jf = nf1; jf <= nf2; ++jf) { x[jf - 1] = sp[jf - 1].real(); y[jf - 1] = sp[jf - 1].imag(); }
  for (int i = ns / 2 + 2; i <= ns; ++i) { x[i - 1] = x[ns - i + 1]; y[i - 1] = -y[ns - i + 1]; }
  fft(x, y, nslog, +1);
  double fac = 1.0;
  const double expo = std::exp(sigma * dts);
  for (int i = 0; i < nsamp; ++i) { ts[i] = (float)(x[(size_t)ndec * i] * df * fac); fac = fac * expo; }
  return 0;
}

// Band-pass event filter: a high-pass Butterworth filter of order n1 and corner
// f1 times a low-pass one of order n2 and corner f2, at the complex frequencies
// f - i sigma / (2 pi), f = (jf-1) df, for jf = nf1..nf2; h[jf-1] (zero below nf1).
// Restates computeEventfilter's BUTTERWORTH_BANDPASS case (f90/eventFilter.f90:40-83)
// with highPassComplexOmegaButterworthFilter and lowPassComplexOmegaButterworthFilter
// (f90/butterworthFilter.f90:49-96), including their running sum of the frequency.
void butterworth_bandpass(int n1, double f1, int n2, double f2, int nf1, int nf2, double df, double sigma, cd* h) {
  const int nf = nf2 - nf1 + 1;
  const double kpi = kPi, two_pi = 2.0 * kPi;        // mc_pid, mc_two_pid (FP32 pi, as everywhere here)
  std::vector<cd> hh(nf, 1.0), hl(nf, 1.0);
  const cd i1(0.0, 1.0);
  for (int pass = 0; pass < 2; ++pass) {
    const int nord = pass == 0 ? n1 : n2;
    const double fc = pass == 0 ? f1 : f2;
    std::vector<cd>& hv = pass == 0 ? hh : hl;
    const cd zfna = (cd((nf1 - 1) * df, 0.0) - i1 * sigma / two_pi) / fc;
    for (int j = 1; j <= nord; ++j) {
      cd zfn = zfna;
      const cd ep = std::exp(i1 * kpi * (double(j) - 0.5) / double(nord));
      for (int i = 0; i < nf; ++i) {
        hv[i] = pass == 0 ? hv[i] * zfn / (zfn - ep) : hv[i] * (-i1) / (zfn - ep);
        zfn = zfn + df / fc;
      }
    }
  }
  for (int jf = 1; jf <= nf2; ++jf) h[jf - 1] = jf < nf1 ? cd(0.0) : hh[jf - nf1] * hl[jf - nf1];
}

// ---- regional (Bessel) branch ----

// cos^2 taper over the last tapfrac of the wavenumbers 1..nwn (1-based j, k = (j-1) dwn).
// Restates taperWavenumberIntegrals, f90/wavenumberIntegrals.f90:499-512.
void taper_wn(int nwn, double dwn, double tapfrac, double* w) {
  int jt = (int)std::lround(nwn * (1.0 - tapfrac));
  if (jt > nwn) jt = nwn;
  if (jt < 1) jt = 1;
  const double c1 = (nwn - 1) * dwn * tapfrac * 2.0 / kPi;
  const double c2 = (1.0 - tapfrac) / tapfrac * 0.5 * kPi;
  for (int j = 1; j <= nwn; ++j) {
    if (j < jt) { w[j - 1] = 1.0; continue; }
    const double c = std::cos((j - 1) * dwn / c1 - c2);
    w[j - 1] = c * c;
  }
}

// Kernel of moment wavenumber integral kint (1..42) at wavenumber k for Bessel
// (or Neumann) values b0, b1, b2 of order 0..2 at k x: kint 15..28 take the
// kernels of 1..14 (on the radial-derivative spectra), 29..42 those of the
// x-derivatives.  Restates the case tables of momentWavenumberIntegrals,
// f90/wavenumberIntegrals.f90:353-373 (and :379-399); kint 43..46 (stresses) are not taken.
static double wn_kernel(int kint, double k, double x, double re, double b0, double b1, double b2) {
  const double kx = k * x, kr = k * re;
  switch (kint) {
    case 1: case 2: case 15: case 16: return k * b0;
    case 3: case 17: return 2.0 / re * b1;
    case 4: case 18: return k * 2.0 * b2;
    case 5: case 6: case 19: case 20: case 29: case 30: return k * (-k * re) * b1;
    case 7: case 11: case 21: case 25: case 31: return k * 2.0 * (b0 - b1 / kx);
    case 8: case 22: case 32: return k * 2.0 * k * re * (b1 - 2.0 * b2 / kx);
    case 9: case 13: case 23: case 27: return k * 2.0 * b1 / kx;
    case 10: case 24: return k * 4.0 * k * re * b2 / kx;
    case 12: case 26: return -k * 2.0 * k * re * (b1 - 2.0 * b2 / kx);
    case 14: case 28: return -k * 4.0 * k * re * b2 / kx;
    case 33: case 34: return -k * kr * kr * (b0 - b1 / kx);
    case 35: case 39: return -k * 2.0 * kr * (b1 - b2 / kx);
    case 36: return k * 2.0 * kr * kr * ((6.0 / (kx * kx) - 1.0) * b2 - b1 / kx);
    case 37: case 41: return -k * 2.0 * kr * b2 / kx;
    case 38: return k * 4.0 * kr * kr * (-3.0 * b2 / (kx * kx) + b1 / kx);
    case 40: return -k * 2.0 * kr * kr * ((6.0 / (kx * kx) - 1.0) * b2 - b1 / kx);
    case 42: return -k * 4.0 * kr * kr * (-3.0 * b2 / (kx * kx) + b1 / kx);
  }
  return 0.0;
}

// Kernel of force wavenumber integral kint (1..21) at wavenumber k: 8..14 take
// the kernels of 1..7, 15..21 those of the x-derivatives.  Restates the case
// tables of forceWavenumberIntegrals, f90/wavenumberIntegrals.f90:162-176 (and
// :183-197, :220-234); kint 22, 23 (stresses) are not taken.
static double wn_kernel_force(int kint, double k, double x, double re, double b0, double b1, double b2) {
  const double kx = k * x, kr = k * re;
  switch (kint) {
    case 1: case 8: return k * b0;
    case 2: case 9: return -2.0 * b1 / re;
    case 3: case 10: case 15: return -k * k * b1 * re;
    case 4: case 11: case 16: return -2.0 * k * (b0 - b1 / kx);
    case 5: case 12: return -2.0 * k * b1 / kx;
    case 6: case 13: return 2.0 * k * (b0 - b1 / kx);
    case 7: case 14: return 2.0 * k * b1 / kx;
    case 17: return -k * kr * kr * (b0 - b1 / kx);
    case 18: return -2.0 * k * kr * (-b1 + b2 / kx);
    case 19: return 2.0 * k * kr * b2 / kx;
    case 20: return 2.0 * k * kr * (-b1 + b2 / kx);
    case 21: return -2.0 * k * kr * b2 / kx;
  }
  return 0.0;
}

// The 14 moment (istyp 1) or 7 force (istyp 0) wavenumber integrals at epicentral distance xx (m) for each frequency.
// gr, gi as in moment_sums (wavenumbers of frequency jf: j = 1..nwn[jf-1], k = (j-1) dwn).
// Restates momentWavenumberIntegrals (f90/wavenumberIntegrals.f90:321-477) with
// computeBesselFunctions (f90/besselFunctions.f90:33-51) in the frequency loop of
// besselDisplacementSpectra (f90/displacementSpectra.f90:119-184); for a force,
// forceWavenumberIntegrals (f90/wavenumberIntegrals.f90:122-250).  J0, J1, Y0,
// Y1 come from the C library (the reference vendors Cody's SPECFUN routines).
// wnint: [nf][nsum] complex.
static const int kSpecWn[14] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 8, 9, 6, 7};
static const int kSpecWnForce[7] = {0, 1, 2, 3, 4, 4, 3};
// strain = false: the displacement integrals kint 1..14 / 1..7 on U, V, W spectra;
// strain = true: kint 1..42 / 1..21 on spectra with radial derivatives (U, V,
// UP, VP, W, WP; positionDataGreenFKSpectra through the (component, jump) maps)
void bessel_sums(int istyp, int nf1, int nf2, const int* nwn, long nkf, const float* gr, const float* gi, double dwn,
                 double re, double xx, double tapfrac, cd* wnint, bool strain = false) {
  const int nsum = strain ? (istyp == 1 ? 42 : 21) : (istyp == 1 ? 14 : 7);
  const int njsph = istyp == 1 ? 4 : 2, njtor = istyp == 1 ? 2 : 1;
  auto specOf = [&](int kint) {
    if (!strain) return istyp == 1 ? kSpecWn[kint - 1] : kSpecWnForce[kint - 1];
    const int* cj = istyp == 1 ? kMomentMap[kint - 1] : kForceMap[kint - 1];
    return positionData(cj[0], cj[1], njsph, njtor);
  };
  auto kernel = [&](int kint, double k, double x, double b0, double b1, double b2) {
    return istyp == 1 ? wn_kernel(kint, k, x, re, b0, b1, b2) : wn_kernel_force(kint, k, x, re, b0, b1, b2);
  };
  int nmax = 0;
  for (int jf = nf1; jf <= nf2; ++jf) nmax = std::max(nmax, nwn[jf - 1]);
  // computeBesselFunctions: besselj(j, 0..2) at k = (j-1) dwn, j = 1 .. nwnmax
  std::vector<double> J0(nmax), J1(nmax), J2(nmax);
  J0[0] = 1.0; J1[0] = 0.0; J2[0] = 0.0;
  double wn = dwn;
  for (int j = 2; j <= nmax; ++j) {
    J0[j - 1] = ::j0(wn * xx);
    J1[j - 1] = ::j1(wn * xx);
    J2[j - 1] = -J0[j - 1] + 2.0 / (wn * xx) * J1[j - 1];
    wn = wn + dwn;
  }
  double x = xx;
  if (xx < 3.0 / (dwn * 1.0e9)) x = 3.0 / (dwn * 1.0e9);
  double wn3 = 3.0 / x;
  const int j3 = (int)(wn3 / dwn) + 2;
  wn3 = (j3 - 1) * dwn;
  const double xdk = x * dwn;
  std::vector<double> w(nmax), kern(nmax);
  long off = 0;
  for (int jf = nf1; jf <= nf2; ++jf) {
    const int n = nwn[jf - 1];
    taper_wn(n, dwn, tapfrac, w.data());
    const int j33 = std::min(j3, n);
    const double wne = (n - 1) * dwn;
    for (int kint = 1; kint <= nsum; ++kint) {
      const int s = specOf(kint);
      auto G = [&](int j) { return cd(gr[(long)s * nkf + off + j - 1], gi[(long)s * nkf + off + j - 1]); };
      for (int j = 1; j <= n; ++j) kern[j - 1] = kernel(kint, (j - 1) * dwn, x, J0[j - 1], J1[j - 1], J2[j - 1]);
      cd z = 0.0;
      for (int j = 2; j <= j33 - 1; ++j) z += G(j) * kern[j - 1] * w[j - 1];
      z += 0.5 * G(j33) * kern[j33 - 1] * w[j33 - 1];
      z *= dwn;
      if (j33 != n) {
        double y0 = ::y0(wn3 * x), y1 = ::y1(wn3 * x), y2 = -y0 + 2.0 / (wn3 * x) * y1;
        const double b3 = kernel(kint, wn3, x, y0, y1, y2);
        y0 = ::y0(wne * x); y1 = ::y1(wne * x); y2 = -y0 + 2.0 / (wne * x) * y1;
        const double be = kernel(kint, wne, x, y0, y1, y2);
        cd zs = G(j3) * kern[j3 - 1] * w[j3 - 1];
        for (int j = j3 + 1; j <= n - 1; ++j) zs += 2.0 * G(j) * kern[j - 1] * w[j - 1];
        zs += G(n) * kern[n - 1] * w[n - 1];
        z += zs / x * (1.0 - std::cos(xdk)) / xdk;
        z += (G(n) * be * w[n - 1] - G(j3) * b3 * w[j3 - 1]) / x * (1.0 - std::sin(xdk) / xdk);
      }
      z *= re * re / (2.0 * kPi);
      wnint[(long)(jf - nf1) * nsum + kint - 1] = z;
    }
    off += n;
  }
}

}  // namespace syn

extern "C" {

int orc_legendre(int lmax, double theta, double* plm, double* dpdt, double* pdst, double* d2pdt2, double* dtpdst) {
  if (lmax < 0) return 9;
  syn::legendre(lmax, theta, plm, dpdt, pdst, d2pdt2, dtpdst);
  return 0;
}

int orc_taper(int lmax, double tapfrac, double* w) {
  syn::taper(lmax, tapfrac, w);
  return 0;
}

int orc_fft(int n, int is, double* x, double* y) {
  std::vector<double> a(x, x + (1 << n)), b(y, y + (1 << n));
  syn::fft(a, b, n, is);
  std::memcpy(x, a.data(), sizeof(double) * a.size());
  std::memcpy(y, b.data(), sizeof(double) * b.size());
  return 0;
}

// sp: complex [nf2] (entries below nf1 ignored) -> ts [nsamp]; returns transformFrequencyTime's status
int orc_freq_time(int nf1, int nf2, double df, double sigma, const double* sp, int nsamp, double dt, float* ts) {
  return syn::freq_time(nf1, nf2, df, sigma, reinterpret_cast<const syn::cd*>(sp), nsamp, dt, ts);
}

// hdsum: [nrcv][nf][14] complex (interleaved)
int orc_moment_sums(int nf1, int nf2, const int* nwn, long nkf, const float* gr, const float* gi, int nrcv,
                    const double* theta, double tapfrac, double* hdsum, int nthreads) {
  const int nf = nf2 - nf1 + 1;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int r = 0; r < nrcv; ++r)
    syn::moment_sums(nf1, nf2, nwn, nkf, gr, gi, theta[r], tapfrac,
                     reinterpret_cast<syn::cd*>(hdsum) + (long)r * nf * 14);
  return 0;
}

// hdsum: [nrcv][nf][7] complex (interleaved)
int orc_force_sums(int nf1, int nf2, const int* nwn, long nkf, const float* gr, const float* gi, int nrcv,
                   const double* theta, double tapfrac, double* hdsum, int nthreads) {
  const int nf = nf2 - nf1 + 1;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int r = 0; r < nrcv; ++r)
    syn::force_sums(nf1, nf2, nwn, nkf, gr, gi, theta[r], tapfrac,
                    reinterpret_cast<syn::cd*>(hdsum) + (long)r * nf * 7);
  return 0;
}

// wnint: [nrcv][nf][nsum] complex, nsum = 14 (istyp 1, moment) or 7 (istyp 0, force);
// x: epicentral distances in m
int orc_bessel_sums_src(int istyp, int nf1, int nf2, const int* nwn, long nkf, const float* gr, const float* gi,
                        double dwn, double re, int nrcv, const double* x, double tapfrac, double* wnint,
                        int nthreads) {
  const int nf = nf2 - nf1 + 1, nsum = istyp == 1 ? 14 : 7;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int r = 0; r < nrcv; ++r)
    syn::bessel_sums(istyp, nf1, nf2, nwn, nkf, gr, gi, dwn, re, x[r], tapfrac,
                     reinterpret_cast<syn::cd*>(wnint) + (long)r * nf * nsum);
  return 0;
}

int orc_bessel_sums(int nf1, int nf2, const int* nwn, long nkf, const float* gr, const float* gi, double dwn,
                    double re, int nrcv, const double* x, double tapfrac, double* wnint, int nthreads) {
  return orc_bessel_sums_src(1, nf1, nf2, nwn, nkf, gr, gi, dwn, re, nrcv, x, tapfrac, wnint, nthreads);
}

// Strain (and stress) spectra of one node for nrcv receivers at epicentral
// distance theta[r] and azimuth phi[r], receiver radius re: istyp 1 (nsp 20:
// U, V, UP, VP x 4 jumps, W, WP x 2) or 0 (nsp 10).  Outputs (any may be NULL):
// hdsum [nrcv][nf][nsum] complex, zsp [nrcv][3][nf2], strain [nrcv][3][3][nf2]
// and, with elcon (7 complex moduli at the receiver), stress [nrcv][3][3][nf2]
// (the reference's zsp(nf2, 3) and strainsp(nf2, 3, 3) per receiver).
// dwn > 0: the regional branch (besselDisplacementSpectra, f90/displacementSpectra.f90:
// 125-184): wavenumber integrals kint 1..42 / 1..21 at x = theta rearth (theta
// the epicentral angle, as the reference passes delta)
int orc_strain_spectra_wn(int istyp, int nf1, int nf2, const int* nwn, long nkf, const float* gr, const float* gi,
                          int nrcv, const double* theta, const double* phi, double re, const double* fomt,
                          const double* elcon, double tapfrac, double dwn, double rearth, double* hdsum, double* zsp,
                          double* strain, double* stress, int nthreads) {
  const int nf = nf2 - nf1 + 1, nsum = istyp == 1 ? 42 : 21;
  const syn::cd* el = reinterpret_cast<const syn::cd*>(elcon);
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int r = 0; r < nrcv; ++r) {
    std::vector<syn::cd> h((size_t)nf * nsum);
    if (dwn > 0.0)
      syn::bessel_sums(istyp, nf1, nf2, nwn, nkf, gr, gi, dwn, rearth, theta[r] * rearth, tapfrac, h.data(), true);
    else
      syn::strain_sums(istyp, nf1, nf2, nwn, nkf, gr, gi, theta[r], tapfrac, h.data());
    if (hdsum) std::memcpy(hdsum + (size_t)r * nf * nsum * 2, h.data(), sizeof(syn::cd) * nf * nsum);
    for (int jf = 1; jf <= nf2; ++jf) {
      syn::cd u[4][3] = {}, e[3][3] = {}, t[3][3] = {};
      if (jf >= nf1) {
        syn::strain_parts(istyp, re, theta[r], phi[r], fomt, &h[(size_t)(jf - nf1) * nsum], u);
        syn::spherical_strain(u, re, theta[r], e);
        if (el) syn::spherical_stress(e, el, t);
      }
      for (int a = 0; a < 3; ++a) {
        if (zsp) {
          const size_t o = ((size_t)r * 3 + a) * nf2 + jf - 1;
          zsp[2 * o] = u[0][a].real(); zsp[2 * o + 1] = u[0][a].imag();
        }
        for (int b = 0; b < 3; ++b) {
          const size_t o = (((size_t)r * 3 + a) * 3 + b) * nf2 + jf - 1;
          if (strain) { strain[2 * o] = e[a][b].real(); strain[2 * o + 1] = e[a][b].imag(); }
          if (stress && el) { stress[2 * o] = t[a][b].real(); stress[2 * o + 1] = t[a][b].imag(); }
        }
      }
    }
  }
  return 0;
}

int orc_strain_spectra(int istyp, int nf1, int nf2, const int* nwn, long nkf, const float* gr, const float* gi,
                       int nrcv, const double* theta, const double* phi, double re, const double* fomt,
                       const double* elcon, double tapfrac, double* hdsum, double* zsp, double* strain,
                       double* stress, int nthreads) {
  return orc_strain_spectra_wn(istyp, nf1, nf2, nwn, nkf, gr, gi, nrcv, theta, phi, re, fomt, elcon, tapfrac, 0.0, 0.0,
                               hdsum, zsp, strain, stress, nthreads);
}

// pure strain combination for checks: u[4][3] complex -> e[3][3] complex
int orc_spherical_strain(const double* u, double r, double delta, double* e) {
  syn::cd uu[4][3], ee[3][3];
  for (int p = 0; p < 4; ++p) for (int c = 0; c < 3; ++c) uu[p][c] = syn::cd(u[2 * (3 * p + c)], u[2 * (3 * p + c) + 1]);
  syn::spherical_strain(uu, r, delta, ee);
  for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) { e[2 * (3 * a + b)] = ee[a][b].real(); e[2 * (3 * a + b) + 1] = ee[a][b].imag(); }
  return 0;
}

// hfil: complex [nf2]; filspecs = (high-pass order, corner, low-pass order, corner)
int orc_butterworth_bandpass(const double* filspecs, int nf1, int nf2, double df, double sigma, double* hfil) {
  syn::butterworth_bandpass((int)std::lround(filspecs[0]), filspecs[1], (int)std::lround(filspecs[2]), filspecs[3],
                            nf1, nf2, df, sigma, reinterpret_cast<syn::cd*>(hfil));
  return 0;
}

double orc_besselj(int order, double x) { return order == 0 ? ::j0(x) : ::j1(x); }
double orc_bessely(int order, double x) { return order == 0 ? ::y0(x) : ::y1(x); }

// Full synthetics of one source node: sums, displacement, time shift, velocity,
// ZNE rotation (propdir given) and the frequency-to-time transform.
// Restates the station loop of computeSyntheticSeismograms
// (f90/computeSyntheticSeismograms.f90:311-388) with unit event and station
// filters.  ts: [nrcv][3][nsamp]; zsp (optional): [nrcv][3][nf2] complex.
// istyp 1: moment source, fomt = (Mrr, Mtt, Mff, Mrt, Mrf, Mtf); istyp 0: force,
// fomt = (Fr, Ft, Ff) (evalSingleFrequencyDisplacementSpectra,
// f90/displacementSpectra.f90:200-228).
// dwn > 0: regional mode (besselDisplacementSpectra, f90/displacementSpectra.f90:119-184):
// theta[r] is the epicentral distance in radians as in the global mode and the
// wavenumber integrals are taken at x = theta[r] * re.
// evfilter: complex [nf2] or NULL; stafilter: complex [nrcv][3][nf2] or NULL
// (the event and station filters of :367-371; NULL: unity).
int orc_synthetics_src(int istyp, int nf1, int nf2, double df, double sigma, const int* nwn, long nkf,
                       const float* gr, const float* gi, double dwn, double re, int nrcv, const double* theta,
                       const double* phi, const double* propdir, const double* tred, const double* evfilter,
                       const double* stafilter, const double* fomt, int velocity, int nsamp, double dt, float* ts,
                       double* zsp, int nthreads) {
  const syn::cd* evf = reinterpret_cast<const syn::cd*>(evfilter);
  const syn::cd* stf = reinterpret_cast<const syn::cd*>(stafilter);
  const int nf = nf2 - nf1 + 1, nsum = istyp == 1 ? 14 : 7;
  int bad = 0;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1) reduction(| : bad)
  for (int r = 0; r < nrcv; ++r) {
    std::vector<syn::cd> h((size_t)nf * nsum), z((size_t)3 * nf2, 0.0);
    if (dwn > 0.0) syn::bessel_sums(istyp, nf1, nf2, nwn, nkf, gr, gi, dwn, re, theta[r] * re, 0.2, h.data());
    else if (istyp == 1) syn::moment_sums(nf1, nf2, nwn, nkf, gr, gi, theta[r], 0.2, h.data());
    else syn::force_sums(nf1, nf2, nwn, nkf, gr, gi, theta[r], 0.2, h.data());
    const double tr = tred ? tred[r] : 0.0;
    for (int jf = nf1; jf <= nf2; ++jf) {
      syn::cd u[3];
      if (istyp == 1) syn::moment_displacement(phi[r], fomt, &h[(size_t)(jf - nf1) * nsum], u);
      else syn::force_displacement(phi[r], fomt, &h[(size_t)(jf - nf1) * nsum], u);
      // time shift filter, f90/computeSyntheticSeismograms.f90:361-365
      const double om = 2.0 * syn::kPi * (jf - 1) * df;
      const syn::cd tsf = std::exp(syn::cd(0.0, om * tr)) * std::exp(sigma * tr);
      // event and station filters with the time shift, :367-371
      for (int c = 0; c < 3; ++c) {
        syn::cd fil = tsf;
        if (evf) fil = evf[jf - 1] * fil;
        if (stf) fil = stf[((size_t)r * 3 + c) * nf2 + jf - 1] * fil;
        u[c] *= fil;
      }
      if (velocity) {                                   // :376-378
        const syn::cd zom(om, -sigma);
        for (int c = 0; c < 3; ++c) u[c] = u[c] * syn::cd(0.0, 1.0) * zom;
      }
      if (propdir) {                                    // f90/axesRotation.f90:560-568
        const double ca = std::cos(propdir[r]), sa = std::sin(propdir[r]);
        const syn::cd uz = u[0], ue = sa * u[1] + ca * u[2], un = -ca * u[1] + sa * u[2];
        u[0] = uz; u[1] = un; u[2] = ue;
      }
      for (int c = 0; c < 3; ++c) z[(size_t)c * nf2 + jf - 1] = u[c];
    }
    if (zsp) std::memcpy(zsp + (size_t)r * 3 * nf2 * 2, z.data(), sizeof(syn::cd) * 3 * nf2);
    for (int c = 0; c < 3; ++c)
      bad |= syn::freq_time(nf1, nf2, df, sigma, &z[(size_t)c * nf2], nsamp, dt, ts + ((size_t)r * 3 + c) * nsamp);
  }
  return bad;
}

int orc_synthetics_mode(int nf1, int nf2, double df, double sigma, const int* nwn, long nkf, const float* gr,
                        const float* gi, double dwn, double re, int nrcv, const double* theta, const double* phi,
                        const double* propdir, const double* tred, const double* mt, int velocity, int nsamp,
                        double dt, float* ts, double* zsp, int nthreads) {
  return orc_synthetics_src(1, nf1, nf2, df, sigma, nwn, nkf, gr, gi, dwn, re, nrcv, theta, phi, propdir, tred,
                            nullptr, nullptr, mt, velocity, nsamp, dt, ts, zsp, nthreads);
}

int orc_synthetics(int nf1, int nf2, double df, double sigma, const int* nwn, long nkf, const float* gr,
                   const float* gi, int nrcv, const double* theta, const double* phi, const double* propdir,
                   const double* tred, const double* mt, int velocity, int nsamp, double dt, float* ts, double* zsp,
                   int nthreads) {
  return orc_synthetics_mode(nf1, nf2, df, sigma, nwn, nkf, gr, gi, 0.0, 0.0, nrcv, theta, phi, propdir, tred, mt,
                             velocity, nsamp, dt, ts, zsp, nthreads);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Travel time of the first arrival from a ray table (f90/travelTimes.f90).
// The table is the reference's raytable_travel_times with its FP32 arrays in
// Fortran order: rnod[nnod], rtp[ntp], slowness[ntp] (angular slowness rtp/v),
// delta and ttime (nnod, ntp), element (jn, k) at (jn-1) + (k-1) nnod.  All
// indices below are 1-based as in the reference.
// ---------------------------------------------------------------------------
namespace tt {

// locate (f90/locatePoint.f90:43-75): bisection returning j with x(j) < xx <= x(j+1)
// for an ascending x (0 below, n above); for a descending x the comparisons turn.
template <class V, class F>
static int locate(V xx, int n, F x) {
  int lo = 0, hi = n + 1;
  const bool asc = x(n) > x(1);
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (asc == (xx > x(mid))) lo = mid;
    else hi = mid;
  }
  return lo;
}

struct Table {
  int nnod, ntp, istele;
  const float *rnod, *rtp, *slow, *delta, *ttime;
  double d(int jn, int k) const { return (double)delta[(jn - 1) + (size_t)(k - 1) * nnod]; }
  double t(int jn, int k) const { return (double)ttime[(jn - 1) + (size_t)(k - 1) * nnod]; }
};

// node bracket and weights of a radius (setSourceInfoTravelTimes :128-176 and the
// receiver half of getReceiverTravelTimes :208-236): 0 or an error code
static int bracket(const Table& T, double r, int& j, double& a, double& b) {
  const double tol = 1e-5;
  j = locate((float)r, T.nnod, [&](int i) { return T.rnod[i - 1]; });
  if (j == 0) {
    if (std::fabs(T.rnod[0] - r) > tol * r) return 1;   // below ray table range
    j = 1;
  } else if (j == T.nnod) {
    if (std::fabs(T.rnod[T.nnod - 1] - r) > tol * r) return 2;   // above ray table range
    j = T.nnod - 1;
  }
  const double h = (double)(T.rnod[j] - T.rnod[j - 1]);   // FP32 difference, as REAL arrays give
  a = ((double)T.rnod[j] - r) / h;
  b = 1.0 - a;
  return 0;
}

// getTravelTimes (:373-388) = setSourceInfoTravelTimes + getReceiverTravelTimes
// (:180-367).  Status: 0 ok; 1/2 source radius below/above the table; 3 no
// turning point below the source; 4 none below the receiver; 5/6 receiver
// radius below/above the table; 7 fewer than two usable turning points (the
// reference would read past its arrays); 8 distance out of the turning-point
// range with istele set.
static int travel_time(const Table& T, double rs, double re, double din, double& tt, double& pout) {
  tt = -1.0;
  pout = 0.0;
  int js, je;
  double as, bs, ae, be;
  int st = bracket(T, rs, js, as, bs);
  if (st) return st;
  auto rtp = [&](int i) { return T.rtp[i - 1]; };
  const int jtps = locate((float)rs, T.ntp, rtp);
  if (jtps == 0) return 3;
  const int jtpr = locate((float)re, T.ntp, rtp);
  if (jtpr == 0) return 4;
  st = bracket(T, re, je, ae, be);
  if (st) return st + 4;
  const int jtpmax = std::min(jtps, jtpr);
  if (jtpmax < 2) return 7;
  auto p = [&](int k) { return (double)T.slow[k - 1]; };
  // distance and intercept time for turning point k, down-going from the
  // source and up to the receiver (sgn = +1) or the steep ray without a turn (sgn = -1)
  auto del = [&](int k, double sgn) {
    return as * T.d(js, k) + bs * T.d(js + 1, k) + sgn * (ae * T.d(je, k)) + sgn * (be * T.d(je + 1, k));
  };
  auto tau = [&](int k, double sgn) {
    return as * T.t(js, k) + bs * T.t(js + 1, k) + sgn * (ae * T.t(je, k)) + sgn * (be * T.t(je + 1, k)) -
           p(k) * del(k, sgn);
  };
  double dmax = del(1, 1.0), dmin = dmax;
  for (int k = 2; k <= jtpmax; ++k) {
    const double v = del(k, 1.0);
    dmax = std::max(dmax, v);
    dmin = std::min(dmin, v);
  }
  int jd;
  double ad, bd;
  double sgn = 1.0;
  if (din >= dmin) {
    if (din > dmax) {   // beyond the last turning point: diffracted along it
      tt = tau(1, 1.0) + p(1) * din;
      pout = p(1);
      return 0;
    }
    if (T.istele) {
      jd = locate(din, jtpmax, [&](int k) { return del(k, 1.0); });
    } else {   // findZeroPassTravelTimes(..., 'end') (:395-413): the last sign change
      jd = jtpmax;
      double prod = 1.0;
      while (prod > 0.0 && jd > 1) {
        prod = (din - del(jd - 1, 1.0)) * (din - del(jd, 1.0));
        --jd;
      }
    }
    jd = std::max(1, std::min(jd, jtpmax - 1));
    const double h = del(jd + 1, 1.0) - del(jd, 1.0);
    ad = (del(jd + 1, 1.0) - din) / h;
    bd = 1.0 - ad;
  } else {
    if (T.istele) return 8;
    sgn = -1.0;
    jd = locate(din, jtpmax, [&](int k) { return del(k, -1.0); });
    if (jd == 0 || jd == jtpmax) {   // extrapolate with the end slope (:319-342)
      const bool left = jd == 0;
      jd = left ? 1 : jtpmax - 1;
      const double h = del(jd + 1, -1.0) - del(jd, -1.0);
      if (std::fabs(h) < 1e-4) {
        ad = left ? 1.0 : 0.0;
        bd = left ? 0.0 : 1.0;
      } else {
        ad = (del(jd + 1, -1.0) - din) / h;
        bd = 1.0 - ad;
      }
    } else {
      const double h = del(jd + 1, -1.0) - del(jd, -1.0);
      ad = (del(jd + 1, -1.0) - din) / h;
      bd = 1.0 - ad;
    }
  }
  const double ta = ad * tau(jd, sgn) + bd * tau(jd + 1, sgn);
  const double pa = ad * p(jd) + bd * p(jd + 1);
  tt = ta + pa * din;
  pout = pa;
  return 0;
}

}  // namespace tt

extern "C" int orc_travel_times(int nnod, int ntp, int istele, const float* rnod, const float* rtp, const float* slowness,
                                const float* delta, const float* ttime, double rs, int nrcv, const double* re,
                                const double* din, double* tt, double* slow, int* status) {
  const tt::Table T{nnod, ntp, istele, rnod, rtp, slowness, delta, ttime};
  for (int r = 0; r < nrcv; ++r) status[r] = tt::travel_time(T, rs, re[r], din[r], tt[r], slow[r]);
  return 0;
}
