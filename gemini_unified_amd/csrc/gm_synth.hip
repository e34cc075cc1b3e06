// Synthetics half of the Green-FK path on MI355X: Legendre basis, harmonic
// degree sums of a moment-tensor (14 sums, 10 spectra per node) or single-force
// (7 sums, 5 spectra per node) source as a split-bf16 MFMA contraction, the
// per-frequency displacement combination, and the frequency-to-time transform.
//
// Pipeline of gm_synthetics (one source node, R receivers, nf frequencies):
//   k_twiddle   e^{+2 pi i k / ns}, k < ns, FP64 (pi as in the reference)
//   k_legendre  per (receiver, m): P_l^m, dP/dtheta, P/sin (m = 0..2) by the FP64
//               upward recurrence, times (2l+1)/(4 pi) and 2/(l(l+1)), as
//               8 basis rows; each value is split into three bf16 pieces
//               (hi, mid, lo: 24 significant bits).  Layout
//               [piece][row][Lp/32][Rp][32]: runs of 32 degrees (one
//               contraction chunk), receivers next to each other, so that a
//               128 x 32 tile of the contraction is one contiguous 8 KB block.
//   k_gprep     spectra G_s(l, f) * taper_f(l) in FP64, split into three bf16
//               pieces.  Layout [piece][s][Np][Lp], column n = 2 f' (+1 imag).
//   k_contract  the sums S_j[r][n] = sum_l B_b(j)[r][l] G_s(j)[n][l] on the
//               matrix cores: v_mfma_f32_16x16x32_bf16 over the six piece
//               products whose exponent sum keeps 24 bits (hi*hi, hi*mid,
//               mid*hi, hi*lo, mid*mid, lo*hi).  Products of bf16 pieces are
//               exact in FP32; each 32-degree chunk is accumulated in FP32 by
//               the MFMA and then added to an FP64 running sum (compensation).
//   k_combine   displacement (r, theta, phi) from the sums and the moment
//               tensor or force, time shift, optional velocity, optional ZNE
//               rotation.
//   k_ifft      per trace: Hermitian spectrum -> real time series through an
//               ns/2-point complex FFT in LDS (real-output split), times df
//               e^{sigma t}, stored as FP32.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
#include <algorithm>
#include "gm_common.h"
#include "gm_ctx.h"
#include "../../include/gemini_fk.h"

#pragma clang fp contract(off)

namespace gm_syn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// pi of the reference's synthetics code (mc_pid, mathConstants.f90:34): its
// literal is single precision, so the value is FP32 pi; used for 4 pi, the
// tapers, 2 pi f and the transform's twiddle angles as the reference does
constexpr double kPi = 3.1415927410125732;
constexpr int kRows = 8;       // basis rows
constexpr int kSums = 14;      // moment-tensor degree sums (the larger set)
constexpr int kForceSums = 7;  // single-force degree sums
constexpr int kPieces = 3;     // bf16 pieces per value
constexpr int TM = 128, TN = 64, TK = 32, LDSK = 48;   // tile sizes; LDS row pitch (bf16)
constexpr int kXcd = 8;        // XCDs of the MI355X the contraction's tile order is laid out for
constexpr int kStrip = 8;      // row tiles per strip of that order
constexpr int kMaxFFT = 8192;  // longest time series (ns/2 complex FP64 = 64 KB of LDS, 68 KB padded)

// Sum j of source type t (0 single force, kint 1..7; 1 moment tensor, kint
// 1..14): basis row, spectrum index (offset of (component, jump) in the GFK
// file for dsvmask (U, V, W): force U1 U2 V1 V2 W1, moment U1..U4 V1..V4 W1 W2)
// and the factor relating the sum's weight to the basis row
// (forceHarmonicDegreeSums, harmonicDegreeSums.f90:115-164, with the map at
// :34-36; momentHarmonicDegreeSums, :251-307, with the map at :42-46)
__constant__ int c_row[2][kSums] = {{0, 1, 3, 4, 6, 4, 6}, {0, 0, 1, 2, 3, 3, 4, 5, 6, 7, 4, 5, 6, 7}};
__constant__ int c_spec[2][kSums] = {{0, 1, 2, 3, 4, 4, 3}, {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 8, 9, 6, 7}};
__constant__ double c_fac[2][kSums] = {{1, -1, 1, -1, -1, 1, 1}, {1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 1, -1, 1, -2}};
// regional mode: the wavenumber-integral kernels map onto the 8 distinct rows
// of wnRows with the same row and spectrum tables and these signs
// (forceWavenumberIntegrals, wavenumberIntegrals.f90:162-176;
// momentWavenumberIntegrals, :353-373)
__constant__ double c_fac_wn[2][kSums] = {{1, -1, 1, -1, -1, 1, 1}, {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -1, 1, -1}};
// first degree each spectrum enters a sum with (the sums' loop bounds); the V
// spectra at l = 0 hold 0/0 from the reciprocity factor 1/dll1 and are never read
__constant__ int c_lmin[2][10] = {{0, 1, 1, 1, 1}, {0, 0, 1, 2, 1, 1, 1, 2, 1, 2}};

__device__ __forceinline__ void split3(double v, __bf16& a, __bf16& b, __bf16& c) {
  a = (__bf16)(float)v;
  double r = v - (double)(float)a;
  b = (__bf16)(float)r;
  r = r - (double)(float)b;
  c = (__bf16)(float)r;
}

__global__ void k_twiddle(int ns, double* __restrict__ tw) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  double s, c;
  sincos((2.0 * kPi / ns) * k, &s, &c);
  tw[2 * k] = c;
  tw[2 * k + 1] = s;
}

// One (receiver, order m) per thread, m = blockIdx.y: the three recurrences in
// l (m = 0, 1, 2) are independent, so they run on three lanes.  mode 0: basis
// pieces Bs[piece][row][Lp/32][Rp][32] of the rows of order m (rows 0, 3 for m = 0;
// 1, 4, 6 for m = 1; 2, 5, 7 for m = 2) for degrees 0..Lp-1 (zero beyond
// lmax); mode 1: FP64 arrays [R][3][lmax+1] of plm, dpdt, pdst, d2pdt2, dtpdst
// (computeLegendreFunctions, legendreFunctions.f90:48-157, same operation order).
// P_l^m (m = 0..2, no Condon-Shortley sign) from P_{l-1}^m = p1 and P_{l-2}^m = p2
// by the upward recurrence (legendreFunctions.f90:87-100); pmm = P_m^m.  The
// reference divides by l - m; here the numerator is multiplied by 1/(l - m),
// which does not depend on p1, p2 and leaves the chain from one degree to the
// next two multiplications and a subtraction long (within an ulp of the quotient).
__device__ __forceinline__ double legendreNext(int l, int m, int lmax, double ct, double pmm, double p1, double p2) {
  if (l > lmax || l < m) return 0.0;
  if (l == m) return pmm;
  if (l == m + 1) return ct * (double)(2 * m + 1) * pmm;
  const double rl = 1.0 / (double)(l - m);
  return (ct * (double)(2 * l - 1) * p1 - (double)(l + m - 1) * p2) * rl;
}
__device__ __forceinline__ double legendrePmm(int m, double st) {
  return m == 0 ? 1.0 : (m == 1 ? (1.0 * 1.0) * st : ((1.0 * 1.0) * st * 3.0) * st);
}

// Start values of the degree segments of k_legendre.  Segment s covers degrees
// [s seglen, (s+1) seglen).  Beyond degree m + 1 the recurrence is linear in the
// pair v_l = (P_l^m, P_{l-1}^m): v_l = M_l v_{l-1}, M_l = [[c_l, -d_l], [1, 0]]
// (M_l = [[0, 0], [1, 0]] beyond lmax).  So, per (receiver r, order m), one
// thread runs segment 0 by the recurrence itself and stores v at its last
// degree, and one thread per segment s = 1..nseg-2 forms the product T_s of its
// M_l.  k_legendre's segment s then starts from T_{s-1} ... T_1 v.  Segments 0
// and 1 continue the recurrence bit for bit, later ones to rounding.
// seeds: [Rp][3][nseg][4] (slot 0: v, slot s: T_s).  Grid (Rp / 64, 3, nseg - 1).
__global__ void __launch_bounds__(64) k_legendre_seed(int R, int Rp, int lmax, const double* __restrict__ theta,
                                                      int nseg, int seglen, double* __restrict__ seeds) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y, sg = blockIdx.z;
  if (r >= Rp) return;
  const double th = r >= R ? 0.5 * kPi : theta[r];
  const double ct = cos(th), st = sin(th);
  double* out = seeds + (((size_t)r * 3 + m) * nseg + sg) * 4;
  if (sg == 0) {
    const double pmm = legendrePmm(m, st);
    double a1 = 0.0, a2 = 0.0;
    for (int l = 0; l < seglen; ++l) {
      const double p = legendreNext(l, m, lmax, ct, pmm, a1, a2);
      a2 = a1;
      a1 = p;
    }
    out[0] = a1;
    out[1] = a2;
    return;
  }
  // T = M_{l1} ... M_{l0}, l0 = sg seglen >= m + 2
  double t00 = 1.0, t01 = 0.0, t10 = 0.0, t11 = 1.0;
  for (int l = sg * seglen; l < (sg + 1) * seglen; ++l) {
    double c = 0.0, d = 0.0;
    if (l <= lmax) {
      const double rl = 1.0 / (double)(l - m);
      c = ct * (double)(2 * l - 1) * rl;
      d = (double)(l + m - 1) * rl;
    }
    const double n00 = c * t00 - d * t10, n01 = c * t01 - d * t11;
    t10 = t00;
    t11 = t01;
    t00 = n00;
    t01 = n01;
  }
  out[0] = t00; out[1] = t01; out[2] = t10; out[3] = t11;
}

// mode 0: basis rows for the contraction; the degrees are cut into segments of
// seglen (blockIdx.z), each continuing the recurrence from its seeds (above).
// mode 2: the same rows of the theta derivatives (dP_lm, d2P_lm/dtheta2,
// d(P_lm/sin)/dtheta in place of P_lm, dP_lm, P_lm/sin), the basis of the
// theta-derivative sums of the strain (harmonicDegreeSums.f90:316-351).
// mode 1: FP64 outputs of computeLegendreFunctions, one segment.
__global__ void __launch_bounds__(64) k_legendre(int R, int Rp, int lmax, int Lp, const double* __restrict__ theta,
                                                 __bf16* __restrict__ Bs, double* __restrict__ o_plm,
                                                 double* __restrict__ o_dpdt, double* __restrict__ o_pdst,
                                                 double* __restrict__ o_d2, double* __restrict__ o_ds, int mode,
                                                 int nseg, int seglen, const double* __restrict__ seeds) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (r >= (mode == 1 ? R : Rp)) return;
  const size_t plane = (size_t)Rp * Lp;
  const int L1 = lmax + 1;
  const bool pad = r >= R;
  const double th = pad ? 0.5 * kPi : theta[r];
  const double ct = cos(th), st = sin(th);
  const bool poles = !(st > 0.0);
  const double rst = poles ? 0.0 : 1.0 / st;
  const double sg = ct < 0.0 ? -1.0 : 1.0;
  const double fourpi = 4.0 * kPi;
  const double pmm = legendrePmm(m, st);
  // basis rows of this order and how many
  const int nrow = m == 0 ? 2 : 3;
  const int row0 = m, row1 = m + 3, row2 = m + 5;
  const int seg = blockIdx.z;
  double p1 = 0.0, p2 = 0.0;
  if (seg > 0) {
    const double* sd = seeds + ((size_t)r * 3 + m) * nseg * 4;
    p1 = sd[0];
    p2 = sd[1];
    for (int t = 1; t < seg; ++t) {
      const double* T = sd + 4 * t;
      const double q1 = T[0] * p1 + T[1] * p2, q2 = T[2] * p1 + T[3] * p2;
      p1 = q1;
      p2 = q2;
    }
  }
  // one degree of the recurrence (values of this order m at degree l)
  auto step = [&](int l, double& p, double& d, double& sv, double& d2, double& ds) {
    const double ll1 = (double)(l * (l + 1));
    const double sl = (l & 1) ? sg : 1.0, sl1 = (l & 1) ? 1.0 : sg;
    const double q = (double)(l + 2) * (l + 1) * l * (l - 1);
    p = legendreNext(l, m, lmax, ct, pmm, p1, p2);
    d = 0.0; sv = 0.0; d2 = 0.0; ds = 0.0;
    if (l <= lmax) {
      if (!poles) {
        if (l == m) d = (double)m * ct * p * rst;
        else if (l > m) d = ((double)l * ct * p - (double)(l + m) * p1) * rst;
        if (m >= 1 && l >= m) sv = p * rst;
        if (l >= m) d2 = -ct * rst * d + ((double)(m * m) * (rst * rst) - ll1) * p;
        if (m >= 1 && l >= m) ds = rst * (d - ct * rst * p);
      } else {
        if (m == 1) d = 0.5 * ll1 * sl;
        if (m == 1 && l >= 1) sv = 0.5 * ll1 * sl1;
        if (m == 0) d2 = 0.5 * ll1 * sl;
        if (m == 2) d2 = 0.25 * q * sl;
        if (m == 2 && l >= 1) ds = 0.125 * q * sl1;
      }
    }
    p2 = p1;
    p1 = p;
  };
  if (mode == 1) {
    for (int l = 0; l < L1; ++l) {
      double p, d, sv, d2, ds;
      step(l, p, d, sv, d2, ds);
      const size_t o = ((size_t)r * 3 + m) * L1 + l;
      o_plm[o] = p; o_dpdt[o] = d; o_pdst[o] = sv; o_d2[o] = d2; o_ds[o] = ds;
    }
    return;
  }
  // mode 0: runs of 32 degrees (Lp and seglen are multiples of 32), each row's
  // values split into bf16 pieces in registers, 8 to a 16-byte vector, and the
  // run's four vectors stored together: one receiver's 64-byte run of the layout
  const int lend = min(Lp, (seg + 1) * seglen);
  for (int l32 = seg * seglen; l32 < lend; l32 += 32) {
    bf16x8 v[4][kPieces][3];
#pragma unroll
    for (int e32 = 0; e32 < 32; ++e32) {
      const int gq = e32 >> 3, e = e32 & 7;
      const int l = l32 + e32;
      double p, d, sv, d2, ds;
      step(l, p, d, sv, d2, ds);
      const double ll1 = (double)(l * (l + 1));
      double b[3];
      const double w = (2.0 * l + 1.0) / fourpi;
      const double g = l >= 1 ? 2.0 / ll1 : 0.0;
      const bool dv = mode == 2;
      const double q0 = dv ? d : p, q1 = dv ? d2 : d, q2 = dv ? ds : sv;
      if (m == 0) {
        b[0] = q0 * w;
        b[1] = l >= 1 ? q1 * w : 0.0;
        b[2] = 0.0;
      } else {
        const bool on = l >= m;
        b[0] = on ? q0 * g * w : 0.0;
        b[1] = on ? q1 * g * w : 0.0;
        b[2] = on ? q2 * g * w : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double x = (pad || l > lmax) ? 0.0 : b[k];
        __bf16 h0, h1, h2;
        split3(x, h0, h1, h2);
        v[gq][0][k][e] = h0; v[gq][1][k][e] = h1; v[gq][2][k][e] = h2;
      }
    }
#pragma unroll
    for (int pc = 0; pc < kPieces; ++pc)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (k >= nrow) continue;
        const int row = k == 0 ? row0 : (k == 1 ? row1 : row2);
        bf16x8* dst = (bf16x8*)(Bs + ((size_t)pc * kRows + row) * plane + ((size_t)(l32 >> 5) * Rp + r) * 32);
#pragma unroll
        for (int gq = 0; gq < 4; ++gq) dst[gq] = v[gq][pc][k];
      }
  }
}

// Tapered spectra split into bf16 pieces: Gs[piece][s][Np][Lp].
// taperHarmonicDegreeSums (harmonicDegreeSums.f90:356-369) per frequency.
// Regional mode (dwn > 0): taperWavenumberIntegrals (wavenumberIntegrals.f90:
// 499-512) per frequency; the k = 0 column (index 0) is never summed.
__device__ __forceinline__ double taperWn(int l, int n, double dwn, double tapfrac) {
  long jt = lround(n * (1.0 - tapfrac));
  if (jt > n) jt = n;
  if (jt < 1) jt = 1;
  if (l + 1 < jt) return 1.0;
  const double c1 = (n - 1) * dwn * tapfrac * 2.0 / kPi;
  const double c2 = (1.0 - tapfrac) / tapfrac * 0.5 * kPi;
  const double c = cos(l * dwn / c1 - c2);
  return c * c;
}

// One thread per (frequency column pair f, degree l): the taper is computed
// once and applied to the real and imaginary parts of every spectrum; columns
// 2f and 2f + 1 of every spectrum plane are written, zero beyond the
// frequency's degrees and in the padding columns (f >= nf).  Grid (Lp / 256
// rounded up, Np / 2).
// Plane s of Gs takes spectrum smap.s[s] of the node's spectra (the identity for
// files with U, V, W; the displacement or the radial-derivative spectra of files
// with derivatives; -1 for the absent W of a receiver in the ocean).
struct SpecMap { int s[10]; };
__global__ void k_gprep(int nsp, int nf, int Np, int Lp, const int* __restrict__ nwn, const long long* __restrict__ koff,
                        long long nkf, const float* __restrict__ gr, const float* __restrict__ gi, double tapfrac,
                        double dwn, int istyp, SpecMap smap, __bf16* __restrict__ Gs) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (l >= Lp) return;
  const size_t per_s = (size_t)Np * Lp, plane = per_s * nsp;
  const bool regional = dwn > 0.0;
  const int n = f < nf ? nwn[f] : 0;
  double w = 0.0;
  long long k = 0;
  if (l < n) {
    k = koff[f] + l;
    if (regional) {
      w = taperWn(l, n, dwn, tapfrac);
    } else {
      const int lmax = n - 1;
      long jt = lround(lmax * (1.0 - tapfrac));
      if (jt > lmax) jt = lmax;
      if (jt < 1) jt = 1;
      w = 1.0;
      if (l >= jt) {
        const double c1 = lmax * tapfrac * 2.0 / kPi;
        const double c2 = (1.0 - tapfrac) / tapfrac * 0.5 * kPi;
        const double c = cos((double)l / c1 - c2);
        w = c * c;
      }
    }
  }
  for (int s = 0; s < nsp; ++s) {
    // an absent spectrum (smap.s[s] < 0: W of a receiver in the ocean) is zero
    const bool on = smap.s[s] >= 0 && l < n && (regional ? l >= 1 : l >= c_lmin[istyp][s]);
    const long long so = on ? (long long)smap.s[s] * nkf + k : 0;
    const double vr = on ? (double)gr[so] * w : 0.0;
    const double vi = on ? (double)gi[so] * w : 0.0;
    const size_t o = (size_t)s * per_s + (size_t)(2 * f) * Lp + l;
    __bf16 a, b, c;
    split3(vr, a, b, c);
    Gs[o] = a;
    Gs[o + plane] = b;
    Gs[o + 2 * plane] = c;
    split3(vi, a, b, c);
    Gs[o + Lp] = a;
    Gs[o + Lp + plane] = b;
    Gs[o + Lp + 2 * plane] = c;
  }
}

// H[j][Rp][Np] (FP64) = c_fac[t][j] * sum over degrees of basis row c_row[t][j]
// times spectrum c_spec[t][j] (the j-th sum of source type t = istyp); one
// block per (degree split, sum, tile) of Np/TN x Rp/TM tiles, 256 threads: 4
// waves, wave w the rows [32 w, 32 w + 32) of the tile and all TN = 64 columns
// (2 x 4 MFMA tiles of 16 x 16).  The basis fragments go from global memory
// straight to registers, two chunks of TK degrees ahead: in the basis layout
// [piece][row][Lp/32][Rp][32] the 16 rows x 32 degrees of a fragment are one
// contiguous 1 KB block, so a wave loads it fully coalesced and no other wave
// needs it.  Only the spectrum tile (64 columns x 32 degrees x 3 pieces) is
// staged in LDS, one chunk ahead, shared by the four waves.  Each output tile takes the six
// piece products in the same order, and folds its FP32 partial into the FP64
// sum after every chunk.  With nsplit > 1 split s covers chunks [s kper, (s+1) kper)
// of TK degrees and writes its partial sums to plane s of H ([nsplit][nsum][Rp][Np]),
// which k_sumsplits adds in split order.
__global__ void __launch_bounds__(256, 2) k_contract(int Rp, int Np, int Lp, int nsp, const __bf16* __restrict__ Bs,
                                                  const __bf16* __restrict__ Gs, double* __restrict__ H, int regional,
                                                  int istyp, int nsplit, int kper) {
  // the spectrum tile, double-buffered by chunk parity: one barrier per chunk
  __shared__ __attribute__((aligned(16))) __bf16 sB2[2][kPieces][TN][LDSK];
  // XCD-aware tile order.  The hardware deals blocks to the kXcd XCDs
  // round-robin by block index, so block b runs on XCD b % kXcd as the
  // (b / kXcd)-th of its XCD.  Each XCD is given one contiguous range of the
  // (sum, tile) list, ordered in strips of kStrip row tiles with the row
  // tile fastest: blocks that run together on one XCD then share column
  // tiles of the spectra and row tiles of the basis in that XCD's L2.
  const int nsum = istyp ? kSums : kForceSums;
  const int gx = Np / TN, gy = Rp / TM, ntiles = gx * gy, nwork = nsum * ntiles * nsplit;
  const int b = blockIdx.x, xcd = b % kXcd, sq = b / kXcd;
  const int per = (nwork + kXcd - 1) / kXcd, rem = nwork - (per - 1) * kXcd;   // rem XCDs get per items
  const int wk = xcd < rem ? xcd * per + sq : rem * per + (xcd - rem) * (per - 1) + sq;
  const int split = wk / (nsum * ntiles), wks = wk % (nsum * ntiles);
  const int j = wks / ntiles, tile = wks % ntiles;
  const int kbeg = split * kper * TK, kend = min(Lp, (split + 1) * kper * TK);
  const int strip = tile / (kStrip * gx), sidx = tile % (kStrip * gx);
  const int sh = min(kStrip, gy - strip * kStrip);
  const int ty = strip * kStrip + sidx % sh, tx = sidx / sh;
  const int rowb = c_row[istyp][j], spec = c_spec[istyp][j];
  const double fac = regional ? c_fac_wn[istyp][j] : c_fac[istyp][j];
  const int r0 = ty * TM, n0 = tx * TN;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const size_t planeA = (size_t)Rp * Lp, planeG = (size_t)Np * Lp;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  // this lane's basis fragment of chunk k0, row fragment a, piece pc:
  // abase + pc apiece + (k0 / 32) Rp 32 + 16 a 32
  const __bf16* __restrict__ abase = Bs + (size_t)rowb * planeA + (size_t)(r0 + 32 * w + fr) * 32 + fk;
  const size_t apiece = (size_t)kRows * planeA;
  // global -> register staging of the spectrum tile (3 pieces x 64 columns x 4
  // groups of 8 degrees) and the basis fragments, written out with literal
  // indices: as loops the registers were indexed at run time and kept in scratch
  uint4 rb0, rb1, rb2;
#define GM_B_SRC(i, k0)                                                                                        \
  (Gs + ((size_t)((t + 256 * (i)) >> 8) * nsp + spec) * planeG +                                             \
   (size_t)(n0 + (((t + 256 * (i)) & 255) >> 2)) * Lp + (k0) + 8 * ((t + 256 * (i)) & 3))
#define GM_B_DST(i) (&sB[(t + 256 * (i)) >> 8][((t + 256 * (i)) & 255) >> 2][8 * ((t + 256 * (i)) & 3)])
#define GM_SB(k0) (sB2[(((k0) - kbeg) / TK) & 1])
#define GM_BLOAD(k0)                                                                                           \
  do {                                                                                                         \
    rb0 = *(const uint4*)GM_B_SRC(0, k0); rb1 = *(const uint4*)GM_B_SRC(1, k0);                                \
    rb2 = *(const uint4*)GM_B_SRC(2, k0);                                                                      \
  } while (0)
#define GM_LSTORE()                                                                                            \
  do { *(uint4*)GM_B_DST(0) = rb0; *(uint4*)GM_B_DST(1) = rb1; *(uint4*)GM_B_DST(2) = rb2; } while (0)
#define GM_A1(Q, a, pc, p) Q##a##pc = *(const bf16x8*)((p) + (pc) * apiece + (a) * 512)
#define GM_ALOAD(Q, k0)                                                                                        \
  do {                                                                                                         \
    const __bf16* p_ = abase + (size_t)((k0) >> 5) * Rp * 32;                                                  \
    GM_A1(Q, 0, 0, p_); GM_A1(Q, 0, 1, p_); GM_A1(Q, 0, 2, p_);                                                \
    GM_A1(Q, 1, 0, p_); GM_A1(Q, 1, 1, p_); GM_A1(Q, 1, 2, p_);                                                \
  } while (0)
  // the six piece products of the output tiles (a, b), pieces 0 = high .. 2 = low,
  // stage by stage over the eight tiles: eight independent accumulation chains
  // in flight, each tile's products in the same order
#define GM_MF(Q, a, bb, pa, pb)                                                                                \
  acc[a][bb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Q##a##pa, bq##bb##pb, acc[a][bb], 0, 0, 0)
#define GM_STAGE(Q, pa, pb)                                                                                    \
  do {                                                                                                         \
    GM_MF(Q, 0, 0, pa, pb); GM_MF(Q, 1, 0, pa, pb); GM_MF(Q, 0, 1, pa, pb); GM_MF(Q, 1, 1, pa, pb);             \
    GM_MF(Q, 0, 2, pa, pb); GM_MF(Q, 1, 2, pa, pb); GM_MF(Q, 0, 3, pa, pb); GM_MF(Q, 1, 3, pa, pb);             \
  } while (0)
#define GM_BQ(bb)                                                                                              \
  const bf16x8 bq##bb##0 = *(const bf16x8*)&sB[0][16 * (bb) + fr][fk];                                        \
  const bf16x8 bq##bb##1 = *(const bf16x8*)&sB[1][16 * (bb) + fr][fk];                                        \
  const bf16x8 bq##bb##2 = *(const bf16x8*)&sB[2][16 * (bb) + fr][fk]
  // one chunk: spectrum tile to LDS, next chunk's loads issued, the 48 MFMAs
  // on the current fragments Q, the fold into FP64
#define GM_STEP(Q, N, k0)                                                                                      \
  do {                                                                                                         \
    __bf16(*__restrict__ sB)[TN][LDSK] = GM_SB(k0);                                                            \
    GM_LSTORE();                                                                                               \
    __syncthreads();                                                                                           \
    if ((k0) + TK < kend) GM_BLOAD((k0) + TK);                                                                \
    if ((k0) + 2 * TK < kend) GM_ALOAD(N, (k0) + 2 * TK);                                                       \
    {                                                                                                          \
      GM_BQ(0); GM_BQ(1); GM_BQ(2); GM_BQ(3);                                                                  \
      GM_STAGE(Q, 2, 0); GM_STAGE(Q, 1, 1); GM_STAGE(Q, 0, 2);                                                 \
      GM_STAGE(Q, 1, 0); GM_STAGE(Q, 0, 1); GM_STAGE(Q, 0, 0);                                                 \
    }                                                                                                          \
    _Pragma("unroll") for (int a_ = 0; a_ < 2; ++a_)                                                           \
    _Pragma("unroll") for (int b_ = 0; b_ < 4; ++b_) {                                                         \
      _Pragma("unroll") for (int e_ = 0; e_ < 4; ++e_) sum[a_][b_][e_] += (double)acc[a_][b_][e_];             \
      acc[a_][b_] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};                                                           \
    }                                                                                                          \
  } while (0)
  floatx4 acc[2][4];
  double sum[2][4][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int bb = 0; bb < 4; ++bb) {
      acc[a][bb] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int e = 0; e < 4; ++e) sum[a][bb][e] = 0.0;
    }
  bf16x8 qa00, qa01, qa02, qa10, qa11, qa12;
  bf16x8 qb00, qb01, qb02, qb10, qb11, qb12;
  bf16x8 qc00, qc01, qc02, qc10, qc11, qc12;
  GM_BLOAD(kbeg);
  GM_ALOAD(qa, kbeg);
  if (kbeg + TK < kend) GM_ALOAD(qb, kbeg + TK);
  for (int k0 = kbeg; k0 < kend; k0 += 3 * TK) {
    GM_STEP(qa, qc, k0);
    if (k0 + TK < kend) GM_STEP(qb, qa, k0 + TK);
    if (k0 + 2 * TK < kend) GM_STEP(qc, qb, k0 + 2 * TK);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int bb = 0; bb < 4; ++bb)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = r0 + 32 * w + a * 16 + (lane >> 4) * 4 + e;
        const int col = n0 + bb * 16 + (lane & 15);
        H[(((size_t)split * nsum + j) * Rp + row) * Np + col] = sum[a][bb][e] * fac;
      }
}
#undef GM_BLOAD
#undef GM_LSTORE
#undef GM_B_SRC
#undef GM_B_DST
#undef GM_A1
#undef GM_ALOAD
#undef GM_MF
#undef GM_STAGE
#undef GM_BQ
#undef GM_SB
#undef GM_STEP

// H[i] = sum over s < nsplit of Hp[s][i], in split order, i < n (nsplit > 1)
__global__ void k_sumsplits(size_t n, int nsplit, const double* __restrict__ Hp, double* __restrict__ H) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a = Hp[i];
  for (int sp = 1; sp < nsplit; ++sp) a += Hp[(size_t)sp * n + i];
  H[i] = a;
}

struct cd { double re, im; };
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd crs(cd a, double s) { return {a.re * s, a.im * s}; }

// ---- regional branch: wavenumber integrals --------------------------------

// the 8 distinct kernels of the 14 moment wavenumber integrals at k for
// Bessel (or Neumann) values b0..b2 of k x (wavenumberIntegrals.f90:353-373)
__device__ __forceinline__ void wnRows(double k, double x, double re, double b0, double b1, double b2, double* w) {
  w[0] = k * b0;                                          // kint 1, 2
  w[1] = 2.0 / re * b1;                                   // 3
  w[2] = k * 2.0 * b2;                                    // 4
  w[3] = k * (-k * re) * b1;                              // 5, 6
  w[4] = k * 2.0 * (b0 - b1 / (k * x));                   // 7, 11
  w[5] = k * 2.0 * k * re * (b1 - 2.0 * b2 / (k * x));    // 8, -12
  w[6] = k * 2.0 * b1 / (k * x);                          // 9, 13
  w[7] = k * 4.0 * k * re * b2 / (k * x);                 // 10, -14
}

// R d/dx of the eight rows (the kernels of the x-derivative integrals, moment
// kint 29..42 and force 15..21, momentWavenumberIntegrals :353-373): row 0 -> 3,
// 1 -> 4, 2 -> 5 of wnRows, and five rows of their own; the same (row,
// spectrum, factor) tables then give those integrals
__device__ __forceinline__ void wnRowsDeriv(double k, double x, double re, double b0, double b1, double b2, double* w) {
  const double kx = k * x, kr = k * re;
  w[0] = k * (-k * re) * b1;                                          // kint 29, 30
  w[1] = k * 2.0 * (b0 - b1 / kx);                                    // 31
  w[2] = k * 2.0 * k * re * (b1 - 2.0 * b2 / kx);                     // 32
  w[3] = -k * kr * kr * (b0 - b1 / kx);                               // 33, 34
  w[4] = -k * 2.0 * kr * (b1 - b2 / kx);                              // 35, 39
  w[5] = k * 2.0 * kr * kr * ((6.0 / (kx * kx) - 1.0) * b2 - b1 / kx);  // 36, -40
  w[6] = -k * 2.0 * kr * b2 / kx;                                     // 37, 41
  w[7] = k * 4.0 * kr * kr * (-3.0 * b2 / (kx * kx) + b1 / kx);        // 38, -42
}

// quadrature geometry of one receiver (momentWavenumberIntegrals, :333-346, 427)
struct WnGeom { double x, xdk, cF, sF; int j3; };
__device__ __forceinline__ WnGeom wnGeom(double xx, double dwn) {
  WnGeom g;
  g.x = xx < 3.0 / (dwn * 1.0e9) ? 3.0 / (dwn * 1.0e9) : xx;
  g.j3 = (int)((3.0 / g.x) / dwn) + 2;
  g.xdk = g.x * dwn;
  g.cF = (1.0 - cos(g.xdk)) / g.xdk / g.x;               // Filon weight of the end points (interior: 2 cF)
  g.sF = (1.0 - sin(g.xdk) / g.xdk) / g.x;               // weight of the Neumann end terms
  return g;
}

// Basis rows of the regional contraction: B_row[r][j] = quadrature weight(j) x
// kernel_row(k_j), j = 1-based wavenumber index, k_j = (j-1) dwn.  Weights:
// 2 <= j < j3: dwn (trapezoid); j3: dwn/2 + cF (trapezoid end, Filon start);
// j > j3: 2 cF (Filon interior).  They are those of a frequency whose last
// wavenumber lies beyond j3; k_wnfix corrects each frequency's last point.
// One thread per (receiver, wavenumber); layout [piece][row][Lp/32][Rp][32] as
// for the Legendre basis, with the thread index running in that order.
// deriv = 1: the rows of the x-derivative integrals (wnRowsDeriv).
__global__ void k_bessel_basis(int R, int Rp, int Lp, const double* __restrict__ xr, double dwn, double re,
                               __bf16* __restrict__ Bs, int deriv) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)Rp * Lp) return;
  const unsigned rest = (unsigned)(idx >> 5);
  const int r = (int)(rest % (unsigned)Rp), l = (int)(rest / (unsigned)Rp) * 32 + (int)(idx & 31);
  const size_t plane = (size_t)Rp * Lp;
  double w[kRows] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (r < R && l >= 1) {
    const double xx = xr[r];
    const WnGeom g = wnGeom(xx, dwn);
    const int j = l + 1;
    double q;
    if (j < g.j3) q = dwn;
    else if (j == g.j3) q = 0.5 * dwn + g.cF;
    else q = 2.0 * g.cF;
    const double k = l * dwn;
    const double b0 = j0(k * xx), b1 = j1(k * xx);
    const double b2 = -b0 + 2.0 / (k * xx) * b1;
    if (deriv) wnRowsDeriv(k, g.x, re, b0, b1, b2, w);
    else wnRows(k, g.x, re, b0, b1, b2, w);
#pragma unroll
    for (int i = 0; i < kRows; ++i) w[i] *= q;
  }
#pragma unroll
  for (int i = 0; i < kRows; ++i) {
    __bf16 a, b, c;
    split3(w[i], a, b, c);
    Bs[((size_t)0 * kRows + i) * plane + idx] = a;
    Bs[((size_t)1 * kRows + i) * plane + idx] = b;
    Bs[((size_t)2 * kRows + i) * plane + idx] = c;
  }
}

// End corrections and the R^2/(2 pi) factor of the regional sums, in place on
// H[j][Rp][Np], one thread per (receiver, frequency), n = the frequency's last
// wavenumber index:
//  - n > j3: the last point carries cF instead of 2 cF, plus the Neumann end
//    terms (G(n) Y-kernel(k_n) - G(j3) Y-kernel(k_j3)) sF;
//  - n == j3: trapezoid only, the point carries dwn/2 instead of dwn/2 + cF;
//  - n < j3: trapezoid only, the last point carries dwn/2 instead of dwn.
// ISTYP: 0 single force (7 sums), 1 moment tensor (14 sums); deriv = 1: the
// kernels of the x-derivative integrals (wnRowsDeriv).
template <int ISTYP>
__global__ void k_wnfix(int R, int nf, int Rp, int Np, const int* __restrict__ nwn, const long long* __restrict__ koff,
                        long long nkf, const float* __restrict__ gr, const float* __restrict__ gi,
                        const double* __restrict__ xr, double dwn, double re, double tapfrac, SpecMap smap,
                        double* __restrict__ H, int deriv) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= R * nf) return;
  const int r = id / nf, f = id % nf;
  const double xx = xr[r];
  const WnGeom g = wnGeom(xx, dwn);
  const int n = nwn[f];
  auto rows = [&](double k, double b0, double b1, double b2, double* w) {
    if (deriv) wnRowsDeriv(k, g.x, re, b0, b1, b2, w);
    else wnRows(k, g.x, re, b0, b1, b2, w);
  };
  auto rowsAt = [&](int j, double* w) {      // Bessel kernels at wavenumber index j
    const double k = (j - 1) * dwn;
    const double b0 = j0(k * xx), b1 = j1(k * xx);
    rows(k, b0, b1, -b0 + 2.0 / (k * xx) * b1, w);
  };
  auto spec = [&](int s, int j) -> cd {      // tapered spectrum s at wavenumber index j (absent: zero)
    if (smap.s[s] < 0) return cd{0.0, 0.0};
    const long long k = (long long)smap.s[s] * nkf + koff[f] + (j - 1);
    const double t = taperWn(j - 1, n, dwn, tapfrac);
    return cd{(double)gr[k] * t, (double)gi[k] * t};
  };
  double wn_[kRows], w3[kRows], yn[kRows], y3[kRows];
  const bool filon = g.j3 < n;
  rowsAt(n, wn_);
  if (filon) {
    const double wn3 = (g.j3 - 1) * dwn, wne = (n - 1) * dwn;
    double n0 = y0(wn3 * g.x), n1 = y1(wn3 * g.x);
    rows(wn3, n0, n1, -n0 + 2.0 / (wn3 * g.x) * n1, y3);
    n0 = y0(wne * g.x); n1 = y1(wne * g.x);
    rows(wne, n0, n1, -n0 + 2.0 / (wne * g.x) * n1, yn);
  } else if (n == g.j3) {
    rowsAt(g.j3, w3);
  }
  const double scale = re * re / (2.0 * kPi);
  // fully unrolled: the kernel-row index is then a constant and the four
  // 8-row tables stay in registers (a run-time index put them in scratch)
  constexpr int kN = ISTYP ? kSums : kForceSums;
  constexpr int kRowOf[2][kSums] = {{0, 1, 3, 4, 6, 4, 6}, {0, 0, 1, 2, 3, 3, 4, 5, 6, 7, 4, 5, 6, 7}};   // = c_row
#pragma unroll
  for (int j = 0; j < kN; ++j) {
    const int row = kRowOf[ISTYP][j], s = c_spec[ISTYP][j];
    const double fac = c_fac_wn[ISTYP][j];
    const size_t o = ((size_t)j * Rp + r) * Np + 2 * f;
    cd z{H[o], H[o + 1]};
    const cd gn = spec(s, n);
    double corr;
    if (filon) corr = -g.cF;                       // 2 cF/2 -> cF/2 at the last point
    else if (n == g.j3) corr = -g.cF;              // (dwn/2 + cF/2 x 2) -> dwn/2
    else corr = -0.5 * dwn;                        // dwn -> dwn/2
    const double kern = (n == g.j3 && !filon) ? w3[row] : wn_[row];
    z.re += fac * corr * kern * gn.re;
    z.im += fac * corr * kern * gn.im;
    if (filon) {
      const cd g3 = spec(s, g.j3);
      z.re += fac * (gn.re * yn[row] - g3.re * y3[row]) * g.sF;
      z.im += fac * (gn.im * yn[row] - g3.im * y3[row]) * g.sF;
    }
    H[o] = z.re * scale;
    H[o + 1] = z.im * scale;
  }
}


// Displacement spectra Z[r][3][nf2] (zero below nf1).
// forceSingleFrequencyDisplacement (istyp 0, fomt = Fr, Ft, Ff) or
// momentSingleFrequencyDisplacement (istyp 1, fomt = Mrr, Mtt, Mff, Mrt, Mrf,
// Mtf) (singleFrequencyDisplacement.f90:35-46, 108-125), time shift and the
// event and station filters (computeSyntheticSeismograms.f90:361-371; evf:
// complex [nf2], stf: complex [R][3][nf2], NULL: unity), velocity (:376-378),
// ZNE rotation (axesRotation.f90:560-568).
// Strain and stress spectra of one (r, f), from the 3 x nsum sums H[part][j]
// (part 0: the displacement sums, 1: on the radial-derivative spectra, 2: on
// the theta-derivative basis): displacement u and its derivatives d/dr u,
// 1/r d/dtheta u, 1/(r sin theta) d/dphi u (moment/force Single, RadDeriv,
// ThetaDeriv and PhiDeriv FrequencyDisplacement, singleFrequencyDisplacement.f90:
// 35-194), the spherical strain (:223-241) and, with elcon (A, C, F, L, N at the
// receiver), the stress (:253-266).  Z [R][3][nf2], E and T [R][3][3][nf2]
// complex, zero below nf1.
__global__ void k_strain(int R, int nf1, int nf2, int Rp, int Np, int nsum, const double* __restrict__ H,
                         const double* __restrict__ theta, const double* __restrict__ phi, double re,
                         const double* __restrict__ fm, const double* __restrict__ elcon, int istyp,
                         double* __restrict__ Z, double* __restrict__ E, double* __restrict__ T) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= R * nf2) return;
  const int r = id / nf2, jf = id % nf2 + 1;
  cd u[4][3], e[3][3], t[3][3];
  for (int a = 0; a < 3; ++a) {
    for (int p = 0; p < 4; ++p) u[p][a] = cd{0.0, 0.0};
    for (int b = 0; b < 3; ++b) { e[a][b] = cd{0.0, 0.0}; t[a][b] = cd{0.0, 0.0}; }
  }
  if (jf >= nf1) {
    const int f = jf - nf1;
    auto h = [&](int part, int j) {
      const size_t o = (((size_t)part * nsum + j) * Rp + r) * Np + 2 * f;
      return cd{H[o], H[o + 1]};
    };
    const double ph = phi[r], del = theta[r];
    const double sp = sin(ph), cp = cos(ph);
    const double rst = re * sin(del);
    if (istyp) {
      const double s2 = sin(2.0 * ph), c2 = cos(2.0 * ph);
      const double a = fm[3] * cp + fm[4] * sp, b = (fm[2] - fm[1]) * c2 - 2.0 * fm[5] * s2;
      const double c = fm[4] * cp - fm[3] * sp, d = (fm[2] - fm[1]) * s2 + 2.0 * fm[5] * c2;
      for (int part = 0; part < 3; ++part) {
        u[part][0] = cadd(cadd(cadd(crs(h(part, 0), fm[0]), crs(h(part, 1), fm[1] + fm[2])), crs(h(part, 2), a)),
                          crs(h(part, 3), b));
        u[part][1] = cadd(cadd(cadd(crs(h(part, 4), fm[0]), crs(h(part, 5), fm[1] + fm[2])),
                               crs(cadd(h(part, 6), h(part, 8)), a)), crs(cadd(h(part, 7), h(part, 9)), b));
        u[part][2] = cadd(crs(cadd(h(part, 10), h(part, 12)), c), crs(cadd(h(part, 11), h(part, 13)), d));
      }
      const double ap = -fm[3] * sp + fm[4] * cp, bp = -2.0 * (fm[2] - fm[1]) * s2 - 4.0 * fm[5] * c2;
      const double cq = -fm[4] * sp - fm[3] * cp, dq = 2.0 * (fm[2] - fm[1]) * c2 - 4.0 * fm[5] * s2;
      u[3][0] = cadd(crs(h(0, 2), ap), crs(h(0, 3), bp));
      u[3][1] = cadd(crs(cadd(h(0, 6), h(0, 8)), ap), crs(cadd(h(0, 7), h(0, 9)), bp));
      u[3][2] = cadd(crs(cadd(h(0, 10), h(0, 12)), cq), crs(cadd(h(0, 11), h(0, 13)), dq));
    } else {
      const double a = fm[1] * cp + fm[2] * sp, b = fm[1] * sp - fm[2] * cp;
      for (int part = 0; part < 3; ++part) {
        u[part][0] = cadd(crs(h(part, 0), fm[0]), crs(h(part, 1), a));
        u[part][1] = cadd(crs(h(part, 2), fm[0]), crs(cadd(h(part, 3), h(part, 4)), a));
        u[part][2] = crs(cadd(h(part, 5), h(part, 6)), b);
      }
      const double ap = -fm[1] * sp + fm[2] * cp, bp = fm[1] * cp + fm[2] * sp;
      u[3][0] = crs(h(0, 1), ap);
      u[3][1] = crs(cadd(h(0, 3), h(0, 4)), ap);
      u[3][2] = crs(cadd(h(0, 5), h(0, 6)), bp);
    }
    for (int a = 0; a < 3; ++a) {
      u[2][a] = cd{u[2][a].re / re, u[2][a].im / re};
      u[3][a] = cd{u[3][a].re / rst, u[3][a].im / rst};
    }
    const double rtandel = re * tan(del);
    const cd *z = u[0], *zr = u[1], *zt = u[2], *zf = u[3];
    auto dv = [](cd x, double y) { return cd{x.re / y, x.im / y}; };
    e[0][0] = zr[0];
    e[1][1] = cadd(zt[1], dv(z[0], re));
    e[2][2] = cadd(cadd(zf[2], dv(z[0], re)), dv(z[1], rtandel));
    e[0][1] = crs(csub(cadd(zt[0], zr[1]), dv(z[1], re)), 0.5);
    e[0][2] = crs(csub(cadd(zf[0], zr[2]), dv(z[2], re)), 0.5);
    e[1][2] = crs(csub(cadd(zf[1], zt[2]), dv(z[2], rtandel)), 0.5);
    e[2][1] = e[1][2];
    e[2][0] = e[0][2];
    e[1][0] = e[0][1];
    if (elcon) {
      cd el[5];
      for (int k = 0; k < 5; ++k) el[k] = cd{elcon[2 * k], elcon[2 * k + 1]};
      const cd tr2 = cadd(e[1][1], e[2][2]);
      t[0][0] = cadd(cmul(el[1], e[0][0]), cmul(el[2], tr2));
      t[1][1] = csub(cadd(cmul(el[2], e[0][0]), cmul(el[0], tr2)), cmul(crs(el[4], 2.0), e[2][2]));
      t[2][2] = csub(cadd(cmul(el[2], e[0][0]), cmul(el[0], tr2)), cmul(crs(el[4], 2.0), e[1][1]));
      t[0][1] = cmul(crs(el[3], 2.0), e[0][1]);
      t[0][2] = cmul(crs(el[3], 2.0), e[0][2]);
      t[1][2] = cmul(crs(el[4], 2.0), e[1][2]);
      t[1][0] = t[0][1];
      t[2][0] = t[0][2];
      t[2][1] = t[1][2];
    }
  }
  for (int a = 0; a < 3; ++a) {
    const size_t oz = (((size_t)r * 3 + a) * nf2 + jf - 1) * 2;
    Z[oz] = u[0][a].re; Z[oz + 1] = u[0][a].im;
    for (int b = 0; b < 3; ++b) {
      const size_t o = ((((size_t)r * 3 + a) * 3 + b) * nf2 + jf - 1) * 2;
      E[o] = e[a][b].re; E[o + 1] = e[a][b].im;
      if (T) { T[o] = t[a][b].re; T[o + 1] = t[a][b].im; }
    }
  }
}

__global__ void k_combine(int R, int nf1, int nf2, int Rp, int Np, const double* __restrict__ H,
                          const double* __restrict__ phi, const double* __restrict__ propdir,
                          const double* __restrict__ tred, const double* __restrict__ evf,
                          const double* __restrict__ stf, const double* __restrict__ mt, int velocity, double df,
                          double sigma, int istyp, double* __restrict__ Z) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= R * nf2) return;
  const int r = id / nf2, jf = id % nf2 + 1;
  cd u[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  if (jf >= nf1) {
    const int f = jf - nf1;
    const int nsum = istyp ? kSums : kForceSums;
    cd h[kSums];
#pragma unroll
    for (int j = 0; j < kSums; ++j) {
      const size_t o = ((size_t)j * Rp + r) * Np + 2 * f;
      h[j] = j < nsum ? cd{H[o], H[o + 1]} : cd{0.0, 0.0};
    }
    const double ph = phi[r];
    const double sp = sin(ph), cp = cos(ph);
    if (istyp) {
      const double s2 = sin(2.0 * ph), c2 = cos(2.0 * ph);
      const double a = mt[3] * cp + mt[4] * sp, b = (mt[2] - mt[1]) * c2 - 2.0 * mt[5] * s2;
      const double c = mt[4] * cp - mt[3] * sp, d = (mt[2] - mt[1]) * s2 + 2.0 * mt[5] * c2;
      u[0] = cadd(cadd(cadd(crs(h[0], mt[0]), crs(h[1], mt[1] + mt[2])), crs(h[2], a)), crs(h[3], b));
      u[1] = cadd(cadd(cadd(crs(h[4], mt[0]), crs(h[5], mt[1] + mt[2])), crs(cadd(h[6], h[8]), a)),
                  crs(cadd(h[7], h[9]), b));
      u[2] = cadd(crs(cadd(h[10], h[12]), c), crs(cadd(h[11], h[13]), d));
    } else {
      const double a = mt[1] * cp + mt[2] * sp, b = mt[1] * sp - mt[2] * cp;
      u[0] = cadd(crs(h[0], mt[0]), crs(h[1], a));
      u[1] = cadd(crs(h[2], mt[0]), crs(cadd(h[3], h[4]), a));
      u[2] = crs(cadd(h[5], h[6]), b);
    }
    const double om = 2.0 * kPi * (jf - 1) * df;
    if (tred) {
      const double tr = tred[r];
      double sn, cs;
      sincos(om * tr, &sn, &cs);
      const double ex = exp(sigma * tr);
      const cd tsf{cs * ex, sn * ex};
#pragma unroll
      for (int k = 0; k < 3; ++k) u[k] = cmul(u[k], tsf);
    }
    if (evf) {
      const cd e{evf[2 * (jf - 1)], evf[2 * (jf - 1) + 1]};
#pragma unroll
      for (int k = 0; k < 3; ++k) u[k] = cmul(u[k], e);
    }
    if (stf) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const size_t o = (((size_t)r * 3 + k) * nf2 + jf - 1) * 2;
        u[k] = cmul(u[k], cd{stf[o], stf[o + 1]});
      }
    }
    if (velocity) {
      const cd iz{sigma, om};      // i * (om - i sigma)
#pragma unroll
      for (int k = 0; k < 3; ++k) u[k] = cmul(u[k], iz);
    }
    if (propdir) {
      const double ca = cos(propdir[r]), sa = sin(propdir[r]);
      const cd ue = cadd(crs(u[1], sa), crs(u[2], ca)), un = cadd(crs(u[1], -ca), crs(u[2], sa));
      u[1] = un;
      u[2] = ue;
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t o = (((size_t)r * 3 + k) * nf2 + jf - 1) * 2;
    Z[o] = u[k].re;
    Z[o + 1] = u[k].im;
  }
}

// R-point DFT in registers, v_q <- sum_r v_r e^{+2 pi i q r / R}: radix-2
// decimation in time over the bit-reversed inputs.  The twiddles are powers of
// e^{2 pi i / 16}: w16[p] = e^{2 pi i p / 16}, p < 8, taken once per thread from
// the transform's table so that every angle uses the same pi.  All indices are
// compile-time constants once unrolled, so v stays in registers.
template <int R>
__device__ __forceinline__ void dftRegs(cd* v, const cd* w16) {
  constexpr int LOG = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
  cd u[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    int br = 0;
#pragma unroll
    for (int b = 0; b < LOG; ++b) br |= ((r >> b) & 1) << (LOG - 1 - b);
    u[br] = v[r];
  }
#pragma unroll
  for (int st = 1; st <= LOG; ++st) {
    const int half = 1 << (st - 1);
#pragma unroll
    for (int g = 0; g < R; g += 2 * half)
#pragma unroll
      for (int pos = 0; pos < half; ++pos) {
        const int i0 = g + pos, i1 = i0 + half;
        const cd x1 = pos == 0 ? u[i1] : cmul(w16[pos * (16 >> st)], u[i1]);   // e^{2 pi i pos / (2 half)}
        const cd x0 = u[i0];
        u[i0] = cadd(x0, x1);
        u[i1] = csub(x0, x1);
      }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = u[r];
}

// LDS slot of transform point i: one 16-byte pad after every 16 points, so that
// the first pass's writes at a stride of 16 points spread over the banks
__device__ __forceinline__ int fftSlot(int i) { return i + (i >> 4); }
constexpr int kFFTSlots = kMaxFFT / 2 + kMaxFFT / 32;

// One pass of the Stockham auto-sort transform over sx (M points, sub-transforms
// of length Ns done): work item j < M/R reads R points at stride M/R, turns them
// by e^{2 pi i (j mod Ns) r / (Ns R)}, transforms them in registers and writes
// them at stride Ns from (j / Ns) Ns R + j mod Ns.  Every point is read once and
// written once per pass, so the pass runs in place: all reads, barrier, all
// writes, barrier.  M / R <= blockDim for every pass of this kernel.  The
// turning factors w^r, w = e^{2 pi i (j mod Ns) / (Ns R)} = tw[(j mod Ns) 2M / (Ns R)],
// are formed as running products of w (relative error below 1e-15 for r < 16).
template <int R>
__device__ __forceinline__ void stockhamPass(cd* sx, int M, int Ns, const double* __restrict__ tw, const cd* w16) {
  const int nw = M / R, j = threadIdx.x;
  cd v[R];
  if (j < nw) {
    const int t = (j % Ns) * (2 * M / (Ns * R));
    const cd w{tw[2 * t], tw[2 * t + 1]};
    cd wr = w;
    v[0] = sx[fftSlot(j)];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      v[r] = cmul(sx[fftSlot(j + r * nw)], wr);
      if (r + 1 < R) wr = cmul(wr, w);
    }
    dftRegs<R>(v, w16);
  }
  __syncthreads();
  if (j < nw) {
    const int base = (j / Ns) * Ns * R + j % Ns;
#pragma unroll
    for (int r = 0; r < R; ++r) sx[fftSlot(base + r * Ns)] = v[r];
  }
  __syncthreads();
}

// One trace per block: spectrum X_j (j = nf1-1 .. nf2-1, Hermitian mirror,
// zero elsewhere) -> x_t = sum_j X_j e^{+2 pi i j t / ns}, t < ns, then
// ts = x df e^{sigma t dt} (transformFrequencyTime, frequencyTime.f90:55-113).
// Real output from an M = ns/2 point complex transform of
// Z_k = (X_k + conj X_{M-k}) + i W^k (X_k - conj X_{M-k}), W = e^{2 pi i / ns}:
// z_n = x_{2n} + i x_{2n+1}.  The M-point transform is a Stockham auto-sort
// FFT in radix-16 passes (and one radix 2, 4 or 8 pass for the rest of log2 M),
// each pass a 16-point DFT in registers between two LDS barriers: three passes
// for M = 4096 where radix 2 took twelve.
// ns: transform length, nsamp: samples kept per trace, every ndec-th (ndec = 1
// or, in the reference's decimation branch, ns / nsamp), dt: step of the
// transform (dts there).
__global__ void __launch_bounds__(256) k_ifft(int nf1, int nf2, int ns, int logm, const double* __restrict__ Z,
                                              const double* __restrict__ tw, double df, double sigma, double dt,
                                              int nsamp, int ndec, float* __restrict__ ts) {
  __shared__ cd sx[kFFTSlots];
  const int M = ns >> 1;
  const double* X = Z + (size_t)blockIdx.x * nf2 * 2;
  auto xin = [&](int j) -> cd {
    if (j < nf1 - 1 || j > nf2 - 1) return {0.0, 0.0};
    return {X[2 * j], X[2 * j + 1]};
  };
  // unrolled so that the loads of several k are in flight together; between
  // nf2 - 1 and M - nf2 + 1 both X_k and X_{M-k} vanish and nothing is loaded
#pragma unroll 4
  for (int k = threadIdx.x; k < M; k += blockDim.x) {
    cd z{0.0, 0.0};
    if (k <= nf2 - 1 || k >= M - nf2 + 1) {
      // the reference transforms the complex zero-frequency bin and keeps the
      // real part of the result, so only its real part reaches the samples
      const cd a0 = xin(k);
      const cd a = k == 0 ? cd{a0.re, 0.0} : a0;
      const cd bc = k == 0 ? cd{0.0, 0.0} : xin(M - k);
      const cd b{bc.re, -bc.im};
      const cd wk{tw[2 * k], tw[2 * k + 1]};
      const cd s = cadd(a, b), d = csub(a, b);
      const cd wd = cmul(wk, d);
      z = cd{s.re - wd.im, s.im + wd.re};            // s + i * wd
    }
    sx[fftSlot(k)] = z;
  }
  __syncthreads();
  // e^{2 pi i p / 16} = tw[p ns / 16]; for ns < 16 only the p that are
  // multiples of 16 / ns occur (the DFTs are at most M = ns / 2 points)
  cd w16[8];
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    const int q = p * ns;
    w16[p] = (q & 15) == 0 ? cd{tw[2 * (q >> 4)], tw[2 * (q >> 4) + 1]} : cd{1.0, 0.0};
  }
  int Ns = 1, left = logm;
  for (; left >= 4; left -= 4, Ns *= 16) stockhamPass<16>(sx, M, Ns, tw, w16);
  if (left == 3) stockhamPass<8>(sx, M, Ns, tw, w16);
  else if (left == 2) stockhamPass<4>(sx, M, Ns, tw, w16);
  else if (left == 1) stockhamPass<2>(sx, M, Ns, tw, w16);
  float* out = ts + (size_t)blockIdx.x * nsamp;
  if (ndec > 1 || nsamp != ns) {
    // decimation branch of transformFrequencyTime (frequencyTime.f90:104-109),
    // taken whenever the transform is longer than the trace (ndec may be 1):
    // sample i is x_{ndec i}, damped by exp(sigma dts i) as written there
    for (int i = threadIdx.x; i < nsamp; i += blockDim.x) {
      const int t = ndec * i;
      const cd z = sx[fftSlot(t >> 1)];
      out[i] = (float)(((t & 1) ? z.im : z.re) * df * exp(sigma * dt * i));
    }
    return;
  }
  // e^{sigma dt t} at t = 2n and 2n + 1, n = threadIdx.x + i blockDim.x: one
  // exponential per thread, then a product per step of 2 blockDim samples
  const double a = sigma * dt;
  double e0 = df * exp(a * (2 * (int)threadIdx.x));
  const double e1 = exp(a), estep = exp(a * (2 * (int)blockDim.x));
  for (int n = threadIdx.x; n < M; n += blockDim.x) {
    const cd z = sx[fftSlot(n)];
    out[2 * n] = (float)(z.re * e0);
    out[2 * n + 1] = (float)(z.im * (e0 * e1));
    e0 *= estep;
  }
}

inline size_t rup(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace gm_syn

using namespace gm_syn;

// ---------------------------------------------------------------------------
// host side

namespace {

struct SynPlan {
  int R, Rp, L1, Lp, nf, Np, nsp;
  int istyp, nsum;       // source type from nsp (0 force: 5 spectra, 7 sums; 1 moment: 10 spectra, 14 sums)
  int nsp_in;            // spectra per node of the input (nsp, or 20 / 10 with stresses or derivatives)
  SpecMap smap;          // input spectrum of each of the nsp planes (U, V, W)
  long long nkf;
  std::vector<int> nwnf;
  std::vector<long long> koff;
};

int plan_sums(gm_ctx* c, int nf1, int nf2, const int* nwn, int nsp, long long nkf, int nrcv, SynPlan& P,
              int flags = 0) {
  if (nf1 < 1 || nf2 < nf1) return gm_fail(c, GM_ERR_ARG, "invalid frequency index range %d..%d", nf1, nf2);
  if (!nwn) return gm_fail(c, GM_ERR_ARG, "nwn is required");
  const bool stress = (flags & GM_FLAG_STRESS_SPECTRA) != 0, deriv = (flags & GM_FLAG_DERIV_SPECTRA) != 0;
  const bool ocean = (flags & GM_FLAG_OCEAN_SPECTRA) != 0;
  if (stress && deriv) return gm_fail(c, GM_ERR_ARG, "a file has either stresses or derivatives");
  if (ocean && deriv) return gm_fail(c, GM_ERR_ARG, "files with derivatives have no receiver in the ocean");
  // numberGreenFKSpectra (greenFKSpectra.f90:216-220): dsvmask (U, V, W) gives 2 x 4 + 2 = 10 for a
  // moment source and 2 x 2 + 1 = 5 for a single force; with stresses or derivatives, 20 and 10;
  // a receiver in the ocean (ncsph = 2, nctor = 0: no S, no W, :129-133) 8 and 4, with stresses 12 and 6
  if (ocean) {
    const int nm = stress ? 12 : 8;
    if (nsp != nm && nsp != nm / 2)
      return gm_fail(c, GM_ERR_ARG, "spectra of a receiver in the ocean%s need nsp = %d (moment) or %d (force), got %d",
                     stress ? " with stresses" : "", nm, nm / 2, nsp);
  } else if (stress || deriv) {
    if (nsp != 20 && nsp != 10)
      return gm_fail(c, GM_ERR_ARG, "spectra with %s need nsp = 20 (moment) or 10 (force), got %d",
                     stress ? "stresses" : "derivatives", nsp);
  } else if (nsp != 10 && nsp != 5) {
    return gm_fail(c, GM_ERR_UNSUPPORTED,
                   "the sums need nsp = 10 (moment: U, V x 4 jumps, W x 2) or 5 (force: U, V x 2 jumps, W x 1), got %d", nsp);
  }
  if (nrcv < 1) return gm_fail(c, GM_ERR_ARG, "nrcv must be positive");
  P.nf = nf2 - nf1 + 1;
  P.nsp_in = nsp;
  P.istyp = ocean ? (nsp == (stress ? 12 : 8) ? 1 : 0) : (stress || deriv) ? (nsp == 20 ? 1 : 0) : (nsp == 10 ? 1 : 0);
  P.nsp = P.istyp ? 10 : 5;
  // positionDataGreenFKSpectra of U(j), V(j), W(j): njsph = 4 / 2, njtor = 2 / 1
  {
    const int nj = P.istyp ? 4 : 2, nt = P.istyp ? 2 : 1;
    const int vOff = (stress ? 2 : 1) * nj, wOff = 4 * nj * (stress || deriv ? 1 : 0) + (stress || deriv ? 0 : 2 * nj);
    for (int k = 0; k < 10; ++k) P.smap.s[k] = 0;
    for (int j = 0; j < nj; ++j) { P.smap.s[j] = j; P.smap.s[nj + j] = vOff + j; }
    for (int j = 0; j < nt; ++j) P.smap.s[2 * nj + j] = ocean ? -1 : wOff + j;
  }
  P.nsum = P.istyp ? kSums : kForceSums;
  P.nwnf.resize(P.nf);
  P.koff.resize(P.nf);
  long long off = 0;
  int lmax = 0;
  for (int f = 0; f < P.nf; ++f) {
    const int n = nwn[nf1 - 1 + f];
    if (n < 1 || n > 46000) return gm_fail(c, GM_ERR_ARG, "nwn[%d] = %d out of range [1, 46000]", nf1 + f, n);
    P.nwnf[f] = n;
    P.koff[f] = off;
    off += n;
    lmax = std::max(lmax, n - 1);
  }
  if (nkf < off) return gm_fail(c, GM_ERR_ARG, "nkf = %lld is smaller than the sum of nwn (%lld)", nkf, off);
  P.nkf = nkf;
  P.R = nrcv;
  P.Rp = (int)rup(nrcv, TM);
  P.L1 = lmax + 1;
  P.Lp = (int)rup(P.L1, TK);
  P.Np = (int)rup(2 * (size_t)P.nf, TN);
  return GM_OK;
}

struct Carve {
  size_t off = 0;
  size_t take(size_t bytes) { size_t o = off; off += rup(bytes, 256); return o; }
};

int ensure_sbuf(gm_ctx* c, size_t bytes) {
  if (bytes <= c->sbuf_size) return GM_OK;
  if (c->sbuf) { (void)hipFree(c->sbuf); c->sbuf = nullptr; c->sbuf_size = 0; }
  if (hipMalloc(&c->sbuf, bytes) != hipSuccess) return gm_fail(c, GM_ERR_NOMEM, "device allocation of %zu bytes failed", bytes);
  c->sbuf_size = bytes;
  return GM_OK;
}

// basis, spectra and the contraction into H; offsets of the carved buffer in `o`
struct SumBufs { size_t theta, Bs, Gs, H, nwn, koff, gr, gi, seeds, Hp; };

// degree splits of the contraction: long degree ranges with few (sum, tile)
// blocks (the regional wavenumber sums) are cut so that about 6000 blocks share
// the chip, each split at least 16 chunks of TK degrees
inline void contractSplits(const SynPlan& P, int& nsplit, int& kper) {
  const int nchunks = P.Lp / TK;
  const long blocks = (long)P.nsum * (P.Np / TN) * (P.Rp / TM);
  int s = (int)std::min<long>(8, std::max<long>(1, (6000 + blocks - 1) / blocks));
  s = std::max(1, std::min(s, nchunks / 16));
  kper = (nchunks + s - 1) / s;
  nsplit = (nchunks + kper - 1) / kper;
}

// degree segments of the Legendre basis: about 8 (at least 64 degrees each),
// seglen a multiple of 32 (Lp is one)
inline void legendreSegments(int Lp, int& nseg, int& seglen) {
  const int want = std::max(1, std::min(8, Lp / 64));
  seglen = (int)rup((size_t)((Lp + want - 1) / want), 32);
  nseg = (Lp + seglen - 1) / seglen;
}

// dwn == 0: harmonic degree sums at epicentral distances theta (radians);
// dwn > 0: wavenumber integrals at epicentral distances theta (metres), rearth = re
int run_sums(gm_ctx* c, const SynPlan& P, const double* theta, const float* gr, const float* gi, bool dev_in,
             char* base, const SumBufs& o, hipStream_t s, double dwn = 0.0, double re = 0.0) {
  HIPCHK(c, hipMemcpyAsync(base + o.theta, theta, sizeof(double) * P.R, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + o.nwn, P.nwnf.data(), sizeof(int) * P.nf, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + o.koff, P.koff.data(), sizeof(long long) * P.nf, hipMemcpyHostToDevice, s));
  const float *dgr = gr, *dgi = gi;
  if (!dev_in) {
    HIPCHK(c, hipMemcpyAsync(base + o.gr, gr, sizeof(float) * P.nsp_in * P.nkf, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(base + o.gi, gi, sizeof(float) * P.nsp_in * P.nkf, hipMemcpyHostToDevice, s));
    dgr = (const float*)(base + o.gr);
    dgi = (const float*)(base + o.gi);
  }
  const bool regional = dwn > 0.0;
  const double* dth = (const double*)(base + o.theta);
  HIPCHK(c, hipEventRecord(c->sev[0], s));
  if (regional) {
    const size_t nb = (size_t)P.Rp * P.Lp;
    hipLaunchKernelGGL(k_bessel_basis, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, P.R, P.Rp, P.Lp, dth, dwn,
                       re, (__bf16*)(base + o.Bs), 0);
  } else {
    int nseg, seglen;
    legendreSegments(P.Lp, nseg, seglen);
    double* seeds = (double*)(base + o.seeds);
    if (nseg > 1)
      hipLaunchKernelGGL(k_legendre_seed, dim3((P.Rp + 63) / 64, 3, nseg - 1), dim3(64), 0, s, P.R, P.Rp, P.L1 - 1,
                         dth, nseg, seglen, seeds);
    hipLaunchKernelGGL(k_legendre, dim3((P.Rp + 63) / 64, 3, nseg), dim3(64), 0, s, P.R, P.Rp, P.L1 - 1, P.Lp, dth,
                       (__bf16*)(base + o.Bs), nullptr, nullptr, nullptr, nullptr, nullptr, 0, nseg, seglen,
                       (const double*)seeds);
  }
  HIPCHK(c, hipEventRecord(c->sev[1], s));
  hipLaunchKernelGGL(k_gprep, dim3((unsigned)((P.Lp + 255) / 256), (unsigned)(P.Np / 2)), dim3(256), 0, s, P.nsp, P.nf,
                     P.Np, P.Lp,
                     (const int*)(base + o.nwn), (const long long*)(base + o.koff), P.nkf, dgr, dgi, 0.2, dwn,
                     P.istyp, P.smap, (__bf16*)(base + o.Gs));
  HIPCHK(c, hipEventRecord(c->sev[2], s));
  int nsplit, kper;
  contractSplits(P, nsplit, kper);
  double* Hout = (double*)(base + (nsplit > 1 ? o.Hp : o.H));
  hipLaunchKernelGGL(k_contract, dim3((P.Np / TN) * (P.Rp / TM) * P.nsum * nsplit), dim3(256), 0, s, P.Rp, P.Np, P.Lp,
                     P.nsp, (const __bf16*)(base + o.Bs), (const __bf16*)(base + o.Gs), Hout, regional ? 1 : 0,
                     P.istyp, nsplit, kper);
  if (nsplit > 1) {
    const size_t nh = (size_t)P.nsum * P.Rp * P.Np;
    hipLaunchKernelGGL(k_sumsplits, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, nh, nsplit,
                       (const double*)Hout, (double*)(base + o.H));
  }
  if (regional)
    hipLaunchKernelGGL(P.istyp ? k_wnfix<1> : k_wnfix<0>, dim3((P.R * P.nf + 255) / 256), dim3(256), 0, s, P.R, P.nf,
                       P.Rp, P.Np, (const int*)(base + o.nwn), (const long long*)(base + o.koff), P.nkf, dgr, dgi, dth,
                       dwn, re, 0.2, P.smap, (double*)(base + o.H), 0);
  HIPCHK(c, hipEventRecord(c->sev[3], s));
  HIPCHK(c, hipGetLastError());
  return GM_OK;
}

SumBufs carve_sums(const SynPlan& P, bool dev_in, Carve& cv) {
  SumBufs o;
  o.theta = cv.take(sizeof(double) * P.R);
  int nseg, seglen;
  legendreSegments(P.Lp, nseg, seglen);
  o.seeds = cv.take(sizeof(double) * 4 * 3 * (size_t)P.Rp * nseg);
  o.nwn = cv.take(sizeof(int) * P.nf);
  o.koff = cv.take(sizeof(long long) * P.nf);
  o.Bs = cv.take(sizeof(__bf16) * kPieces * kRows * (size_t)P.Rp * P.Lp);
  o.Gs = cv.take(sizeof(__bf16) * kPieces * (size_t)P.nsp * P.Np * P.Lp);
  o.H = cv.take(sizeof(double) * P.nsum * (size_t)P.Rp * P.Np);
  int nsplit, kper;
  contractSplits(P, nsplit, kper);
  o.Hp = nsplit > 1 ? cv.take(sizeof(double) * nsplit * P.nsum * (size_t)P.Rp * P.Np) : 0;
  o.gr = dev_in ? 0 : cv.take(sizeof(float) * P.nsp_in * P.nkf);
  o.gi = dev_in ? 0 : cv.take(sizeof(float) * P.nsp_in * P.nkf);
  return o;
}

}  // namespace

extern "C" int gm_legendre_batch(gm_ctx* c, int nrcv, const double* theta, int lmax, double* plm, double* dpdt,
                                 double* pdst, double* d2pdt2, double* dtpdst) {
  if (!c) return GM_ERR_ARG;
  if (nrcv < 1 || !theta) return gm_fail(c, GM_ERR_ARG, "nrcv must be positive and theta given");
  if (lmax < 0 || lmax > 46000) return gm_fail(c, GM_ERR_ARG, "lmax = %d out of range [0, 46000]", lmax);
  if (!plm || !dpdt || !pdst || !d2pdt2 || !dtpdst) return gm_fail(c, GM_ERR_ARG, "all five output arrays are required");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t n = (size_t)nrcv * 3 * (lmax + 1);
  Carve cv;
  const size_t ot = cv.take(sizeof(double) * nrcv);
  size_t oo[5];
  for (int i = 0; i < 5; ++i) oo[i] = cv.take(sizeof(double) * n);
  int rc = ensure_sbuf(c, cv.off);
  if (rc) return rc;
  char* base = (char*)c->sbuf;
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(base + ot, theta, sizeof(double) * nrcv, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_legendre, dim3((nrcv + 63) / 64, 3), dim3(64), 0, s, nrcv, nrcv, lmax, 0,
                     (const double*)(base + ot), nullptr, (double*)(base + oo[0]), (double*)(base + oo[1]),
                     (double*)(base + oo[2]), (double*)(base + oo[3]), (double*)(base + oo[4]), 1, 1, 1 << 30,
                     nullptr);
  HIPCHK(c, hipGetLastError());
  double* outs[5] = {plm, dpdt, pdst, d2pdt2, dtpdst};
  for (int i = 0; i < 5; ++i) HIPCHK(c, hipMemcpyAsync(outs[i], base + oo[i], sizeof(double) * n, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  return GM_OK;
}

static int sums_impl(gm_ctx* c, int istyp, int nf1, int nf2, const int* nwn, int nsp, long long nkf,
                     const float* gfkspr, const float* gfkspi, int nrcv, const double* theta, int flags, double* hdsum,
                     double dwn, double re) {
  if (!c) return GM_ERR_ARG;
  SynPlan P;
  int rc = plan_sums(c, nf1, nf2, nwn, nsp, nkf, nrcv, P, flags);
  if (rc) return rc;
  if (P.istyp != istyp)
    return gm_fail(c, GM_ERR_ARG, istyp ? "moment sums need a moment source's spectra (nsp = 10, or 20 with stresses or derivatives)"
                                        : "force sums need a force's spectra (nsp = 5, or 10 with stresses or derivatives)");
  if (!gfkspr || !gfkspi || !theta || !hdsum) return gm_fail(c, GM_ERR_ARG, "spectra, theta and hdsum are required");
  const bool dev = (flags & GM_FLAG_DEVICE_PTRS) != 0;
  HIPCHK(c, hipSetDevice(c->device));
  Carve cv;
  SumBufs o = carve_sums(P, dev, cv);
  if ((rc = ensure_sbuf(c, cv.off))) return rc;
  char* base = (char*)c->sbuf;
  hipStream_t s = c->stream;
  if ((rc = run_sums(c, P, theta, gfkspr, gfkspi, dev, base, o, s, dwn, re))) return rc;
  // H[j][Rp][Np] -> hdsum[j][R][nf] complex
  for (int j = 0; j < P.nsum; ++j)
    HIPCHK(c, hipMemcpy2DAsync(hdsum + (size_t)j * P.R * P.nf * 2, sizeof(double) * 2 * P.nf,
                               base + o.H + sizeof(double) * (size_t)j * P.Rp * P.Np, sizeof(double) * P.Np,
                               sizeof(double) * 2 * P.nf, P.R, dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  float ms;
  for (int i = 0; i < 3; ++i) { (void)hipEventElapsedTime(&ms, c->sev[i], c->sev[i + 1]); c->stiming[i] = ms; }
  c->stiming[3] = c->stiming[4] = 0.0;
  (void)hipEventElapsedTime(&ms, c->sev[0], c->sev[3]);
  c->stiming[5] = ms;
  return GM_OK;
}

static int synth_impl(gm_ctx* c, int nf1, int nf2, double df, double sigma, const int* nwn, int nsp, long long nkf,
                      const float* gfkspr, const float* gfkspi, int nrcv, const double* theta, const double* phi,
                      const double* propdir, const double* tred, const double* evfilter, const double* stafilter,
                      int istyp, const double* mt, int velocity, int nsamp, double dt, int flags, float* ts,
                      double* zsp, double dwn, double re) {
  if (!c) return GM_ERR_ARG;
  if (istyp != 0 && istyp != 1) return gm_fail(c, GM_ERR_ARG, "istyp must be 0 (FORCE) or 1 (MOMENT)");
  SynPlan P;
  int rc = plan_sums(c, nf1, nf2, nwn, nsp, nkf, nrcv, P, flags);
  if (rc) return rc;
  if (P.istyp != istyp)
    return gm_fail(c, GM_ERR_UNSUPPORTED, "istyp %d does not match nsp = %d spectra of this layout", istyp, nsp);
  if (!gfkspr || !gfkspi || !theta || !phi || !mt || !ts)
    return gm_fail(c, GM_ERR_ARG, "spectra, theta, phi, fomt and ts are required");
  if (!(df > 0.0) || !(dt > 0.0)) return gm_fail(c, GM_ERR_ARG, "df and dt must be positive");
  // transformFrequencyTime (frequencyTime.f90:68-85): transform length ns and
  // decimation ndec; the reference takes the first logarithm in single precision
  int nslog = (int)std::ceil((double)std::log((float)nf2) / std::log(2.0)) + 1;
  int ns = 1 << nslog, ndec = 1;
  double dt_fft = dt;
  if (ns <= nsamp) {
    ns = nsamp;
    nslog = (int)std::ceil(-std::log(df * dt) / std::log(2.0));
    if (nslog < 2 || (1 << nslog) != nsamp)
      return gm_fail(c, GM_ERR_ARG, "nsamp = %d must equal 2^ceil(-log2(df dt)) = %d", nsamp, nslog >= 2 ? 1 << nslog : 0);
  } else {
    dt_fft = 1.0 / (df * ns);
    ndec = (int)std::lround(dt / dt_fft);
    if (std::fabs(dt - ndec * dt_fft) > 1.e-5)
      return gm_fail(c, GM_ERR_ARG, "Non-integer ratio of desired dt and dt defined from frequency spectrum");
    // ndec = 1 is allowed: nsamp < ns samples at the transform's own step, the first nsamp kept
    if (ndec < 1 || (size_t)(nsamp - 1) * ndec >= (size_t)ns)
      return gm_fail(c, GM_ERR_ARG, "nsamp = %d samples at every %d-th of %d do not fit the transform", nsamp, ndec, ns);
  }
  if (ns > kMaxFFT) return gm_fail(c, GM_ERR_UNSUPPORTED, "transform length %d exceeds %d", ns, kMaxFFT);
  const bool dev = (flags & GM_FLAG_DEVICE_PTRS) != 0;
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  if (c->tw_n != ns) {
    if (c->dtw) { (void)hipFree(c->dtw); c->dtw = nullptr; c->tw_n = 0; }
    if (hipMalloc(&c->dtw, sizeof(double) * 2 * ns) != hipSuccess)
      return gm_fail(c, GM_ERR_NOMEM, "twiddle allocation failed");
    hipLaunchKernelGGL(k_twiddle, dim3((ns + 255) / 256), dim3(256), 0, s, ns, c->dtw);
    c->tw_n = ns;
  }
  Carve cv;
  SumBufs o = carve_sums(P, dev, cv);
  const size_t o_phi = cv.take(sizeof(double) * P.R), o_pd = cv.take(sizeof(double) * P.R);
  const size_t o_tr = cv.take(sizeof(double) * P.R), o_mt = cv.take(sizeof(double) * 6);
  const size_t o_ev = cv.take(sizeof(double) * 2 * (size_t)nf2);
  const size_t o_st = stafilter ? cv.take(sizeof(double) * 2 * 3 * (size_t)P.R * nf2) : 0;
  const size_t o_z = cv.take(sizeof(double) * 2 * 3 * (size_t)P.R * nf2);
  const size_t o_ts = dev ? 0 : cv.take(sizeof(float) * 3 * (size_t)P.R * nsamp);
  if ((rc = ensure_sbuf(c, cv.off))) return rc;
  char* base = (char*)c->sbuf;
  HIPCHK(c, hipMemcpyAsync(base + o_phi, phi, sizeof(double) * P.R, hipMemcpyHostToDevice, s));
  if (propdir) HIPCHK(c, hipMemcpyAsync(base + o_pd, propdir, sizeof(double) * P.R, hipMemcpyHostToDevice, s));
  if (tred) HIPCHK(c, hipMemcpyAsync(base + o_tr, tred, sizeof(double) * P.R, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + o_mt, mt, sizeof(double) * (P.istyp ? 6 : 3), hipMemcpyHostToDevice, s));
  if (evfilter) HIPCHK(c, hipMemcpyAsync(base + o_ev, evfilter, sizeof(double) * 2 * nf2, hipMemcpyHostToDevice, s));
  if (stafilter)
    HIPCHK(c, hipMemcpyAsync(base + o_st, stafilter, sizeof(double) * 2 * 3 * (size_t)P.R * nf2, hipMemcpyHostToDevice, s));
  if ((rc = run_sums(c, P, theta, gfkspr, gfkspi, dev, base, o, s, dwn, re))) return rc;
  double* dZ = (double*)(base + o_z);
  hipLaunchKernelGGL(k_combine, dim3((unsigned)(((size_t)P.R * nf2 + 255) / 256)), dim3(256), 0, s, P.R, nf1, nf2,
                     P.Rp, P.Np, (const double*)(base + o.H), (const double*)(base + o_phi),
                     propdir ? (const double*)(base + o_pd) : nullptr, tred ? (const double*)(base + o_tr) : nullptr,
                     evfilter ? (const double*)(base + o_ev) : nullptr,
                     stafilter ? (const double*)(base + o_st) : nullptr, (const double*)(base + o_mt),
                     velocity ? 1 : 0, df, sigma, P.istyp, dZ);
  HIPCHK(c, hipEventRecord(c->sev[4], s));
  float* dts = dev ? ts : (float*)(base + o_ts);
  int logm = 0;
  while ((1 << (logm + 1)) <= ns / 2) ++logm;
  hipLaunchKernelGGL(k_ifft, dim3(3 * P.R), dim3(256), 0, s, nf1, nf2, ns, logm, (const double*)dZ, c->dtw, df, sigma,
                     dt_fft, nsamp, ndec, dts);
  HIPCHK(c, hipEventRecord(c->sev[5], s));
  HIPCHK(c, hipGetLastError());
  if (!dev) HIPCHK(c, hipMemcpyAsync(ts, dts, sizeof(float) * 3 * (size_t)P.R * nsamp, hipMemcpyDeviceToHost, s));
  if (zsp)
    HIPCHK(c, hipMemcpyAsync(zsp, dZ, sizeof(double) * 2 * 3 * (size_t)P.R * nf2,
                             dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  float ms;
  for (int i = 0; i < 5; ++i) { (void)hipEventElapsedTime(&ms, c->sev[i], c->sev[i + 1]); c->stiming[i] = ms; }
  (void)hipEventElapsedTime(&ms, c->sev[0], c->sev[5]);
  c->stiming[5] = ms;
  return GM_OK;
}

// Strain spectra: three contractions of the 14 (moment) or 7 (force) sum
// pattern over the node's spectra with radial derivatives (the displacement
// spectra, the derivative spectra, and the displacement spectra against the
// theta-derivative basis give kint 1..14, 15..28 and 29..42 / 1..7, 8..14 and
// 15..21), then k_strain.
// dwn > 0: the regional branch (besselDisplacementSpectra, displacementSpectra.f90:125-184):
// theta holds the epicentral distances in metres, the sums are wavenumber integrals with the
// kernels of wnRows / wnRowsDeriv, and the strain takes the angle theta / rearth
static int strain_impl(gm_ctx* c, int nf1, int nf2, const int* nwn, int nsp, long long nkf, const float* gfkspr,
                       const float* gfkspi, int nrcv, const double* theta, const double* phi, double re, int istyp,
                       const double* fomt, const double* elcon, int flags, double* hdsum, double* zsp,
                       double* strainsp, double* stresssp, double dwn = 0.0, double rearth = 0.0) {
  if (!c) return GM_ERR_ARG;
  if (istyp != 0 && istyp != 1) return gm_fail(c, GM_ERR_ARG, "istyp must be 0 (FORCE) or 1 (MOMENT)");
  // numberGreenFKSpectra with derivflag = 1 (greenFKSpectra.f90:109-111, 216-220): U, V, UP, VP x njsph, W, WP x njtor
  if (nsp != (istyp ? 20 : 10))
    return gm_fail(c, GM_ERR_UNSUPPORTED, "strain spectra need nsp = %d (U, V, dU/dr, dV/dr, W, dW/dr), got %d",
                   istyp ? 20 : 10, nsp);
  if (!gfkspr || !gfkspi || !theta || !phi || !fomt || !strainsp)
    return gm_fail(c, GM_ERR_ARG, "spectra, theta, phi, fomt and strainsp are required");
  if (!(re > 0.0)) return gm_fail(c, GM_ERR_ARG, "receiver radius must be positive");
  const bool regional = dwn > 0.0;
  if (regional && !(rearth > 0.0)) return gm_fail(c, GM_ERR_ARG, "rearth must be positive");
  if (regional)
    for (int r = 0; r < nrcv; ++r)
      if (!(theta[r] > 0.0)) return gm_fail(c, GM_ERR_ARG, "epicentral distance %d must be positive", r);
  SynPlan P;
  int rc = plan_sums(c, nf1, nf2, nwn, istyp ? 10 : 5, nkf, nrcv, P);   // 10 / 5 planes per contraction
  if (rc) return rc;
  const bool dev = (flags & GM_FLAG_DEVICE_PTRS) != 0;
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const size_t plane_h = (size_t)P.nsum * P.Rp * P.Np;
  int nsplit, kper;
  contractSplits(P, nsplit, kper);
  int nseg, seglen;
  legendreSegments(P.Lp, nseg, seglen);
  Carve cv;
  const size_t o_th = cv.take(sizeof(double) * P.R), o_ph = cv.take(sizeof(double) * P.R);
  const size_t o_dl = regional ? cv.take(sizeof(double) * P.R) : o_th;     // the angles of the strain
  const size_t o_seeds = cv.take(sizeof(double) * 4 * 3 * (size_t)P.Rp * nseg);
  const size_t o_nwn = cv.take(sizeof(int) * P.nf), o_koff = cv.take(sizeof(long long) * P.nf);
  const size_t nbs = sizeof(__bf16) * kPieces * kRows * (size_t)P.Rp * P.Lp;
  const size_t ngs = sizeof(__bf16) * kPieces * (size_t)P.nsp * P.Np * P.Lp;
  const size_t o_B0 = cv.take(nbs), o_B1 = cv.take(nbs), o_G0 = cv.take(ngs), o_G1 = cv.take(ngs);
  const size_t o_H = cv.take(sizeof(double) * 3 * plane_h);
  const size_t o_Hp = nsplit > 1 ? cv.take(sizeof(double) * nsplit * plane_h) : 0;
  const size_t o_gr = dev ? 0 : cv.take(sizeof(float) * nsp * (size_t)nkf);
  const size_t o_gi = dev ? 0 : cv.take(sizeof(float) * nsp * (size_t)nkf);
  const size_t o_fm = cv.take(sizeof(double) * 6), o_el = cv.take(sizeof(double) * 14);
  const size_t nz = 2 * 3 * (size_t)P.R * nf2, ne = 2 * 9 * (size_t)P.R * nf2;
  const size_t o_Z = cv.take(sizeof(double) * nz), o_E = cv.take(sizeof(double) * ne);
  const size_t o_T = elcon ? cv.take(sizeof(double) * ne) : 0;
  if ((rc = ensure_sbuf(c, cv.off))) return rc;
  char* base = (char*)c->sbuf;
  HIPCHK(c, hipMemcpyAsync(base + o_th, theta, sizeof(double) * P.R, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + o_ph, phi, sizeof(double) * P.R, hipMemcpyHostToDevice, s));
  std::vector<double> delta;
  if (regional) {
    delta.resize(P.R);
    for (int r = 0; r < P.R; ++r) delta[r] = theta[r] / rearth;
    HIPCHK(c, hipMemcpyAsync(base + o_dl, delta.data(), sizeof(double) * P.R, hipMemcpyHostToDevice, s));
  }
  HIPCHK(c, hipMemcpyAsync(base + o_nwn, P.nwnf.data(), sizeof(int) * P.nf, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + o_koff, P.koff.data(), sizeof(long long) * P.nf, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + o_fm, fomt, sizeof(double) * (istyp ? 6 : 3), hipMemcpyHostToDevice, s));
  if (elcon) HIPCHK(c, hipMemcpyAsync(base + o_el, elcon, sizeof(double) * 14, hipMemcpyHostToDevice, s));
  const float *dgr = gfkspr, *dgi = gfkspi;
  if (!dev) {
    HIPCHK(c, hipMemcpyAsync(base + o_gr, gfkspr, sizeof(float) * nsp * (size_t)nkf, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(base + o_gi, gfkspi, sizeof(float) * nsp * (size_t)nkf, hipMemcpyHostToDevice, s));
    dgr = (const float*)(base + o_gr);
    dgi = (const float*)(base + o_gi);
  }
  const double* dth = (const double*)(base + o_th);
  HIPCHK(c, hipEventRecord(c->sev[0], s));
  double* seeds = (double*)(base + o_seeds);
  if (regional) {
    // the wavenumber rows and their x-derivatives (R d/dx)
    const size_t nb = (size_t)P.Rp * P.Lp;
    for (int d = 0; d < 2; ++d)
      hipLaunchKernelGGL(k_bessel_basis, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, P.R, P.Rp, P.Lp, dth,
                         dwn, rearth, (__bf16*)(base + (d ? o_B1 : o_B0)), d);
  } else {
    if (nseg > 1)
      hipLaunchKernelGGL(k_legendre_seed, dim3((P.Rp + 63) / 64, 3, nseg - 1), dim3(64), 0, s, P.R, P.Rp, P.L1 - 1,
                         dth, nseg, seglen, seeds);
    for (int mode = 0; mode <= 2; mode += 2)
      hipLaunchKernelGGL(k_legendre, dim3((P.Rp + 63) / 64, 3, nseg), dim3(64), 0, s, P.R, P.Rp, P.L1 - 1, P.Lp, dth,
                         (__bf16*)(base + (mode ? o_B1 : o_B0)), nullptr, nullptr, nullptr, nullptr, nullptr, mode,
                         nseg, seglen, (const double*)seeds);
  }
  HIPCHK(c, hipEventRecord(c->sev[1], s));
  // spectrum offsets (positionDataGreenFKSpectra with dsvmask U, V, UP, VP, W, WP)
  SpecMap mdisp, mder;
  if (istyp) {
    const int a[10] = {0, 1, 2, 3, 4, 5, 6, 7, 16, 17}, b[10] = {8, 9, 10, 11, 12, 13, 14, 15, 18, 19};
    for (int k = 0; k < 10; ++k) { mdisp.s[k] = a[k]; mder.s[k] = b[k]; }
  } else {
    const int a[10] = {0, 1, 2, 3, 8, 0, 0, 0, 0, 0}, b[10] = {4, 5, 6, 7, 9, 0, 0, 0, 0, 0};
    for (int k = 0; k < 10; ++k) { mdisp.s[k] = a[k]; mder.s[k] = b[k]; }
  }
  for (int g = 0; g < 2; ++g)
    hipLaunchKernelGGL(k_gprep, dim3((unsigned)((P.Lp + 255) / 256), (unsigned)(P.Np / 2)), dim3(256), 0, s, P.nsp,
                       P.nf, P.Np, P.Lp, (const int*)(base + o_nwn), (const long long*)(base + o_koff), (long long)nkf,
                       dgr, dgi, 0.2, regional ? dwn : 0.0, istyp, g ? mder : mdisp, (__bf16*)(base + (g ? o_G1 : o_G0)));
  HIPCHK(c, hipEventRecord(c->sev[2], s));
  // part 0: displacement spectra x basis; 1: derivative spectra x basis; 2: displacement spectra x theta-derivative basis
  const size_t pb[3] = {o_B0, o_B0, o_B1}, pg[3] = {o_G0, o_G1, o_G0};
  for (int part = 0; part < 3; ++part) {
    double* Hpart = (double*)(base + o_H) + part * plane_h;
    hipLaunchKernelGGL(k_contract, dim3((P.Np / TN) * (P.Rp / TM) * P.nsum * nsplit), dim3(256), 0, s, P.Rp, P.Np,
                       P.Lp, P.nsp, (const __bf16*)(base + pb[part]), (const __bf16*)(base + pg[part]),
                       nsplit > 1 ? (double*)(base + o_Hp) : Hpart, regional ? 1 : 0, istyp, nsplit, kper);
    if (nsplit > 1)
      hipLaunchKernelGGL(k_sumsplits, dim3((unsigned)((plane_h + 255) / 256)), dim3(256), 0, s, plane_h, nsplit,
                         (const double*)(base + o_Hp), Hpart);
    if (regional)      // end points, Neumann terms and R^2 / (2 pi) of this part's integrals
      hipLaunchKernelGGL(istyp ? k_wnfix<1> : k_wnfix<0>, dim3((P.R * P.nf + 255) / 256), dim3(256), 0, s, P.R, P.nf,
                         P.Rp, P.Np, (const int*)(base + o_nwn), (const long long*)(base + o_koff), (long long)nkf,
                         dgr, dgi, dth, dwn, rearth, 0.2, part == 1 ? mder : mdisp, Hpart, part == 2 ? 1 : 0);
  }
  HIPCHK(c, hipEventRecord(c->sev[3], s));
  hipLaunchKernelGGL(k_strain, dim3((unsigned)(((size_t)P.R * nf2 + 255) / 256)), dim3(256), 0, s, P.R, nf1, nf2,
                     P.Rp, P.Np, P.nsum, (const double*)(base + o_H), (const double*)(base + o_dl),
                     (const double*)(base + o_ph), re,
                     (const double*)(base + o_fm), elcon ? (const double*)(base + o_el) : nullptr, istyp,
                     (double*)(base + o_Z), (double*)(base + o_E), elcon ? (double*)(base + o_T) : nullptr);
  HIPCHK(c, hipEventRecord(c->sev[4], s));
  HIPCHK(c, hipGetLastError());
  const hipMemcpyKind kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (hdsum)          // H[part][j][Rp][Np] -> hdsum[part * nsum + j][R][nf] complex
    for (int j = 0; j < 3 * P.nsum; ++j)
      HIPCHK(c, hipMemcpy2DAsync(hdsum + (size_t)j * P.R * P.nf * 2, sizeof(double) * 2 * P.nf,
                                 base + o_H + sizeof(double) * (size_t)j * P.Rp * P.Np, sizeof(double) * P.Np,
                                 sizeof(double) * 2 * P.nf, P.R, kind, s));
  if (zsp) HIPCHK(c, hipMemcpyAsync(zsp, base + o_Z, sizeof(double) * nz, kind, s));
  HIPCHK(c, hipMemcpyAsync(strainsp, base + o_E, sizeof(double) * ne, kind, s));
  if (elcon && stresssp) HIPCHK(c, hipMemcpyAsync(stresssp, base + o_T, sizeof(double) * ne, kind, s));
  HIPCHK(c, hipEventRecord(c->sev[5], s));
  HIPCHK(c, hipStreamSynchronize(s));
  float ms;
  for (int i = 0; i < 5; ++i) { (void)hipEventElapsedTime(&ms, c->sev[i], c->sev[i + 1]); c->stiming[i] = ms; }
  (void)hipEventElapsedTime(&ms, c->sev[0], c->sev[5]);
  c->stiming[5] = ms;
  return GM_OK;
}

extern "C" int gm_strain_spectra(gm_ctx* c, int nf1, int nf2, const int* nwn, int nsp, long long nkf,
                                 const float* gfkspr, const float* gfkspi, int nrcv, const double* theta,
                                 const double* phi, double re, int istyp, const double* fomt, const double* elcon,
                                 int flags, double* hdsum, double* zsp, double* strainsp, double* stresssp) {
  return strain_impl(c, nf1, nf2, nwn, nsp, nkf, gfkspr, gfkspi, nrcv, theta, phi, re, istyp, fomt, elcon, flags,
                     hdsum, zsp, strainsp, stresssp);
}

extern "C" int gm_strain_spectra_regional(gm_ctx* c, int nf1, int nf2, const int* nwn, int nsp, long long nkf,
                                          const float* gfkspr, const float* gfkspi, int nrcv, const double* x,
                                          const double* phi, double re, double dwn, double rearth, int istyp,
                                          const double* fomt, const double* elcon, int flags, double* hdsum,
                                          double* zsp, double* strainsp, double* stresssp) {
  if (!c) return GM_ERR_ARG;
  if (!(dwn > 0.0)) return gm_fail(c, GM_ERR_ARG, "dwn must be positive");
  if (!x) return gm_fail(c, GM_ERR_ARG, "x is required");
  return strain_impl(c, nf1, nf2, nwn, nsp, nkf, gfkspr, gfkspi, nrcv, x, phi, re, istyp, fomt, elcon, flags, hdsum,
                     zsp, strainsp, stresssp, dwn, rearth);
}

extern "C" int gm_moment_sums(gm_ctx* c, int nf1, int nf2, const int* nwn, int nsp, long long nkf, const float* gfkspr,
                              const float* gfkspi, int nrcv, const double* theta, int flags, double* hdsum) {
  return sums_impl(c, 1, nf1, nf2, nwn, nsp, nkf, gfkspr, gfkspi, nrcv, theta, flags, hdsum, 0.0, 0.0);
}

extern "C" int gm_force_sums(gm_ctx* c, int nf1, int nf2, const int* nwn, int nsp, long long nkf, const float* gfkspr,
                             const float* gfkspi, int nrcv, const double* theta, int flags, double* hdsum) {
  return sums_impl(c, 0, nf1, nf2, nwn, nsp, nkf, gfkspr, gfkspi, nrcv, theta, flags, hdsum, 0.0, 0.0);
}

static int bessel_impl(gm_ctx* c, int istyp, int nf1, int nf2, const int* nwn, int nsp, long long nkf,
                       const float* gfkspr, const float* gfkspi, double dwn, double rearth, int nrcv, const double* x,
                       int flags, double* wnint) {
  if (!c) return GM_ERR_ARG;
  if (!(dwn > 0.0) || !(rearth > 0.0)) return gm_fail(c, GM_ERR_ARG, "dwn and rearth must be positive");
  for (int r = 0; r < nrcv && x; ++r)
    if (!(x[r] > 0.0)) return gm_fail(c, GM_ERR_ARG, "epicentral distance %d must be positive", r);
  return sums_impl(c, istyp, nf1, nf2, nwn, nsp, nkf, gfkspr, gfkspi, nrcv, x, flags, wnint, dwn, rearth);
}

extern "C" int gm_bessel_sums(gm_ctx* c, int nf1, int nf2, const int* nwn, int nsp, long long nkf, const float* gfkspr,
                              const float* gfkspi, double dwn, double rearth, int nrcv, const double* x, int flags,
                              double* wnint) {
  return bessel_impl(c, 1, nf1, nf2, nwn, nsp, nkf, gfkspr, gfkspi, dwn, rearth, nrcv, x, flags, wnint);
}

extern "C" int gm_force_bessel_sums(gm_ctx* c, int nf1, int nf2, const int* nwn, int nsp, long long nkf,
                                    const float* gfkspr, const float* gfkspi, double dwn, double rearth, int nrcv,
                                    const double* x, int flags, double* wnint) {
  return bessel_impl(c, 0, nf1, nf2, nwn, nsp, nkf, gfkspr, gfkspi, dwn, rearth, nrcv, x, flags, wnint);
}

extern "C" int gm_synthetics(gm_ctx* c, int nf1, int nf2, double df, double sigma, const int* nwn, int nsp,
                             long long nkf, const float* gfkspr, const float* gfkspi, int nrcv, const double* theta,
                             const double* phi, const double* propdir, const double* tred, const double* evfilter,
                             const double* stafilter, int istyp, const double* fomt, int velocity, int nsamp,
                             double dt, int flags, float* ts, double* zsp) {
  return synth_impl(c, nf1, nf2, df, sigma, nwn, nsp, nkf, gfkspr, gfkspi, nrcv, theta, phi, propdir, tred, evfilter,
                    stafilter, istyp, fomt, velocity, nsamp, dt, flags, ts, zsp, 0.0, 0.0);
}

extern "C" int gm_synthetics_regional(gm_ctx* c, int nf1, int nf2, double df, double sigma, const int* nwn, int nsp,
                                      long long nkf, const float* gfkspr, const float* gfkspi, double dwn,
                                      double rearth, int nrcv, const double* x, const double* phi,
                                      const double* propdir, const double* tred, const double* evfilter,
                                      const double* stafilter, int istyp, const double* fomt, int velocity, int nsamp,
                                      double dt, int flags, float* ts, double* zsp) {
  if (!c) return GM_ERR_ARG;
  if (!(dwn > 0.0) || !(rearth > 0.0)) return gm_fail(c, GM_ERR_ARG, "dwn and rearth must be positive");
  for (int r = 0; r < nrcv && x; ++r)
    if (!(x[r] > 0.0)) return gm_fail(c, GM_ERR_ARG, "epicentral distance %d must be positive", r);
  return synth_impl(c, nf1, nf2, df, sigma, nwn, nsp, nkf, gfkspr, gfkspi, nrcv, x, phi, propdir, tred, evfilter,
                    stafilter, istyp, fomt, velocity, nsamp, dt, flags, ts, zsp, dwn, rearth);
}

extern "C" int gm_event_filter(int filtype, const double* filspecs, int nf1, int nf2, double df, double sigma,
                               double* hfil) {
  if (filtype != GM_FILTER_BUTTERWORTH_BANDPASS)
    return gm_fail(nullptr, GM_ERR_ARG, "Unknown event filter type specified (%d)", filtype);
  if (!filspecs || !hfil || nf1 < 1 || nf2 < nf1 || !(df > 0.0))
    return gm_fail(nullptr, GM_ERR_ARG, "event filter: filspecs and hfil required, 1 <= nf1 <= nf2, df > 0");
  // high-pass (order filspecs[0], corner filspecs[1]) times low-pass (order
  // filspecs[2], corner filspecs[3]) Butterworth responses at the damped
  // frequencies z = (f - i sigma / 2 pi) / fc: each order-n factor is the
  // product over the unit roots e_j = e^{i pi (j - 1/2) / n} of z / (z - e_j)
  // (high-pass) or -i / (z - e_j) (low-pass).  The frequency is formed from
  // its index, not by the reference's running sum.
  const int nord[2] = {(int)std::lround(filspecs[0]), (int)std::lround(filspecs[2])};
  const double fc[2] = {filspecs[1], filspecs[3]};
  for (int k = 0; k < 2; ++k)
    if (nord[k] < 0 || (nord[k] > 0 && !(fc[k] > 0.0)))
      return gm_fail(nullptr, GM_ERR_ARG, "event filter: orders must be >= 0 and corner frequencies positive");
  for (int jf = 1; jf <= nf2; ++jf) {
    double hr = 0.0, hi = 0.0;
    if (jf >= nf1) {
      hr = 1.0;
      for (int k = 0; k < 2; ++k) {
        const double zr = (jf - 1) * df / fc[k], zi = -sigma / (2.0 * kPi) / fc[k];
        for (int j = 1; j <= nord[k]; ++j) {
          const double a = kPi * (j - 0.5) / nord[k];
          const double dr = zr - std::cos(a), di = zi - std::sin(a);
          const double nr = k == 0 ? zr : 0.0, ni = k == 0 ? zi : -1.0;     // numerator z or -i
          const double q = dr * dr + di * di;
          const double fr = (nr * dr + ni * di) / q, fi = (ni * dr - nr * di) / q;
          const double tr = hr * fr - hi * fi;
          hi = hr * fi + hi * fr;
          hr = tr;
        }
      }
    }
    hfil[2 * (jf - 1)] = hr;
    hfil[2 * (jf - 1) + 1] = hi;
  }
  return GM_OK;
}

extern "C" int gm_last_synth_timing(gm_ctx* c, double* ms6) {
  if (!c || !ms6) return GM_ERR_ARG;
  for (int i = 0; i < 6; ++i) ms6[i] = c->stiming[i];
  return GM_OK;
}

// ---------------------------------------------------------------------------
// Travel times from a ray table for the automatic time shift
// (getTravelTimes, f90/travelTimes.f90:128-388), one receiver per lane.  The
// table keeps the reference's FP32 arrays in Fortran order: rnod[nnod],
// rtp[ntp], slowness[ntp], delta and ttime (nnod, ntp) with element (jn, k) at
// (jn-1) + (k-1) nnod.  Indices below are 1-based as in the reference.  The
// arithmetic is left unfused so that each sum is formed in the reference's order.
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)

namespace {

// bisection for j with x(j) < xx <= x(j+1) on an ascending x (0 below, n
// above), comparisons turned for a descending x (f90/locatePoint.f90:43-111)
template <class V, class F>
__device__ int ttLocate(V xx, int n, F x) {
  int lo = 0, hi = n + 1;
  const bool asc = x(n) > x(1);
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (asc == (xx > x(mid))) lo = mid;
    else hi = mid;
  }
  return lo;
}

struct TtTable {
  int nnod, ntp, istele;
  const float *rnod, *rtp, *slow, *delta, *ttime;
};

// node bracket and linear weights of a radius, extrapolating up to 1e-5 r
// beyond the table: 0, or 1 / 2 for a radius below / above it
__device__ int ttBracket(const TtTable& T, double r, int& j, double& a, double& b) {
  const double tol = 1e-5;
  j = ttLocate((float)r, T.nnod, [&](int i) { return T.rnod[i - 1]; });
  if (j == 0) {
    if (fabs(T.rnod[0] - r) > tol * r) return 1;
    j = 1;
  } else if (j == T.nnod) {
    if (fabs(T.rnod[T.nnod - 1] - r) > tol * r) return 2;
    j = T.nnod - 1;
  }
  const float hf = T.rnod[j] - T.rnod[j - 1];
  a = ((double)T.rnod[j] - r) / (double)hf;
  b = 1.0 - a;
  return 0;
}

__global__ void __launch_bounds__(256) k_travel_times(TtTable T, double rs, int nrcv, const double* __restrict__ re,
                                                      const double* __restrict__ din, double* __restrict__ tt,
                                                      double* __restrict__ pout, int* __restrict__ status) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrcv) return;
  const double x = din[r];
  tt[r] = -1.0;
  pout[r] = 0.0;
  int js, je, st;
  double as, bs, ae, be;
  if ((st = ttBracket(T, rs, js, as, bs))) { status[r] = st; return; }
  auto rtp = [&](int i) { return T.rtp[i - 1]; };
  const int jtps = ttLocate((float)rs, T.ntp, rtp);
  if (jtps == 0) { status[r] = 3; return; }
  const int jtpr = ttLocate((float)re[r], T.ntp, rtp);
  if (jtpr == 0) { status[r] = 4; return; }
  if ((st = ttBracket(T, re[r], je, ae, be))) { status[r] = st + 4; return; }
  const int jtpmax = min(jtps, jtpr);
  if (jtpmax < 2) { status[r] = 7; return; }
  const int nn = T.nnod;
  auto dat = [&](const float* a, int jn, int k) { return (double)a[(jn - 1) + (size_t)(k - 1) * nn]; };
  auto p = [&](int k) { return (double)T.slow[k - 1]; };
  // distance and intercept time via turning point k: down from the source and
  // up to the receiver (sgn = 1), or the steep ray without a turn (sgn = -1)
  auto del = [&](int k, double sgn) {
    return as * dat(T.delta, js, k) + bs * dat(T.delta, js + 1, k) + sgn * (ae * dat(T.delta, je, k)) +
           sgn * (be * dat(T.delta, je + 1, k));
  };
  auto tau = [&](int k, double sgn) {
    return as * dat(T.ttime, js, k) + bs * dat(T.ttime, js + 1, k) + sgn * (ae * dat(T.ttime, je, k)) +
           sgn * (be * dat(T.ttime, je + 1, k)) - p(k) * del(k, sgn);
  };
  double dmax = del(1, 1.0), dmin = dmax;
  for (int k = 2; k <= jtpmax; ++k) {
    const double v = del(k, 1.0);
    dmax = fmax(dmax, v);
    dmin = fmin(dmin, v);
  }
  int jd;
  double ad, bd, sgn = 1.0;
  if (x >= dmin) {
    if (x > dmax) {   // beyond the deepest turning point: diffracted along it
      tt[r] = tau(1, 1.0) + p(1) * x;
      pout[r] = p(1);
      status[r] = 0;
      return;
    }
    if (T.istele) {
      jd = ttLocate(x, jtpmax, [&](int k) { return del(k, 1.0); });
    } else {   // the last sign change of del - x, searched from the end
      jd = jtpmax;
      double prod = 1.0;
      while (prod > 0.0 && jd > 1) {
        prod = (x - del(jd - 1, 1.0)) * (x - del(jd, 1.0));
        --jd;
      }
    }
    jd = max(1, min(jd, jtpmax - 1));
    const double h = del(jd + 1, 1.0) - del(jd, 1.0);
    ad = (del(jd + 1, 1.0) - x) / h;
    bd = 1.0 - ad;
  } else {
    if (T.istele) { status[r] = 8; return; }
    sgn = -1.0;
    jd = ttLocate(x, jtpmax, [&](int k) { return del(k, -1.0); });
    const bool left = jd == 0, right = jd == jtpmax;
    if (left) jd = 1;
    if (right) jd = jtpmax - 1;
    const double h = del(jd + 1, -1.0) - del(jd, -1.0);
    if ((left || right) && fabs(h) < 1e-4) {   // flat end: take the end point
      ad = left ? 1.0 : 0.0;
      bd = left ? 0.0 : 1.0;
    } else {
      ad = (del(jd + 1, -1.0) - x) / h;
      bd = 1.0 - ad;
    }
  }
  const double ta = ad * tau(jd, sgn) + bd * tau(jd + 1, sgn);
  const double pa = ad * p(jd) + bd * p(jd + 1);
  tt[r] = ta + pa * x;
  pout[r] = pa;
  status[r] = 0;
}

const char* ttMessage(int st) {
  switch (st) {
    case 1: return "source radius below ray table range";
    case 2: return "source radius above ray table range";
    case 3: return "no turning point below source";
    case 4: return "no turning point below receiver";
    case 5: return "receiver radius below ray table range";
    case 6: return "receiver radius above ray table range";
    case 7: return "fewer than two turning points below source and receiver";
    default: return "Delta out of turning point range";
  }
}

}  // namespace

extern "C" int gm_travel_times(gm_ctx* c, int nnod, int ntp, int istele, const float* rnod, const float* rtp,
                               const float* slowness, const float* delta, const float* ttime, double rs, int nrcv,
                               const double* re, const double* deltain, double* tt, double* slow, int* status) {
  if (!c) return GM_ERR_ARG;
  if (nnod < 2 || ntp < 2 || !rnod || !rtp || !slowness || !delta || !ttime)
    return gm_fail(c, GM_ERR_ARG, "ray table: nnod >= 2, ntp >= 2 and all five arrays required");
  if (nrcv < 0 || (nrcv > 0 && (!re || !deltain || !tt)))
    return gm_fail(c, GM_ERR_ARG, "nrcv >= 0 with re, deltain and tt required");
  if (!(rs > 0.0)) return gm_fail(c, GM_ERR_ARG, "source radius must be positive");
  if (nrcv == 0) return GM_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t tn = (size_t)nnod * ntp;
  Carve cv;
  const size_t orn = cv.take(sizeof(float) * nnod), ortp = cv.take(sizeof(float) * ntp),
               osl = cv.take(sizeof(float) * ntp), odl = cv.take(sizeof(float) * tn), otm = cv.take(sizeof(float) * tn),
               ore = cv.take(sizeof(double) * nrcv), odn = cv.take(sizeof(double) * nrcv),
               ott = cv.take(sizeof(double) * nrcv), op = cv.take(sizeof(double) * nrcv), ost = cv.take(sizeof(int) * nrcv);
  int rc = ensure_sbuf(c, cv.off);
  if (rc) return rc;
  char* base = (char*)c->sbuf;
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(base + orn, rnod, sizeof(float) * nnod, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + ortp, rtp, sizeof(float) * ntp, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + osl, slowness, sizeof(float) * ntp, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + odl, delta, sizeof(float) * tn, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + otm, ttime, sizeof(float) * tn, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + ore, re, sizeof(double) * nrcv, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(base + odn, deltain, sizeof(double) * nrcv, hipMemcpyHostToDevice, s));
  const TtTable T{nnod, ntp, istele ? 1 : 0, (const float*)(base + orn), (const float*)(base + ortp),
                  (const float*)(base + osl), (const float*)(base + odl), (const float*)(base + otm)};
  hipLaunchKernelGGL(k_travel_times, dim3((nrcv + 255) / 256), dim3(256), 0, s, T, rs, nrcv,
                     (const double*)(base + ore), (const double*)(base + odn), (double*)(base + ott),
                     (double*)(base + op), (int*)(base + ost));
  HIPCHK(c, hipGetLastError());
  std::vector<int> stv(nrcv);
  HIPCHK(c, hipMemcpyAsync(tt, base + ott, sizeof(double) * nrcv, hipMemcpyDeviceToHost, s));
  if (slow) HIPCHK(c, hipMemcpyAsync(slow, base + op, sizeof(double) * nrcv, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(stv.data(), base + ost, sizeof(int) * nrcv, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (status) {
    std::copy(stv.begin(), stv.end(), status);
    return GM_OK;
  }
  for (int r = 0; r < nrcv; ++r)
    if (stv[r]) return gm_fail(c, GM_ERR_ARG, "travel time of receiver %d: %s", r, ttMessage(stv[r]));
  return GM_OK;
}
