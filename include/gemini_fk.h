/*
 * gemini_fk.h — C ABI of the MI355X Green-FK engine (libgemini_fk.so).
 *
 * Drop-in boundary for the per-(omega, l) radial problem of GEMINI_UNIFIED and
 * its frequency loop.  Plain pointers and sizes only; every function returns an
 * int status (0 ok, > 0 error code, see GM_ERR_*), mirroring the reference's
 * error_message level-2 convention (f90/errorMessage.f90:45-258): the caller
 * checks the status and aborts or reports, the callee never terminates the
 * process.  gm_last_error() returns the message of the last failure.
 *
 * Array conventions follow the reference's Fortran column-major layouts so an
 * iso_c_binding caller can pass its own allocations with c_loc:
 *   complex values are interleaved (re, im) doubles ("double complex");
 *   per-(omega, l) outputs are indexed [node][component][k] in C order, i.e.
 *   Fortran (k, component, node), k = running index over all (frequency, l)
 *   pairs of the call, l fastest — the layout of gfre/gfim in
 *   f90/computeGreenFKSpectraForSynthetics.f90:473-489 and of the GFK file
 *   datasets GreenFKSpectraReal/Imag (nkfmax, nsp, nnod).
 * Node and layer indices are 1-based as in the reference.
 *
 * Threading: one gm_ctx per host thread; all work runs asynchronously on the
 * context's HIP stream and is synchronised before a call returns.
 */
#ifndef GEMINI_FK_H
#define GEMINI_FK_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct gm_ctx gm_ctx;

/* error codes */
#define GM_OK 0
#define GM_ERR_MINSTEP 1      /* "Steps size smaller than minimum" (episode/integrationStep.f90:177-180) */
#define GM_ERR_MAXSTEP 2      /* "Maximum number of steps reached" (episode/integrationStep.f90:186) */
#define GM_ERR_NODEIDX 3      /* invalid source node index (f90/toroidalMinors.f90:90-97) */
#define GM_ERR_CF 4           /* continued fraction did not converge (f90/initialValues.f90:515,572) */
#define GM_ERR_RANGE 5        /* node/discontinuity bookkeeping (episode/propagateOde.f90:190-200,314-320) */
#define GM_ERR_SPLINE 6       /* tridiagonal system could not be solved (f90/cubicSpline.f90:104) */
#define GM_ERR_STATE 7        /* solid-fluid state sequence (f90/spheroidalMinors.f90:268-271) */
#define GM_ERR_UNSUPPORTED 8  /* configuration outside the implemented branches */
#define GM_ERR_ARG 9          /* invalid argument */
#define GM_ERR_HIP 10         /* HIP runtime failure */
#define GM_ERR_NOMEM 11       /* device allocation failed */

/* attenuation modes (f90/complexElasticConstants.f90:118-131) */
#define GM_ELASTIC 0
#define GM_ATTENUATION_ONLY 1
#define GM_DISPERSION_ONLY 2
#define GM_ATTENUATION_AND_DISPERSION 3

/* flags of the batch calls */
#define GM_FLAG_DEVICE_PTRS 1  /* output pointers are device (HBM) pointers */
/* Spectrum layout of the synthetics calls' input (dsvmask of the GFK file, greenFKSpectra.f90:
 * 107-111, 204-212).  Neither flag: U, V (x njsph jumps), W (x njtor): nsp = 10 (moment) or 5
 * (force).  GM_FLAG_STRESS_SPECTRA: U, R, V, S and W, T (files written with STRESSFLAG,
 * dsvstep = 1): nsp = 20 or 10.  GM_FLAG_DERIV_SPECTRA: U, V, dU/dr, dV/dr and W, dW/dr (the
 * multi-source driver's files, derivflag = 1): nsp = 20 or 10.  GM_FLAG_OCEAN_SPECTRA: a
 * receiver in the ocean (ncsph = 2, nctor = 0, greenFKSpectra.f90:129-133): U, V, no W: nsp = 8
 * or 4; with GM_FLAG_STRESS_SPECTRA U, R, V: nsp = 12 or 6.  The displacement sums read U, V
 * and W wherever the layout puts them, an absent W as zero. */
#define GM_FLAG_STRESS_SPECTRA 2
#define GM_FLAG_DERIV_SPECTRA 4
#define GM_FLAG_OCEAN_SPECTRA 8
#define GM_FILTER_BUTTERWORTH_BANDPASS 1  /* event filter type 'BUTTERWORTH_BANDPASS' (f90/eventFilter.f90:56) */

/* options (gm_set_option) */
#define GM_OPT_VARIANT 1       /* 0: exact (no FMA contraction, bitwise operation order), 1: fast (FMA) */
#define GM_OPT_PERSISTENT 2    /* 0 (default): one lane per job, 1: persistent substep-scheduled kernels */
#define GM_OPT_UNIT 3          /* (omega, l) instances per work unit of the persistent kernels, 64..256
                                  (default 256) */
#define GM_OPT_WAVES 4         /* waves per SIMD the persistent kernels are built for: 1 or 2 */
#define GM_OPT_PROPOSE_EVERY 5 /* learned proposer: consult the network after every N-th accepted step,
                                  and at a sub-interval start once N accepted steps have passed since the
                                  last consultation (default 1: every step) */
#define GM_OPT_SPH_SPLIT 6     /* blocks per 256-instance work unit of the spheroidal kernel: 1, 2
                                  or 4 (default) */
#define GM_OPT_ORDER_TARGET 7  /* target try count of the step rule of the toroidal and spheroidal
                                  kernels without a network: 6 (default) is the reference's
                                  (complexBulirschStep.f90:346-359), 3..7 allowed; on the fused
                                  seven-sequence sweep every step costs the same, so 7 takes the
                                  largest steps */

#define GM_OPT_GRAPH 8         /* 1: the kernel stages of a batch with the plain step rule (no network,
                                  no step log) run as one hipGraph, captured on the first call and
                                  replayed while the next call's shapes, buffers and parameters are
                                  the same (results unchanged; gm_last_timing then reports the graph's
                                  time as its first and last entries); 0 (default): stage-by-stage
                                  launches */

#define GM_OPT_ADJOINT_PAIR 9  /* 1: the unit-jump Green functions of a (omega, l) from two adjoint
                                  integrations per leg instead of four: the four start values
                                  (unitJumpGreenFunctions.f90:306-330) span a plane, so the four
                                  adjoints are combinations of two (half the adjoint work; results
                                  differ from the reference's at the level of eps, not bitwise);
                                  0 (default): one adjoint per unit jump as the reference */

/* work counters returned in stats[10] (same meaning as the CPU oracle's) */
#define GM_NSTAT 10            /* propagations, accepted steps, rejected steps, MMP sequences,
                                  matvecs, sysmat evaluations, Neville cells, algorithmic flops,
                                  flagged zero determinants, integration sub-intervals */

/* Context on HIP device `device`.  Replaces the module-global state the
 * reference keeps between calls (f90/splineEarthmodelCoefficients.f90:28-41). */
int gm_ctx_create(int device, gm_ctx** out);
int gm_ctx_destroy(gm_ctx* ctx);
int gm_set_option(gm_ctx* ctx, int option, int value);
/* copies the last error message of ctx (or of a failed gm_ctx_create when ctx is NULL) */
int gm_last_error(gm_ctx* ctx, char* buf, int len);

/* Node earth model, SI units, knots bottom-up; iktop is 1-based.
 * Replaces createNodeEarthmodel's result (f90/nodeEarthmodel.f90:98-257).
 * Limits: nk <= 512 knots, nlay <= 32 layers, and the per-frequency spline
 * image that every block stages in LDS, (nk+1)*192 + (nlay+1)*12 bytes, must
 * fit the device's LDS per block (GM_ERR_UNSUPPORTED otherwise; at most 99 KB
 * within these limits, so always on MI355X's 160 KB).  Up to about 205 knots,
 * four spheroidal blocks share a CU. */
int gm_model_set(gm_ctx* ctx, int nk, const double* rk, const double* ro, const double* vpv, const double* vph,
                 const double* vsv, const double* vsh, const double* eta, const double* qkinv, const double* qminv,
                 int nlay, const int* iktop, const double* rb, const int* iflso, int nic, int noc, int fsflag,
                 int equidis, double rearth, double ric, double roc, double fref);

/* External radial nodes (ascending radii, state 1 = fluid, 1-based layer).
 * Replaces the external_radial_nodes object (f90/externalRadialNodes.f90:62-72). */
int gm_nodes_set(gm_ctx* ctx, int nnod, const double* rnod, const int* state, const int* layer);

/* Density ro[nnod], complex moduli elcon[nnod][7] (A, C, F, L, N, kappa, mu)
 * and inverse Q qinv[nnod][2] (QK, QM) at the nodes for frequency f and
 * attenuation mode attmode, host arrays.  Replaces
 * elasticConstantsExternalRadialNodes (f90/externalRadialNodes.f90:422-433),
 * which the reference calls with f = fref, 'ELASTIC' for the GFK file's
 * externalNodes group (:208-265). */
int gm_node_properties(gm_ctx* ctx, double f, int attmode, double* ro, double* elcon, double* qinv);

/* Toroidal and spheroidal minors for all (f_q, l), l = lmin[q] .. lmin[q]+nl[q]-1,
 * with omega = 2*pi32*f_q - i*sigma and dll1 = (sqrt(l(l+1))/R*R)^2.
 * Replaces computeToroidalMinors (f90/toroidalMinors.f90:71) and
 * computeSpheroidalMinors (f90/spheroidalMinors.f90:72) called in the degree loop
 * (f90/computeGreenFKSpectraForSynthetics.f90:410-429), including the
 * per-frequency computeSplineEarthmodelCoefficients (:391).
 * Outputs (any may be NULL), k over all instances:
 *   u_tor, d_tor  [nnod][2][k] complex        dettor [nnod][k] complex
 *   u_sph, d_sph  [nnod][6][k] complex        ulg, dlg [nnod][k]   detsph [nnod][k] complex
 *   status [k] (0 or GM_ERR_*), stats[GM_NSTAT] (accumulated). */
int gm_minors_batch(gm_ctx* ctx, int nfreq, const double* freqs, const int* lmin, const int* nl, double sigma,
                    double eps, int attmode, int jsl, int jsr, int flags, double* u_tor, double* d_tor,
                    double* dettor, double* u_sph, double* d_sph, double* ulg, double* dlg, double* detsph,
                    int* status, long long* stats);

/* Full Green-FK columns: minors, unit-jump Green functions with the spheroidal
 * adjoints, reciprocity, source terms and packing, for a receiver at node jre
 * and a MOMENT (istyp = 1) or FORCE (istyp = 0) source type.
 * Replaces the degree loop body of computeGreenFKSpectraForSynthetics
 * (f90/computeGreenFKSpectraForSynthetics.f90:410-494) together with
 * computeUnitJumpGreenFunctions (f90/unitJumpGreenFunctions.f90:62),
 * spheroidalReciprocity/toroidalReciprocity (f90/reciprocity.f90:82,169) and
 * oneNodeSourceTerms (f90/sourceTerms.f90:59).
 * gf: complex FP64 [nnod][nsp][k]; gfre/gfim: FP32 [nnod][nsp][k] (the file's
 * values); nsp = 10 (moment) or 5 (force): U, V (x njsph jumps), W (x njtor).
 * A receiver in the ocean (a fluid node of the top layer, :293-300): nsp = 8 or
 * 4, U and V = -R / (r om^2 rho) of the spheroidal jumps, r and rho taken at the
 * node as the reference does (:465); a receiver in the fluid core returns
 * GM_ERR_UNSUPPORTED.  GM_FLAG_STRESS_SPECTRA in flags: the columns with
 * stresses (STRESSFLAG, dsvstep = 1, :209-212), U, R, V, S (x njsph) and W, T
 * (x njtor), nsp = 20 or 10; in the ocean U, R, V and njsph zero S slots,
 * nsp = 16 or 8, of which a reader takes the first 12 or 6
 * (greenFKSpectra.f90:107-133).  Any output may be NULL. */
int gm_gfk_batch(gm_ctx* ctx, int nfreq, const double* freqs, const int* lmin, const int* nl, double sigma,
                 double eps, int attmode, int jre, int istyp, int flags, double* gf, float* gfre, float* gfim,
                 int* status, long long* stats);

/* The Green-FK columns of gm_gfk_batch with GM_FLAG_DEVICE_PTRS, allocating
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Every buffer of the batch lives in one per-device array
 * of 128 MiB (GM_NOALLOC_BYTES) that the library holds until the process ends;
 * a batch that needs more (about 560 B x nodes per (omega, l) plus 330 B x
 * knots per frequency) or has more than 1024 frequencies is rejected with
 * GM_ERR_ARG before any launch.  gf, gfre, gfim and status[k] are device
 * pointers (status may be NULL), freqs, lmin, nl and stats host pointers.  All
 * calls on one GPU share the array and it keeps the last call's contents:
 * calls that overlap on one GPU race. */
int gm_gfk_batch_noalloc(gm_ctx* ctx, int nfreq, const double* freqs, const int* lmin, const int* nl, double sigma,
                         double eps, int attmode, int jre, int istyp, double* gf, float* gfre, float* gfim,
                         int* status, long long* stats);

/* Regional (wavenumber) mode of gm_minors_batch / gm_gfk_batch: the instances
 * of frequency q are the 1-based wavenumber indices iwn = iwnmin[q] ..
 * iwnmin[q]+nwn[q]-1 with k = (iwn-1) dwn and dll1 = (k R)^2
 * (computeGreenFKSpectraForSynthetics.f90:410-413 with GLOBAL = 0; the
 * reference starts at iwn = 2 and writes zeros for k = 0).  Everything else is
 * as for the degree-based calls. */
int gm_minors_batch_wn(gm_ctx* ctx, int nfreq, const double* freqs, const int* iwnmin, const int* nwn, double dwn,
                       double sigma, double eps, int attmode, int jsl, int jsr, int flags, double* u_tor,
                       double* d_tor, double* dettor, double* u_sph, double* d_sph, double* ulg, double* dlg,
                       double* detsph, int* status, long long* stats);
int gm_gfk_batch_wn(gm_ctx* ctx, int nfreq, const double* freqs, const int* iwnmin, const int* nwn, double dwn,
                    double sigma, double eps, int attmode, int jre, int istyp, int flags, double* gf, float* gfre,
                    float* gfim, int* status, long long* stats);

/* The unit-jump Green functions of a jump at node jre themselves, for every
 * instance: computeUnitJumpGreenFunctions (f90/unitJumpGreenFunctions.f90:62-380)
 * after the minors, with the spheroidal adjoints.  Arguments as gm_gfk_batch
 * (no source type); outputs ujs complex [nnod][jump 4][component 4][k]
 * (ujgfsph(4, njs, nnod) per instance, njs = 4, or 2 for a jump in the ocean
 * with jumps 3, 4 zero) and ujt complex [nnod][jump 2][component 2][k]
 * (ujgftor, zero for a jump in the ocean), zero for an instance whose status
 * is not 0.  The per-(omega, l) Fortran module unitJumpGreenFunctions serves
 * the reference's calls from it. */
int gm_ujgf_batch(gm_ctx* ctx, int nfreq, const double* freqs, const int* lmin, const int* nl, double sigma,
                  double eps, int attmode, int jre, int flags, double* ujs, double* ujt, int* status,
                  long long* stats);
/* gm_ujgf_batch over wavenumber indices (regional mode, as gm_gfk_batch_wn). */
int gm_ujgf_batch_wn(gm_ctx* ctx, int nfreq, const double* freqs, const int* iwnmin, const int* nwn, double dwn,
                     double sigma, double eps, int attmode, int jre, int flags, double* ujs, double* ujt,
                     int* status, long long* stats);

/* Multi-source Green-FK columns with radial derivatives: the degree loop body
 * of computeGreenFKSpectraForASKI (f90/computeGreenFKSpectraForASKI.f90:
 * 458-590).  Minors over the source node range min(jsnod)..max(jsnod)
 * (computeToroidalMinors / computeSpheroidalMinors with jsl < jsr); then for
 * each of the nsnod + 1 sources (nsnod MOMENT (istyp = 1) or FORCE (0) sources
 * at the 1-based nodes jsnod[js], and a FORCE at the receiver node jre, which
 * must be one of jsnod) the unit-jump Green functions of a jump at that node
 * (computeUnitJumpGreenFunctions, f90/unitJumpGreenFunctions.f90:62), the
 * source terms there (oneNodeSourceTerms, f90/sourceTerms.f90:59), and at
 * every node the direct Green functions U, V and W with their radial
 * derivatives (computeDsvDerivatives, f90/dsvDerivatives.f90:38; at fluid
 * nodes V and dV/dr from getHorcompComplexFluid, f90/fluidOdeSystem.f90:171).
 * Outputs [nsnod+1][nnod][nspmax][k] (the reference's gfre(nwn, nspmax, nnod,
 * nsnod+1)), nspmax = 20 for istyp = 1 and 10 for istyp = 0; per node
 * isp = U(j) V(j) dU/dr(j) dV/dr(j) W(j) dW/dr(j) with njsph = 4 / 2 and
 * njtor = 2 / 1 jumps (moment / force), zero beyond a source's
 * 4 njsph + 2 njtor.  Source nodes lie in the solid mantle or in the ocean (a
 * source there has two (U, R) jumps and no toroidal motion, :359-376: 8 spectra);
 * a source in the fluid core or below the CMB returns GM_ERR_UNSUPPORTED. */
int gm_aski_batch(gm_ctx* ctx, int nfreq, const double* freqs, const int* lmin, const int* nl, double sigma,
                  double eps, int attmode, int nsnod, const int* jsnod, int jre, int istyp, int flags, double* gf,
                  float* gfre, float* gfim, int* status, long long* stats);

/* Kernel time of the last batch call in milliseconds (HIP events on the
 * context stream), per stage: 0 tables, 1 setup, 2 toroidal, 3 spheroidal,
 * 4 adjoint+assembly, 5 total. */
int gm_last_timing(gm_ctx* ctx, double* ms6);

/* Work counters of the last batch call restricted to one ODE system
 * (0 minors, 1 toroidal, 2 fluid, 3 radial, 4 adjoint), for per-kernel rooflines. */
int gm_last_system_stats(gm_ctx* ctx, int sys, long long* stats);

/* ---- learned step proposer (no counterpart in the reference) ----------------
 * A 9 -> 64 -> 64 -> 64 -> 2 FP32 MLP (ReLU) proposes log10(h/R) and the try
 * count for each Bulirsch-Stoer step of the spheroidal minors system; the
 * reference error test still accepts or rejects every step and a rejected step
 * falls back to the reference halving.  Features: log10 omega, l/1000, r/R,
 * vp, vs (km/s), rho (g/cc), log10 Q_mu, log10(h_prev/R), previous try count. */
int gm_mlp_nparams(void);
/* forward pass of n samples X[n][9] -> Y[n][2] on `device` */
int gm_mlp_forward(int device, const float* params, int n, const float* X, float* Y);
/* loss = mean_b (w0 (y0-t0)^2 + w1 (y1-t1)^2) and its parameter gradient */
int gm_mlp_grad(int device, const float* params, int n, const float* X, const float* T, float w0, float w1,
                float* loss, float* grad);
/* Adam training on the GPU; params in/out; losses[epochs] = mean loss per epoch */
int gm_mlp_train(int device, float* params, int n, const float* X, const float* T, float w0, float w1, int epochs,
                 int batch, float lr, unsigned seed, float* losses);
/* install or remove (enable = 0) the proposer of a context;
 * norm = mu[9], sd[9] of the features and mu[2], sd[2] of the outputs.
 * enable = 1: the network proposes step size and try count during the
 *   spheroidal integration (every GM_OPT_PROPOSE_EVERY accepted steps), on the
 *   sequential step engine;
 * enable = 2: per-instance policy: before the minors kernels the network is
 *   evaluated once per (omega, l) on its features at the spheroidal start
 *   radius (no step history), and its try count (5..7) becomes that
 *   instance's target of the reference step rule in the toroidal,
 *   spheroidal and adjoint kernels, which stay on the fused sweep; with
 *   kcheck = 1 its step size also becomes the first step of the spheroidal
 *   up pass (kcheck = 1 in the every-step mode: the first error test is at
 *   the proposed try count - 1).
 * In both modes the reference error test accepts or rejects every step. */
int gm_set_proposer(gm_ctx* ctx, const float* params, const float* norm, float safety, int kcheck, int enable);
/* run the plain minors integration over the given (f, l) (or, with dwn > 0,
 * the wavenumber indices of gm_minors_batch_wn) and record every
 * spheroidal-minors step: 9 features, log10(h/R), log10(err_k/eps) for the
 * error tests after 3..7 sequences (all 7 sequences are computed), accepted
 * try (0 = rejected), instance index, running step number of the instance,
 * log10(distance to the end of the sub-interval / R), 1 if the step was cut to
 * end the sub-interval: 20 floats per record */
int gm_collect_steplog(gm_ctx* ctx, int nfreq, const double* freqs, const int* lmin, const int* nl, double dwn,
                       double sigma, double eps, int attmode, int jsl, int jsr, int maxrec, float* records, int* nrec,
                       long long* stats);

/* ---- synthetics half: Legendre functions, degree sums, frequency-time ---- */

/* Unnormalised associated Legendre functions without the Condon-Shortley sign
 * for m = 0..2 and their theta derivatives at each theta[r] (radians), FP64.
 * Replaces computeLegendreFunctions (f90/legendreFunctions.f90:48-157).
 * Outputs [nrcv][3][lmax+1] each (m = 0 rows of pdst and dtpdst are zero). */
int gm_legendre_batch(gm_ctx* ctx, int nrcv, const double* theta, int lmax, double* plm, double* dpdt, double* pdst,
                      double* d2pdt2, double* dtpdst);

/* The 14 moment-tensor harmonic degree sums (kint 1..14) of one source node
 * for receivers at epicentral distances theta[r] (radians) and frequency
 * indices jf = nf1..nf2 (Fortran numbering, f = (jf-1) df) with nwn[jf-1]
 * degrees each.  gfkspr/gfkspi: FP32 [nsp = 10][nkf] (or 20 with
 * GM_FLAG_STRESS_SPECTRA or GM_FLAG_DERIV_SPECTRA), k over the frequencies
 * nf1..nf2 with degrees fastest (readDataNodeGreenFKSpectra's layout,
 * f90/greenFKSpectra.f90:171-212).  hdsum: complex [14][nrcv][nf2-nf1+1].
 * Replaces momentHarmonicDegreeSums with taperHarmonicDegreeSums
 * (f90/harmonicDegreeSums.f90:231-369) in the frequency loop of
 * sphericalHarmonicsDisplacementSpectra (f90/displacementSpectra.f90:49-106).
 * The contraction runs on the matrix cores in split bf16 with FP32
 * accumulation per 32 degrees and FP64 accumulation across them.
 * GM_FLAG_DEVICE_PTRS: gfkspr, gfkspi and hdsum are device pointers. */
int gm_moment_sums(gm_ctx* ctx, int nf1, int nf2, const int* nwn, int nsp, long long nkf, const float* gfkspr,
                   const float* gfkspi, int nrcv, const double* theta, int flags, double* hdsum);

/* The 7 single-force harmonic degree sums (kint 1..7, displacement) of one
 * source node, arguments as for gm_moment_sums with nsp = 5 spectra (U, V x 2
 * jumps, W x 1: the FK spectra of a FORCE source, istyp = 0; 10 with the
 * layout flags).
 * hdsum: complex [7][nrcv][nf2-nf1+1].  Replaces forceHarmonicDegreeSums with
 * taperHarmonicDegreeSums (f90/harmonicDegreeSums.f90:95-167, 356-369) in the
 * frequency loop of sphericalHarmonicsDisplacementSpectra
 * (f90/displacementSpectra.f90:49-106). */
int gm_force_sums(gm_ctx* ctx, int nf1, int nf2, const int* nwn, int nsp, long long nkf, const float* gfkspr,
                  const float* gfkspi, int nrcv, const double* theta, int flags, double* hdsum);

/* Seismograms of one source node: degree sums, then the moment-tensor or force
 * combination (evalSingleFrequencyDisplacementSpectra,
 * f90/displacementSpectra.f90:200-228) by the GFK file's source type istyp:
 * istyp = 1, moment source, nsp = 10, fomt = Mrr, Mtt, Mff, Mrt, Mrf, Mtf
 * (f90/singleFrequencyDisplacement.f90:108-125); istyp = 0, single force,
 * nsp = 5, fomt = Fr, Ft, Ff (:35-46); with GM_FLAG_STRESS_SPECTRA or
 * GM_FLAG_DERIV_SPECTRA in flags, nsp = 20 or 10 and U, V, W are read where that
 * layout puts them.  Then time shift by tred[r] (NULL: none),
 * event and station filters (evfilter: complex [nf2], stafilter: complex
 * [nrcv][3][nf2] for the components r, theta, phi, i.e. the reference's
 * stafilter(nf2, 3, nstat); NULL: unity; f90/computeSyntheticSeismograms.f90:
 * 367-371), velocity (velocity = 1) and
 * rotation to Z, N, E by propdir[r] (NULL: components r, theta, phi), then the
 * frequency-to-time transform to nsamp FP32 samples (transformFrequencyTime,
 * f90/frequencyTime.f90:55-113): nsamp = 2^ceil(-log2(df dt)), or, when nf2
 * needs more samples (ns = 2^(ceil(log2 nf2) + 1) > nsamp), every ndec-th
 * sample of the ns-sample transform, ndec = dt df ns an integer; ns <= 8192.  Replaces the station loop body of
 * f90/computeSyntheticSeismograms.f90:311-388 (without the travel-time
 * automatic shift: tred is given).
 * ts: [nrcv][3][nsamp]; zsp (optional): complex [nrcv][3][nf2].
 * GM_FLAG_DEVICE_PTRS: gfkspr, gfkspi, ts and zsp are device pointers. */
int gm_synthetics(gm_ctx* ctx, int nf1, int nf2, double df, double sigma, const int* nwn, int nsp, long long nkf,
                  const float* gfkspr, const float* gfkspi, int nrcv, const double* theta, const double* phi,
                  const double* propdir, const double* tred, const double* evfilter, const double* stafilter,
                  int istyp, const double* fomt, int velocity, int nsamp, double dt, int flags, float* ts,
                  double* zsp);

/* Strain (and stress) spectra of one source node for receivers at epicentral
 * distances theta[r] and azimuths phi[r] (radians) at radius re (m), from
 * spectra with radial derivatives (the multi-source driver's files,
 * derivflag = 1: nsp = 20 for istyp = 1 (U, V, dU/dr, dV/dr x 4 jumps, W,
 * dW/dr x 2) or 10 for istyp = 0 (x 2 and x 1), layout as gm_moment_sums).
 * Replaces sphericalHarmonicsDisplacementSpectra with its strainsp argument
 * (f90/displacementSpectra.f90:49-106, 200-228): momentHarmonicDegreeSums /
 * forceHarmonicDegreeSums for kint 1..42 / 1..21 (f90/harmonicDegreeSums.f90:
 * 95-351; with getIndexArrayHarmonicDegreeSums, :375-420, the stress sums
 * kint 43..46 / 22..23 are not requested by such files), the displacement and
 * its r, theta and phi derivatives and sphericalStrainSingleFrequencyDisplacement
 * (f90/singleFrequencyDisplacement.f90:35-241); with elcon (7 complex moduli
 * A, C, F, L, N, kappa, mu at the receiver; NULL: no stress),
 * sphericalStressSingleFrequencyDisplacement (:253-266).  The three degree-sum
 * contractions run on the matrix cores as in gm_moment_sums.
 * Outputs: hdsum (optional) complex [42 or 21][nrcv][nf2-nf1+1] (kint order);
 * zsp (optional) complex [nrcv][3][nf2]; strainsp and stresssp (optional,
 * with elcon) complex [nrcv][3][3][nf2] (the reference's zsp(nf2, 3) and
 * strainsp(nf2, 3, 3) per receiver), components r, theta, phi, zero below nf1.
 * GM_FLAG_DEVICE_PTRS: spectra and outputs are device pointers. */
int gm_strain_spectra(gm_ctx* ctx, int nf1, int nf2, const int* nwn, int nsp, long long nkf, const float* gfkspr,
                      const float* gfkspi, int nrcv, const double* theta, const double* phi, double re, int istyp,
                      const double* fomt, const double* elcon, int flags, double* hdsum, double* zsp, double* strainsp,
                      double* stresssp);

/* Regional (Bessel) branch: the 14 moment wavenumber integrals (kint 1..14) of
 * one source node at epicentral distances x[r] in metres, for spectra sampled
 * at k = (j-1) dwn, j = 1..nwn[jf-1] (j = 1 is k = 0 and is not summed).
 * Trapezoid rule below k x = 3, Filon's rule with Neumann end terms above,
 * cos^2 taper over the last 20 %, times rearth^2 / (2 pi).  Replaces
 * momentWavenumberIntegrals with taperWavenumberIntegrals
 * (f90/wavenumberIntegrals.f90:321-512) and computeBesselFunctions
 * (f90/besselFunctions.f90:33-51) in besselDisplacementSpectra
 * (f90/displacementSpectra.f90:119-184).  The sum over wavenumbers runs on
 * the matrix cores as for gm_moment_sums.  wnint: complex [14][nrcv][nf2-nf1+1]. */
int gm_bessel_sums(gm_ctx* ctx, int nf1, int nf2, const int* nwn, int nsp, long long nkf, const float* gfkspr,
                   const float* gfkspi, double dwn, double rearth, int nrcv, const double* x, int flags,
                   double* wnint);

/* The 7 single-force wavenumber integrals (kint 1..7) of one source node,
 * arguments as for gm_bessel_sums with nsp = 5 spectra (U, V x 2 jumps, W x 1).
 * Replaces forceWavenumberIntegrals (f90/wavenumberIntegrals.f90:122-250).
 * wnint: complex [7][nrcv][nf2-nf1+1]. */
int gm_force_bessel_sums(gm_ctx* ctx, int nf1, int nf2, const int* nwn, int nsp, long long nkf, const float* gfkspr,
                         const float* gfkspi, double dwn, double rearth, int nrcv, const double* x, int flags,
                         double* wnint);

/* gm_synthetics for the regional branch: receivers at epicentral distances
 * x[r] (metres) and azimuths phi[r]; otherwise as gm_synthetics. */
/* gm_strain_spectra for the regional branch (besselDisplacementSpectra with
 * strainsp, f90/displacementSpectra.f90:125-184): receivers at epicentral
 * distances x[r] (metres, > 0) and azimuths phi[r] at radius re, wavenumbers
 * k = (j-1) dwn.  momentWavenumberIntegrals / forceWavenumberIntegrals for
 * kint 1..42 / 1..21 (f90/wavenumberIntegrals.f90:122-477): the wavenumber
 * integrals of gm_bessel_sums on the displacement and on the radial-derivative
 * spectra, and the x-derivative integrals (kernels R d/dx of those rows); then
 * the strain of the epicentral angle x[r] / rearth as in gm_strain_spectra.
 * Arguments and outputs otherwise as gm_strain_spectra. */
int gm_strain_spectra_regional(gm_ctx* ctx, int nf1, int nf2, const int* nwn, int nsp, long long nkf,
                               const float* gfkspr, const float* gfkspi, int nrcv, const double* x, const double* phi,
                               double re, double dwn, double rearth, int istyp, const double* fomt, const double* elcon,
                               int flags, double* hdsum, double* zsp, double* strainsp, double* stresssp);

int gm_synthetics_regional(gm_ctx* ctx, int nf1, int nf2, double df, double sigma, const int* nwn, int nsp,
                           long long nkf, const float* gfkspr, const float* gfkspi, double dwn, double rearth,
                           int nrcv, const double* x, const double* phi, const double* propdir, const double* tred,
                           const double* evfilter, const double* stafilter, int istyp, const double* fomt, int velocity,
                           int nsamp, double dt, int flags, float* ts, double* zsp);

/* Event filter of the synthetics stage: computeEventfilter
 * (f90/eventFilter.f90:40-83), type GM_FILTER_BUTTERWORTH_BANDPASS with
 * filspecs = (high-pass order, high-pass corner Hz, low-pass order, low-pass
 * corner Hz): the product of highPassComplexOmegaButterworthFilter and
 * lowPassComplexOmegaButterworthFilter (f90/butterworthFilter.f90:49-96) at the
 * damped frequencies (jf-1) df - i sigma / (2 pi).  Order 0 is unity.
 * hfil: complex [nf2], zero below nf1.  Host computation, no context. */
int gm_event_filter(int filtype, const double* filspecs, int nf1, int nf2, double df, double sigma, double* hfil);

/* Travel times for the automatic time shift of the synthetics stage:
 * getTravelTimes (f90/travelTimes.f90:373-388, with setSourceInfoTravelTimes
 * :128-176 and getReceiverTravelTimes :180-367) for a source at radius rs and
 * receivers at radii re[r] and epicentral distances deltain[r] (radians), one
 * receiver per GPU lane.  The ray table is the reference's
 * raytable_travel_times (read by readTableTravelTimes, :68-124): FP32 rnod[nnod],
 * rtp[ntp], slowness[ntp] (angular slowness rtp / v) and delta, ttime (nnod, ntp)
 * in Fortran order; istele restricts turning points to the lower mantle.
 * Outputs tt[r] (seconds, -1 on failure) and, when slow is not NULL, the ray's
 * slowness.  status (optional) gets 0 or the failure per receiver: 1 / 2
 * source radius below / above the table, 3 no turning point below the source,
 * 4 none below the receiver, 5 / 6 receiver radius below / above the table,
 * 7 fewer than two usable turning points, 8 distance out of the turning-point
 * range with istele.  Without status a failure returns GM_ERR_ARG with the
 * reference's message.  The caller forms the reduction time tred = tt - tsbuf
 * (f90/computeSyntheticSeismograms.f90:344-352).  Host pointers. */
int gm_travel_times(gm_ctx* ctx, int nnod, int ntp, int istele, const float* rnod, const float* rtp,
                    const float* slowness, const float* delta, const float* ttime, double rs, int nrcv,
                    const double* re, const double* deltain, double* tt, double* slow, int* status);

/* Kernel time of the last gm_moment_sums / gm_synthetics call in milliseconds:
 * 0 Legendre basis, 1 spectra preparation, 2 contraction, 3 combination,
 * 4 frequency-time transform, 5 total. */
int gm_last_synth_timing(gm_ctx* ctx, double* ms6);

#ifdef __cplusplus
}
#endif
#endif
