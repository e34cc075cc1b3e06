/*
 * gemini_gfkfile.h — GFK file writer and reader (libgemini_gfkfile.so, host
 * code over the HDF5 C library).
 *
 * The Green frequency-wavenumber spectra file that
 * computeGreenFKSpectraForSynthetics writes (f90/computeGreenFKSpectraForSynthetics.f90:236-345,
 * 500-519) and that computeSyntheticSeismograms and the kernel programs read
 * (f90/greenFKSpectra.f90:71-167).  Layout, in the Fortran order the reference
 * writes (C order: dimensions reversed):
 *   attribute  identity                         string (fdate) of the run
 *   dataset    earthModelName, attenuationMode  scalar strings
 *   dataset    reals    float32 (6): rearth, re, sigma, df, dwn, fref
 *   group      externalNodes: rearth (1), radii (nnod), state (nnod), layer (nnod),
 *              density (nnod), inverseQ (2, nnod), realElcon / imagElcon (7, nnod)
 *   dataset    numberOfWavenumbersForFrequency  int32 (nf2), zero below nf1
 *   dataset    integers  int32 (8): nf1, nf2, nkfmax, nwnmax, dsvstep, derivflag,
 *              numtasks, global
 *   attribute  dataSpecificIntegers  int32 (5): istyp, ncsph, nctor, njsph, njtor
 *   dataset    GreenFKSpectraReal / GreenFKSpectraImag  float32 (nkfmax, nsp, nnod),
 *              frequency jf at k offset sum(nwn(1:jf-1)), degrees fastest
 * The spectra of one call of gm_gfk_batch over consecutive frequencies with
 * l = 0 .. nwn(jf)-1 ([node][component][k] in C order) are one hyperslab of
 * these datasets.  The multi-file layout of the ASKI variant (a .meta file
 * with sourceNodeRadii, data files per source node;
 * f90/computeGreenFKSpectraForASKI.f90:277-406) is written with the same calls.
 *
 * Strings are HDF5 Fortran strings (space padded) of the trimmed length, as
 * writeStringHDFWrapper / writeStringAttributeHDFWrapper create them
 * (f90/hdfWrapper.f90:230-328).  Every function returns 0 or a GM_ERR_* code
 * and keeps a message for gm_gfkfile_last_error.
 */
#ifndef GEMINI_GFKFILE_H
#define GEMINI_GFKFILE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct gm_gfkfile gm_gfkfile;

#define GM_ERR_IO 12           /* HDF5 call failed (file missing, dataset absent, wrong shape) */

/* strings of a GFK file (gm_gfkfile_read_string) */
#define GM_GFK_IDENTITY 0
#define GM_GFK_EARTH_MODEL_NAME 1
#define GM_GFK_ATTENUATION_MODE 2

/* Creates (truncates) path for writing.  Replaces createFileParallelAccessHDFWrapper
 * (f90/hdfWrapper.f90:126-144); one writer per file. */
int gm_gfkfile_create(const char* path, gm_gfkfile** out);
/* Opens path read-only (openFileRoHDFWrapper). */
int gm_gfkfile_open(const char* path, gm_gfkfile** out);
/* Closes the file and frees the handle (NULL is a no-op). */
int gm_gfkfile_close(gm_gfkfile* f);
/* Message of the last failure on f, or of a failed create/open when f is NULL. */
int gm_gfkfile_last_error(gm_gfkfile* f, char* buf, int len);

/* Root attribute 'identity' (f90/computeGreenFKSpectraForSynthetics.f90:242-248). */
int gm_gfkfile_write_identity(gm_gfkfile* f, const char* identity);
/* earthModelName, attenuationMode, reals[6], numberOfWavenumbersForFrequency
 * nwn[nf2] (entries below nf1 zero) and integers[8]
 * (f90/computeGreenFKSpectraForSynthetics.f90:252-287).  integers[2] (nkfmax)
 * must equal the sum of nwn and integers[1] (nf2) the length of nwn. */
int gm_gfkfile_write_meta(gm_gfkfile* f, const char* earth_model_name, const char* attenuation_mode,
                          const double* reals, const int* integers, const int* nwn);
/* Group externalNodes (writeHDFExternalRadialNodes, f90/externalRadialNodes.f90:208-265):
 * node_rearth, radii rnod[nnod], state, layer, density ro[nnod], qinv[nnod][2]
 * and elcon[nnod][7] complex (gm_node_properties at fref, ELASTIC); every real
 * stored as float32. */
int gm_gfkfile_write_nodes(gm_gfkfile* f, double node_rearth, int nnod, const double* rnod, const int* state,
                           const int* layer, const double* ro, const double* qinv, const double* elcon);
/* Dataset sourceNodeRadii rsnod[n] of the ASKI meta file
 * (f90/computeGreenFKSpectraForASKI.f90:305-316). */
int gm_gfkfile_write_source_nodes(gm_gfkfile* f, int n, const double* rsnod);
/* Attribute dataSpecificIntegers dsi[5] and the empty spectra datasets
 * (nkfmax, nsp, nnod), nkfmax = sum(nwn[nf1-1 .. nf2-1])
 * (f90/computeGreenFKSpectraForSynthetics.f90:318-335).  Unwritten values read 0. */
int gm_gfkfile_create_spectra(gm_gfkfile* f, const int* dsi, int nsp, int nnod, int nf1, int nf2, const int* nwn);
/* Frequencies jf1..jf2 (1-based indices, nf1 <= jf1 <= jf2 <= nf2): gfre/gfim
 * [nnod][nsp][cnt], cnt = sum(nwn(jf1:jf2)), at k offset sum(nwn(1:jf1-1))
 * (f90/computeGreenFKSpectraForSynthetics.f90:502-519; any order of calls). */
int gm_gfkfile_write_frequencies(gm_gfkfile* f, int jf1, int jf2, const float* gfre, const float* gfim);

/* readMetaGreenFKSpectra (f90/greenFKSpectra.f90:71-137), one path each. */
int gm_gfkfile_read_string(gm_gfkfile* f, int which, char* buf, int len);
int gm_gfkfile_read_meta(gm_gfkfile* f, double* reals, int* integers);
/* nwn[n], n >= nf2 = integers[1] */
int gm_gfkfile_read_nwn(gm_gfkfile* f, int n, int* nwn);
int gm_gfkfile_read_dsi(gm_gfkfile* f, int* dsi);
/* number of external nodes (size of externalNodes/radii) */
int gm_gfkfile_node_count(gm_gfkfile* f, int* nnod);
/* readExternalRadialNodes (f90/externalRadialNodes.f90:271-330): the float32
 * values widened to double; any output may be NULL. */
int gm_gfkfile_read_nodes(gm_gfkfile* f, double* node_rearth, double* rnod, int* state, int* layer, double* ro,
                          double* qinv, double* elcon);
/* size and values of sourceNodeRadii (rsnod NULL: size only) */
int gm_gfkfile_read_source_nodes(gm_gfkfile* f, int* n, double* rsnod);
/* readDataNodeGreenFKSpectra (f90/greenFKSpectra.f90:146-167) for node jr
 * (1-based): spectra isp = 1..nsp, k = k0 .. k0+nk-1 (0-based) into gfkr/gfki
 * [nsp][nk]; the reference reads k0 = 0, nk = nkfmax. */
int gm_gfkfile_read_node(gm_gfkfile* f, int jr, int nsp, long long k0, long long nk, float* gfkr, float* gfki);

/* numberGreenFKSpectra (f90/greenFKSpectra.f90:210-215) from integers[8] and
 * dsi[5] (dsvmask of readMetaGreenFKSpectra :109-133). */
int gm_gfk_number_of_spectra(const int* integers, const int* dsi, int* nsp);

#ifdef __cplusplus
}
#endif
#endif
