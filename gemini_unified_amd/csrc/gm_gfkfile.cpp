// GFK file writer and reader over the HDF5 C library (libgemini_gfkfile.so).
// Host code only: the spectra arrive from gm_gfk_batch in the file's layout,
// so every write is one hyperslab per call.
#include <hdf5.h>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/gemini_fk.h"
#include "../../include/gemini_gfkfile.h"

struct gm_gfkfile {
  hid_t fid = -1;
  hid_t dre = -1, dim = -1;        // spectra datasets
  bool writable = false;
  int nf1 = 0, nf2 = 0, nsp = 0, nnod = 0;
  std::vector<long long> koff;     // koff[jf] = sum(nwn(1:jf-1)), jf = 1 .. nf2+1
  std::string err;
};

static std::string g_open_err;

static int fail(gm_gfkfile* f, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (f) f->err = buf; else g_open_err = buf;
  return code;
}

// HDF5 prints its error stack to stderr unless told otherwise; failures are
// reported through the return codes instead
static void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

// Fortran string type of the trimmed length (H5T_STRING_F: space padded)
static hid_t fortran_string(const char* s, size_t* n) {
  size_t len = strlen(s);
  while (len > 0 && s[len - 1] == ' ') --len;       // len_trim
  *n = len;
  hid_t t = H5Tcopy(H5T_FORTRAN_S1);
  if (t < 0) return t;
  if (H5Tset_size(t, len > 0 ? len : 1) < 0) { H5Tclose(t); return -1; }
  return t;
}

// rank-1..3 dataset of native type `mem` under loc, C dimensions dims
static int write_array(gm_gfkfile* f, hid_t loc, const char* name, hid_t file_type, hid_t mem, int rank,
                       const hsize_t* dims, const void* data) {
  hid_t sp = H5Screate_simple(rank, dims, nullptr);
  if (sp < 0) return fail(f, GM_ERR_IO, "dataspace of %s", name);
  hid_t ds = H5Dcreate2(loc, name, file_type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (ds < 0) { H5Sclose(sp); return fail(f, GM_ERR_IO, "cannot create dataset %s", name); }
  herr_t e = H5Dwrite(ds, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  H5Dclose(ds);
  H5Sclose(sp);
  return e < 0 ? fail(f, GM_ERR_IO, "cannot write dataset %s", name) : GM_OK;
}

static std::vector<float> to_float(const double* x, size_t n) {
  std::vector<float> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = (float)x[i];
  return v;
}

extern "C" int gm_gfkfile_create(const char* path, gm_gfkfile** out) {
  quiet();
  if (!out || !path) return fail(nullptr, GM_ERR_ARG, "path and handle must not be NULL");
  *out = nullptr;
  hid_t fid = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (fid < 0) return fail(nullptr, GM_ERR_IO, "cannot create %s", path);
  gm_gfkfile* f = new gm_gfkfile();
  f->fid = fid;
  f->writable = true;
  *out = f;
  return GM_OK;
}

extern "C" int gm_gfkfile_open(const char* path, gm_gfkfile** out) {
  quiet();
  if (!out || !path) return fail(nullptr, GM_ERR_ARG, "path and handle must not be NULL");
  *out = nullptr;
  hid_t fid = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (fid < 0) return fail(nullptr, GM_ERR_IO, "cannot open %s", path);
  gm_gfkfile* f = new gm_gfkfile();
  f->fid = fid;
  *out = f;
  return GM_OK;
}

extern "C" int gm_gfkfile_close(gm_gfkfile* f) {
  if (!f) return GM_OK;
  herr_t e = 0;
  if (f->dre >= 0) e |= H5Dclose(f->dre);
  if (f->dim >= 0) e |= H5Dclose(f->dim);
  if (f->fid >= 0) e |= H5Fclose(f->fid);
  delete f;
  return e < 0 ? fail(nullptr, GM_ERR_IO, "closing the file failed") : GM_OK;
}

extern "C" int gm_gfkfile_last_error(gm_gfkfile* f, char* buf, int len) {
  if (!buf || len <= 0) return GM_ERR_ARG;
  const std::string& s = f ? f->err : g_open_err;
  snprintf(buf, (size_t)len, "%s", s.c_str());
  return GM_OK;
}

static int need_writer(gm_gfkfile* f) {
  if (!f) return GM_ERR_ARG;
  if (!f->writable) return fail(f, GM_ERR_ARG, "file is open read-only");
  return GM_OK;
}

extern "C" int gm_gfkfile_write_identity(gm_gfkfile* f, const char* identity) {
  quiet();
  if (int rc = need_writer(f)) return rc;
  if (!identity) return fail(f, GM_ERR_ARG, "identity must not be NULL");
  size_t n;
  hid_t t = fortran_string(identity, &n);
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t a = (t >= 0 && sp >= 0) ? H5Acreate2(f->fid, "identity", t, sp, H5P_DEFAULT, H5P_DEFAULT) : -1;
  std::string v(identity, n);
  v.resize(n > 0 ? n : 1, ' ');
  herr_t e = a >= 0 ? H5Awrite(a, t, v.data()) : -1;
  if (a >= 0) H5Aclose(a);
  if (sp >= 0) H5Sclose(sp);
  if (t >= 0) H5Tclose(t);
  return e < 0 ? fail(f, GM_ERR_IO, "cannot write attribute identity") : GM_OK;
}

static int write_string(gm_gfkfile* f, const char* name, const char* val) {
  size_t n;
  hid_t t = fortran_string(val, &n);
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t ds = (t >= 0 && sp >= 0) ? H5Dcreate2(f->fid, name, t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT) : -1;
  std::string v(val, n);
  v.resize(n > 0 ? n : 1, ' ');
  herr_t e = ds >= 0 ? H5Dwrite(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) : -1;
  if (ds >= 0) H5Dclose(ds);
  if (sp >= 0) H5Sclose(sp);
  if (t >= 0) H5Tclose(t);
  return e < 0 ? fail(f, GM_ERR_IO, "cannot write string %s", name) : GM_OK;
}

extern "C" int gm_gfkfile_write_meta(gm_gfkfile* f, const char* earth_model_name, const char* attenuation_mode,
                                     const double* reals, const int* integers, const int* nwn) {
  quiet();
  if (int rc = need_writer(f)) return rc;
  if (!earth_model_name || !attenuation_mode || !reals || !integers || !nwn)
    return fail(f, GM_ERR_ARG, "meta arguments must not be NULL");
  const int nf1 = integers[0], nf2 = integers[1];
  if (nf1 < 1 || nf2 < nf1) return fail(f, GM_ERR_ARG, "frequency indices nf1=%d, nf2=%d", nf1, nf2);
  long long sum = 0;
  for (int j = 0; j < nf2; ++j) {
    if (nwn[j] < 0) return fail(f, GM_ERR_ARG, "negative wavenumber count at frequency %d", j + 1);
    sum += nwn[j];
  }
  if (sum != integers[2])
    return fail(f, GM_ERR_ARG, "nkfmax=%d differs from the sum %lld of the wavenumber counts", integers[2], sum);
  if (int rc = write_string(f, "earthModelName", earth_model_name)) return rc;
  if (int rc = write_string(f, "attenuationMode", attenuation_mode)) return rc;
  std::vector<float> r = to_float(reals, 6);
  hsize_t d6 = 6, d8 = 8, dn = (hsize_t)nf2;
  if (int rc = write_array(f, f->fid, "reals", H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, 1, &d6, r.data())) return rc;
  if (int rc = write_array(f, f->fid, "numberOfWavenumbersForFrequency", H5T_NATIVE_INT, H5T_NATIVE_INT, 1, &dn, nwn))
    return rc;
  return write_array(f, f->fid, "integers", H5T_NATIVE_INT, H5T_NATIVE_INT, 1, &d8, integers);
}

extern "C" int gm_gfkfile_write_nodes(gm_gfkfile* f, double node_rearth, int nnod, const double* rnod,
                                      const int* state, const int* layer, const double* ro, const double* qinv,
                                      const double* elcon) {
  quiet();
  if (int rc = need_writer(f)) return rc;
  if (nnod < 1 || !rnod || !state || !layer || !ro || !qinv || !elcon)
    return fail(f, GM_ERR_ARG, "node arrays must not be NULL and nnod >= 1");
  hid_t g = H5Gcreate2(f->fid, "externalNodes", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (g < 0) return fail(f, GM_ERR_IO, "cannot create group externalNodes");
  const hsize_t d1 = 1, dn = (hsize_t)nnod, dq[2] = {(hsize_t)nnod, 2}, de[2] = {(hsize_t)nnod, 7};
  const float re = (float)node_rearth;
  std::vector<float> fr = to_float(rnod, nnod), fro = to_float(ro, nnod), fq = to_float(qinv, 2 * (size_t)nnod);
  std::vector<float> er(7 * (size_t)nnod), ei(7 * (size_t)nnod);
  for (size_t i = 0; i < 7 * (size_t)nnod; ++i) { er[i] = (float)elcon[2 * i]; ei[i] = (float)elcon[2 * i + 1]; }
  int rc = write_array(f, g, "rearth", H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, 1, &d1, &re);
  if (!rc) rc = write_array(f, g, "radii", H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, 1, &dn, fr.data());
  if (!rc) rc = write_array(f, g, "state", H5T_NATIVE_INT, H5T_NATIVE_INT, 1, &dn, state);
  if (!rc) rc = write_array(f, g, "layer", H5T_NATIVE_INT, H5T_NATIVE_INT, 1, &dn, layer);
  if (!rc) rc = write_array(f, g, "density", H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, 1, &dn, fro.data());
  if (!rc) rc = write_array(f, g, "inverseQ", H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, 2, dq, fq.data());
  if (!rc) rc = write_array(f, g, "realElcon", H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, 2, de, er.data());
  if (!rc) rc = write_array(f, g, "imagElcon", H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, 2, de, ei.data());
  H5Gclose(g);
  return rc;
}

extern "C" int gm_gfkfile_write_source_nodes(gm_gfkfile* f, int n, const double* rsnod) {
  quiet();
  if (int rc = need_writer(f)) return rc;
  if (n < 1 || !rsnod) return fail(f, GM_ERR_ARG, "source node radii must not be empty");
  std::vector<float> v = to_float(rsnod, n);
  hsize_t d = (hsize_t)n;
  return write_array(f, f->fid, "sourceNodeRadii", H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, 1, &d, v.data());
}

extern "C" int gm_gfkfile_create_spectra(gm_gfkfile* f, const int* dsi, int nsp, int nnod, int nf1, int nf2,
                                         const int* nwn) {
  quiet();
  if (int rc = need_writer(f)) return rc;
  if (!dsi || !nwn || nsp < 1 || nnod < 1 || nf1 < 1 || nf2 < nf1)
    return fail(f, GM_ERR_ARG, "invalid spectra dimensions (nsp=%d, nnod=%d, nf1=%d, nf2=%d)", nsp, nnod, nf1, nf2);
  if (f->dre >= 0) return fail(f, GM_ERR_ARG, "spectra datasets exist already");
  std::vector<long long> koff(nf2 + 2, 0);
  for (int jf = 1; jf <= nf2; ++jf) {
    if (nwn[jf - 1] < 0) return fail(f, GM_ERR_ARG, "negative wavenumber count at frequency %d", jf);
    koff[jf + 1] = koff[jf] + nwn[jf - 1];
  }
  const long long nkfmax = koff[nf2 + 1];
  if (nkfmax < 1) return fail(f, GM_ERR_ARG, "no wavenumbers");
  hsize_t d5 = 5;
  hid_t sp = H5Screate_simple(1, &d5, nullptr);
  hid_t a = sp >= 0 ? H5Acreate2(f->fid, "dataSpecificIntegers", H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT) : -1;
  herr_t e = a >= 0 ? H5Awrite(a, H5T_NATIVE_INT, dsi) : -1;
  if (a >= 0) H5Aclose(a);
  if (sp >= 0) H5Sclose(sp);
  if (e < 0) return fail(f, GM_ERR_IO, "cannot write attribute dataSpecificIntegers");
  const hsize_t dims[3] = {(hsize_t)nnod, (hsize_t)nsp, (hsize_t)nkfmax};
  sp = H5Screate_simple(3, dims, nullptr);
  if (sp < 0) return fail(f, GM_ERR_IO, "dataspace of the spectra");
  f->dre = H5Dcreate2(f->fid, "GreenFKSpectraReal", H5T_NATIVE_FLOAT, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  f->dim = H5Dcreate2(f->fid, "GreenFKSpectraImag", H5T_NATIVE_FLOAT, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  H5Sclose(sp);
  if (f->dre < 0 || f->dim < 0) return fail(f, GM_ERR_IO, "cannot create the spectra datasets");
  f->nf1 = nf1; f->nf2 = nf2; f->nsp = nsp; f->nnod = nnod;
  f->koff.swap(koff);
  return GM_OK;
}

// one hyperslab [0:nnod, 0:nsp, k0:k0+cnt] of dataset ds from mem [nnod][nsp][cnt]
static int slab(gm_gfkfile* f, hid_t ds, long long k0, long long cnt, const float* data, const char* name) {
  const hsize_t start[3] = {0, 0, (hsize_t)k0}, count[3] = {(hsize_t)f->nnod, (hsize_t)f->nsp, (hsize_t)cnt};
  hid_t fs = H5Dget_space(ds);
  hid_t ms = H5Screate_simple(3, count, nullptr);
  herr_t e = (fs >= 0 && ms >= 0) ? H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr) : -1;
  if (e >= 0) e = H5Dwrite(ds, H5T_NATIVE_FLOAT, ms, fs, H5P_DEFAULT, data);
  if (ms >= 0) H5Sclose(ms);
  if (fs >= 0) H5Sclose(fs);
  return e < 0 ? fail(f, GM_ERR_IO, "cannot write the slab of %s", name) : GM_OK;
}

extern "C" int gm_gfkfile_write_frequencies(gm_gfkfile* f, int jf1, int jf2, const float* gfre, const float* gfim) {
  quiet();
  if (int rc = need_writer(f)) return rc;
  if (f->dre < 0) return fail(f, GM_ERR_ARG, "spectra datasets must be created first");
  if (jf1 < f->nf1 || jf2 < jf1 || jf2 > f->nf2)
    return fail(f, GM_ERR_ARG, "frequency range %d..%d outside %d..%d", jf1, jf2, f->nf1, f->nf2);
  if (!gfre || !gfim) return fail(f, GM_ERR_ARG, "spectra must not be NULL");
  const long long k0 = f->koff[jf1], cnt = f->koff[jf2 + 1] - k0;
  if (cnt == 0) return GM_OK;
  if (int rc = slab(f, f->dre, k0, cnt, gfre, "GreenFKSpectraReal")) return rc;
  return slab(f, f->dim, k0, cnt, gfim, "GreenFKSpectraImag");
}

// ---------------------------------------------------------------------------
// reading
// ---------------------------------------------------------------------------
static int read_string_obj(gm_gfkfile* f, hid_t obj, bool attr, char* buf, int len, const char* name) {
  hid_t t = attr ? H5Aget_type(obj) : H5Dget_type(obj);
  if (t < 0) return fail(f, GM_ERR_IO, "type of %s", name);
  const size_t n = H5Tget_size(t);
  std::vector<char> v(n + 1, '\0');
  herr_t e = attr ? H5Aread(obj, t, v.data()) : H5Dread(obj, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
  H5Tclose(t);
  if (e < 0) return fail(f, GM_ERR_IO, "cannot read %s", name);
  if ((size_t)len < n + 1) return fail(f, GM_ERR_ARG, "buffer of %d bytes too small for %s (%zu)", len, name, n);
  memcpy(buf, v.data(), n + 1);
  return GM_OK;
}

extern "C" int gm_gfkfile_read_string(gm_gfkfile* f, int which, char* buf, int len) {
  quiet();
  if (!f || !buf || len < 1) return GM_ERR_ARG;
  if (which == GM_GFK_IDENTITY) {
    hid_t a = H5Aopen(f->fid, "identity", H5P_DEFAULT);
    if (a < 0) return fail(f, GM_ERR_IO, "no attribute identity");
    int rc = read_string_obj(f, a, true, buf, len, "identity");
    H5Aclose(a);
    return rc;
  }
  const char* name = which == GM_GFK_EARTH_MODEL_NAME ? "earthModelName"
                     : which == GM_GFK_ATTENUATION_MODE ? "attenuationMode" : nullptr;
  if (!name) return fail(f, GM_ERR_ARG, "unknown string %d", which);
  hid_t ds = H5Dopen2(f->fid, name, H5P_DEFAULT);
  if (ds < 0) return fail(f, GM_ERR_IO, "Path_%s does not exist in HDF file", name);
  int rc = read_string_obj(f, ds, false, buf, len, name);
  H5Dclose(ds);
  return rc;
}

// whole dataset `name` under loc into out (mem type mem), checking its element count
static int read_array(gm_gfkfile* f, hid_t loc, const char* name, hid_t mem, size_t nexpect, void* out,
                      size_t* ngot = nullptr) {
  hid_t ds = H5Dopen2(loc, name, H5P_DEFAULT);
  if (ds < 0) return fail(f, GM_ERR_IO, "Path_%s does not exist in HDF file", name);
  hid_t sp = H5Dget_space(ds);
  const hssize_t n = sp >= 0 ? H5Sget_simple_extent_npoints(sp) : -1;
  if (sp >= 0) H5Sclose(sp);
  if (n < 0) { H5Dclose(ds); return fail(f, GM_ERR_IO, "dataspace of %s", name); }
  if (ngot) *ngot = (size_t)n;
  if (!out) { H5Dclose(ds); return GM_OK; }
  if ((size_t)n != nexpect) {
    H5Dclose(ds);
    return fail(f, GM_ERR_IO, "%s holds %lld values, %zu expected", name, (long long)n, nexpect);
  }
  herr_t e = H5Dread(ds, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, out);
  H5Dclose(ds);
  return e < 0 ? fail(f, GM_ERR_IO, "cannot read %s", name) : GM_OK;
}

extern "C" int gm_gfkfile_read_meta(gm_gfkfile* f, double* reals, int* integers) {
  quiet();
  if (!f) return GM_ERR_ARG;
  if (reals) {
    float r[6];
    if (int rc = read_array(f, f->fid, "reals", H5T_NATIVE_FLOAT, 6, r)) return rc;
    for (int i = 0; i < 6; ++i) reals[i] = r[i];
  }
  if (integers)
    if (int rc = read_array(f, f->fid, "integers", H5T_NATIVE_INT, 8, integers)) return rc;
  return GM_OK;
}

extern "C" int gm_gfkfile_read_nwn(gm_gfkfile* f, int n, int* nwn) {
  quiet();
  if (!f || !nwn) return GM_ERR_ARG;
  size_t got = 0;
  if (int rc = read_array(f, f->fid, "numberOfWavenumbersForFrequency", H5T_NATIVE_INT, 0, nullptr, &got)) return rc;
  if ((size_t)n < got) return fail(f, GM_ERR_ARG, "nwn buffer of %d entries, %zu needed", n, got);
  return read_array(f, f->fid, "numberOfWavenumbersForFrequency", H5T_NATIVE_INT, got, nwn);
}

extern "C" int gm_gfkfile_read_dsi(gm_gfkfile* f, int* dsi) {
  quiet();
  if (!f || !dsi) return GM_ERR_ARG;
  hid_t a = H5Aopen(f->fid, "dataSpecificIntegers", H5P_DEFAULT);
  if (a < 0) return fail(f, GM_ERR_IO, "Path_dataSpecificIntegers does not exist in HDF file");
  hid_t sp = H5Aget_space(a);
  const hssize_t n = sp >= 0 ? H5Sget_simple_extent_npoints(sp) : -1;
  if (sp >= 0) H5Sclose(sp);
  herr_t e = n == 5 ? H5Aread(a, H5T_NATIVE_INT, dsi) : -1;
  H5Aclose(a);
  return e < 0 ? fail(f, GM_ERR_IO, "cannot read dataSpecificIntegers (5 values)") : GM_OK;
}

extern "C" int gm_gfkfile_node_count(gm_gfkfile* f, int* nnod) {
  quiet();
  if (!f || !nnod) return GM_ERR_ARG;
  hid_t g = H5Gopen2(f->fid, "externalNodes", H5P_DEFAULT);
  if (g < 0) return fail(f, GM_ERR_IO, "Path_externalNodes does not exist in HDF file");
  size_t n = 0;
  int rc = read_array(f, g, "radii", H5T_NATIVE_FLOAT, 0, nullptr, &n);
  H5Gclose(g);
  *nnod = (int)n;
  return rc;
}

extern "C" int gm_gfkfile_read_nodes(gm_gfkfile* f, double* node_rearth, double* rnod, int* state, int* layer,
                                     double* ro, double* qinv, double* elcon) {
  quiet();
  if (!f) return GM_ERR_ARG;
  int nnod = 0;
  if (int rc = gm_gfkfile_node_count(f, &nnod)) return rc;
  hid_t g = H5Gopen2(f->fid, "externalNodes", H5P_DEFAULT);
  if (g < 0) return fail(f, GM_ERR_IO, "Path_externalNodes does not exist in HDF file");
  const size_t n = (size_t)nnod;
  std::vector<float> a(7 * n), b(7 * n);
  int rc = GM_OK;
  auto widen = [&](const char* name, size_t cnt, double* out) {
    if (rc || !out) return;
    rc = read_array(f, g, name, H5T_NATIVE_FLOAT, cnt, a.data());
    if (!rc) for (size_t i = 0; i < cnt; ++i) out[i] = a[i];
  };
  widen("rearth", 1, node_rearth);
  widen("radii", n, rnod);
  if (!rc && state) rc = read_array(f, g, "state", H5T_NATIVE_INT, n, state);
  if (!rc && layer) rc = read_array(f, g, "layer", H5T_NATIVE_INT, n, layer);
  widen("density", n, ro);
  widen("inverseQ", 2 * n, qinv);
  if (!rc && elcon) {
    rc = read_array(f, g, "realElcon", H5T_NATIVE_FLOAT, 7 * n, a.data());
    if (!rc) rc = read_array(f, g, "imagElcon", H5T_NATIVE_FLOAT, 7 * n, b.data());
    if (!rc) for (size_t i = 0; i < 7 * n; ++i) { elcon[2 * i] = a[i]; elcon[2 * i + 1] = b[i]; }
  }
  H5Gclose(g);
  return rc;
}

extern "C" int gm_gfkfile_read_source_nodes(gm_gfkfile* f, int* n, double* rsnod) {
  quiet();
  if (!f || !n) return GM_ERR_ARG;
  size_t got = 0;
  if (int rc = read_array(f, f->fid, "sourceNodeRadii", H5T_NATIVE_FLOAT, 0, nullptr, &got)) return rc;
  *n = (int)got;
  if (!rsnod) return GM_OK;
  std::vector<float> v(got);
  if (int rc = read_array(f, f->fid, "sourceNodeRadii", H5T_NATIVE_FLOAT, got, v.data())) return rc;
  for (size_t i = 0; i < got; ++i) rsnod[i] = v[i];
  return GM_OK;
}

extern "C" int gm_gfkfile_read_node(gm_gfkfile* f, int jr, int nsp, long long k0, long long nk, float* gfkr,
                                    float* gfki) {
  quiet();
  if (!f || !gfkr || !gfki) return GM_ERR_ARG;
  const char* names[2] = {"GreenFKSpectraReal", "GreenFKSpectraImag"};
  float* outs[2] = {gfkr, gfki};
  for (int c = 0; c < 2; ++c) {
    hid_t ds = H5Dopen2(f->fid, names[c], H5P_DEFAULT);
    if (ds < 0) return fail(f, GM_ERR_IO, "Path_%s does not exist in HDF file", names[c]);
    hid_t fs = H5Dget_space(ds);
    hsize_t dims[3] = {0, 0, 0};
    const int rank = fs >= 0 ? H5Sget_simple_extent_dims(fs, dims, nullptr) : -1;
    if (rank != 3 || jr < 1 || (hsize_t)jr > dims[0] || nsp < 1 || (hsize_t)nsp > dims[1] || k0 < 0 || nk < 1 ||
        (hsize_t)(k0 + nk) > dims[2]) {
      if (fs >= 0) H5Sclose(fs);
      H5Dclose(ds);
      return fail(f, GM_ERR_ARG, "node %d, %d spectra, k %lld..%lld outside the dataset (%llu, %llu, %llu)", jr, nsp,
                  k0, k0 + nk - 1, (unsigned long long)dims[0], (unsigned long long)dims[1],
                  (unsigned long long)dims[2]);
    }
    const hsize_t start[3] = {(hsize_t)(jr - 1), 0, (hsize_t)k0}, count[3] = {1, (hsize_t)nsp, (hsize_t)nk};
    const hsize_t mdims[2] = {(hsize_t)nsp, (hsize_t)nk};
    hid_t ms = H5Screate_simple(2, mdims, nullptr);
    herr_t e = H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr);
    if (e >= 0) e = H5Dread(ds, H5T_NATIVE_FLOAT, ms, fs, H5P_DEFAULT, outs[c]);
    H5Sclose(ms);
    H5Sclose(fs);
    H5Dclose(ds);
    if (e < 0) return fail(f, GM_ERR_IO, "cannot read node %d of %s", jr, names[c]);
  }
  return GM_OK;
}

extern "C" int gm_gfk_number_of_spectra(const int* integers, const int* dsi, int* nsp) {
  if (!integers || !dsi || !nsp) return GM_ERR_ARG;
  int mask[10] = {0, 1, 0, 1, 0, 0, 0, 1, 0, 0};     // U R V S UP VP W T WP, 1-based
  if (integers[5] == 1) mask[5] = mask[6] = mask[9] = 1;
  if (integers[4] == 1) mask[2] = mask[4] = mask[8] = 1;
  if (dsi[1] == 2) mask[4] = 0;
  if (dsi[2] == 0) mask[7] = mask[8] = mask[9] = 0;
  int s = 0, t = 0;
  for (int i = 1; i <= 6; ++i) s += mask[i];
  for (int i = 7; i <= 9; ++i) t += mask[i];
  *nsp = s * dsi[3] + t * dsi[4];
  return GM_OK;
}
