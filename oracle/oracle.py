"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path
(gemini_unified_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)
L = C.POINTER(C.c_int64)


def build(force: bool = False) -> str:
    srcs = [os.path.join(_HERE, f) for f in ("gemini_oracle.cpp", "synth_oracle.cpp")]
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < max(os.path.getmtime(f) for f in srcs):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", _HERE, "liboracle.so"])
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _lib.orc_create.restype = C.c_void_p
        _lib.orc_create.argtypes = [C.c_int] + [D] * 9 + [C.c_int, I, D, I, C.c_int, C.c_int, C.c_int, C.c_int,
                                                          C.c_double, C.c_double, C.c_double, C.c_double,
                                                          C.c_int, D, I, I]
        _lib.orc_destroy.argtypes = [C.c_void_p]
        _lib.orc_set_frequency.argtypes = [C.c_void_p, C.c_double, C.c_int]
        _lib.orc_get_spline.argtypes = [C.c_void_p, D, D, D, D]
        _lib.orc_eval_elcon.argtypes = [C.c_void_p, C.c_double, C.c_int, D, D, D]
        _lib.orc_layer_index.argtypes = [C.c_void_p, C.c_double, I, I]
        _lib.orc_start_radius.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, D]
        _lib.orc_initial_values.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                            C.c_int, D, D]
        _lib.orc_sphancf.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, D]
        _lib.orc_spheroidal_adjoint.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, D, C.c_double,
                                                C.c_double, C.c_double, D, D, L]
        _lib.orc_zetl.argtypes = [C.c_double, C.c_double, C.c_double, D]
        _lib.orc_toroidal_minors.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                             C.c_double, D, D, D, L]
        _lib.orc_spheroidal_minors.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                               C.c_double, D, D, D, D, D, L]
        _lib.orc_gfk_column.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.c_double, D, D, D, I, L]
        _lib.orc_gfk_degrees.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                         C.c_int, C.c_int, C.c_double, C.c_int, D, I, L]
        _lib.orc_gfk_degrees_stress.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int,
                                                C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, D, I, L]
        _lib.orc_propagate_const.argtypes = [C.c_int, D, C.c_double, C.c_double, C.c_double, C.c_double, D, D, L]
        _lib.orc_num_threads.restype = C.c_int
        _lib.orc_reciprocity_sph.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, D, D]
        _lib.orc_vti.argtypes = [C.c_double] * 9 + [D]
        _lib.orc_minors_degrees.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                            C.c_int, C.c_int, C.c_double, C.c_int, I, L]
        _lib.orc_gfk_wavenumbers.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                             C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, D, I, L]
        _lib.orc_gfk_wavenumbers_stress.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int,
                                                    C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, D,
                                                    I, L]
        _lib.orc_aski_degrees.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                          C.c_int, I, C.c_int, C.c_int, C.c_double, C.c_int, D, I, L]
        _lib.orc_exp10.restype = C.c_double
        _lib.orc_exp10.argtypes = [C.c_double]
        _lib.orc_sysmat_at.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, D]
        _lib.orc_minors_wavenumbers.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                                C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, I, L]
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(D)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(I)


def _c(a):
    """complex128 array -> double* view"""
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.ctypes.data_as(D)


STAT_NAMES = ["propagations", "accepted_steps", "rejected_steps", "mmp_sequences", "matvecs",
              "sysmat_evals", "neville_cells", "flops", "zero_det", "subintervals"]


class Oracle:
    """One model + node set; mirrors the reference module state (spline tables per frequency)."""

    def __init__(self, nem, nodes):
        self.nem, self.nodes = nem, nodes
        self._keep = [np.ascontiguousarray(x, dtype=np.float64) for x in
                      (nem.rk, nem.ro, nem.vpv, nem.vph, nem.vsv, nem.vsh, nem.eta, nem.qkinv, nem.qminv, nem.rb,
                       nodes.rnod)]
        self._ikeep = [np.ascontiguousarray(x, dtype=np.int32) for x in
                       (nem.iktop, nem.iflso, nodes.state, nodes.layer)]
        k, ik = self._keep, self._ikeep
        self.h = lib().orc_create(nem.nk, *(x.ctypes.data_as(D) for x in k[:9]), nem.nlay,
                                  ik[0].ctypes.data_as(I), k[9].ctypes.data_as(D), ik[1].ctypes.data_as(I),
                                  nem.nic, nem.noc, nem.fsflag, nem.equidis, nem.rearth, nem.ric, nem.roc,
                                  nem.fref, nodes.nnod, k[10].ctypes.data_as(D), ik[2].ctypes.data_as(I),
                                  ik[3].ctypes.data_as(I))
        self.stats = np.zeros(10, dtype=np.int64)

    def __del__(self):
        try:
            lib().orc_destroy(self.h)
        except Exception:
            pass

    def set_frequency(self, f, attmode=3):
        rc = lib().orc_set_frequency(self.h, float(f), int(attmode))
        if rc:
            raise RuntimeError(f"oracle spline failed rc={rc}")

    def spline(self):
        nk = self.nem.nk
        ro, ro2 = np.zeros(nk), np.zeros(nk)
        el, el2 = np.zeros((7, nk), np.complex128), np.zeros((7, nk), np.complex128)
        lib().orc_get_spline(self.h, _d(ro) if False else ro.ctypes.data_as(D), ro2.ctypes.data_as(D),
                             _c(el), _c(el2))
        return ro, ro2, el, el2

    def eval_elcon(self, r, nl):
        ro = C.c_double()
        z = np.zeros(7, np.complex128)
        q = np.zeros(2)
        lib().orc_eval_elcon(self.h, float(r), int(nl), C.byref(ro), _c(z), q.ctypes.data_as(D))
        return ro.value, z, q

    def start_radius(self, omre, dll1, tor):
        r = C.c_double()
        lib().orc_start_radius(self.h, float(omre), float(dll1), int(bool(tor)), C.byref(r))
        return r.value

    def initial_values(self, omre, omim, dll1, r, nl, kind):
        y = np.zeros(5, np.complex128)
        h = C.c_double()
        rc = lib().orc_initial_values(self.h, omre, omim, dll1, r, nl, kind, _c(y), C.byref(h))
        return rc, y, h.value

    def toroidal_minors(self, omre, omim, dll1, jsl, jsr, eps):
        n = self.nodes.nnod
        u, d, det = (np.zeros((n, 2), np.complex128), np.zeros((n, 2), np.complex128), np.zeros(n, np.complex128))
        rc = lib().orc_toroidal_minors(self.h, omre, omim, dll1, jsl, jsr, eps, _c(u), _c(d), _c(det),
                                       self.stats.ctypes.data_as(L))
        return rc, u, d, det

    def spheroidal_minors(self, omre, omim, dll1, jsl, jsr, eps):
        n = self.nodes.nnod
        u, d = np.zeros((n, 6), np.complex128), np.zeros((n, 6), np.complex128)
        ulg, dlg, det = np.zeros(n), np.zeros(n), np.zeros(n, np.complex128)
        rc = lib().orc_spheroidal_minors(self.h, omre, omim, dll1, jsl, jsr, eps, _c(u), _c(d),
                                         ulg.ctypes.data_as(D), dlg.ctypes.data_as(D), _c(det),
                                         self.stats.ctypes.data_as(L))
        return rc, u, d, ulg, dlg, det

    def spheroidal_adjoint(self, omre, omim, dll1, ys, ra, re, eps):
        """computeSpheroidalAdjoint from ra to re starting at ys[4]; returns (rc, adj[nnod, 4], yend[4])."""
        n = self.nodes.nnod
        y0 = np.ascontiguousarray(ys, dtype=np.complex128)
        adj, yend = np.zeros((n, 4), np.complex128), np.zeros(4, np.complex128)
        rc = lib().orc_spheroidal_adjoint(self.h, C.c_double(omre), C.c_double(omim), C.c_double(dll1), _c(y0),
                                          C.c_double(ra), C.c_double(re), C.c_double(eps), _c(adj), _c(yend),
                                          self.stats.ctypes.data_as(L))
        return rc, adj, yend

    def nsp(self, jre, istyp):
        """spectra per node: U, V of the spheroidal jumps and W of the toroidal ones; no W for a
        receiver in the ocean (computeGreenFKSpectraForSynthetics.f90:293-313)"""
        njsph, njtor = (4, 2) if istyp == 1 else (2, 1)
        return 2 * njsph + (0 if self.nodes.state[jre - 1] == 1 else njtor)

    def gfk_column(self, omre, omim, dll1, jre, istyp, eps):
        n = self.nodes.nnod
        nsp = self.nsp(jre, istyp)
        gf = np.zeros((n, nsp), np.complex128)
        ujs = np.zeros((n, 4, 4), np.complex128)
        ujt = np.zeros((n, 2, 2), np.complex128)
        flag = C.c_int()
        rc = lib().orc_gfk_column(self.h, omre, omim, dll1, jre, istyp, eps, _c(gf), _c(ujs), _c(ujt),
                                  C.byref(flag), self.stats.ctypes.data_as(L))
        return rc, gf, ujs, ujt, flag.value

    def minors_degrees(self, omre, omim, lmin, nl, jsl, jsr, eps, nthreads=1):
        """Toroidal + spheroidal minors for l = lmin..lmin+nl-1 (results discarded); returns status."""
        status = np.zeros(nl, np.int32)
        lib().orc_minors_degrees(self.h, omre, omim, self.nem.rearth, lmin, nl, jsl, jsr, eps, nthreads,
                                 status.ctypes.data_as(I), self.stats.ctypes.data_as(L))
        return status

    def gfk_wavenumbers(self, omre, omim, iwnmin, nwn, dwn, jre, istyp, eps, nthreads=1):
        """Regional-mode columns for wavenumber indices iwnmin..iwnmin+nwn-1 (k = (iwn-1) dwn); gf[n, isp, i]."""
        n = self.nodes.nnod
        nsp = self.nsp(jre, istyp)
        gf = np.zeros((n, nsp, nwn), np.complex128)
        status = np.zeros(nwn, np.int32)
        rc = lib().orc_gfk_wavenumbers(self.h, C.c_double(omre), C.c_double(omim), C.c_double(self.nem.rearth),
                                       int(iwnmin), int(nwn), C.c_double(dwn), int(jre), int(istyp), C.c_double(eps),
                                       int(nthreads), _c(gf), status.ctypes.data_as(I),
                                       self.stats.ctypes.data_as(L))
        return rc, gf, status

    def minors_wavenumbers(self, omre, omim, iwnmin, nwn, dwn, jsl, jsr, eps, nthreads=1):
        status = np.zeros(nwn, np.int32)
        lib().orc_minors_wavenumbers(self.h, C.c_double(omre), C.c_double(omim), C.c_double(self.nem.rearth),
                                     int(iwnmin), int(nwn), C.c_double(dwn), int(jsl), int(jsr), C.c_double(eps),
                                     int(nthreads), status.ctypes.data_as(I), self.stats.ctypes.data_as(L))
        return status

    def gfk_degrees(self, omre, omim, lmin, nl, jre, istyp, eps, nthreads=1):
        """Columns for l = lmin..lmin+nl-1 at the current frequency; gf[n, isp, il]."""
        n = self.nodes.nnod
        nsp = self.nsp(jre, istyp)
        gf = np.zeros((n, nsp, nl), np.complex128)
        status = np.zeros(nl, np.int32)
        rc = lib().orc_gfk_degrees(self.h, omre, omim, self.nem.rearth, lmin, nl, jre, istyp, eps, nthreads,
                                   _c(gf), status.ctypes.data_as(I), self.stats.ctypes.data_as(L))
        return rc, gf, status

    def gfk_degrees_stress(self, omre, omim, lmin, nl, jre, istyp, eps, nthreads=1):
        """Columns with stresses (STRESSFLAG): gf[n, isp, il] with U, R, V, S (x njsph jumps), W, T (x njtor);
        U, R, V and zero S slots for a receiver in the ocean."""
        n = self.nodes.nnod
        njsph, njtor = (4, 2) if istyp == 1 else (2, 1)
        nsp = 4 * njsph + (0 if self.nodes.state[jre - 1] == 1 else 2 * njtor)
        gf = np.zeros((n, nsp, nl), np.complex128)
        status = np.zeros(nl, np.int32)
        rc = lib().orc_gfk_degrees_stress(self.h, omre, omim, self.nem.rearth, lmin, nl, jre, istyp, eps, nthreads,
                                          _c(gf), status.ctypes.data_as(I), self.stats.ctypes.data_as(L))
        return rc, gf, status

    def gfk_wavenumbers_stress(self, omre, omim, iwnmin, nwn, dwn, jre, istyp, eps, nthreads=1):
        """Regional-mode columns with stresses (STRESSFLAG), k = (iwn-1) dwn; layout of gfk_degrees_stress."""
        n = self.nodes.nnod
        njsph, njtor = (4, 2) if istyp == 1 else (2, 1)
        nsp = 4 * njsph + (0 if self.nodes.state[jre - 1] == 1 else 2 * njtor)
        gf = np.zeros((n, nsp, nwn), np.complex128)
        status = np.zeros(nwn, np.int32)
        rc = lib().orc_gfk_wavenumbers_stress(self.h, C.c_double(omre), C.c_double(omim),
                                              C.c_double(self.nem.rearth), int(iwnmin), int(nwn), C.c_double(dwn),
                                              int(jre), int(istyp), C.c_double(eps), int(nthreads), _c(gf),
                                              status.ctypes.data_as(I), self.stats.ctypes.data_as(L))
        return rc, gf, status

    def aski_degrees(self, omre, omim, lmin, nl, jsnod, jre, istyp, eps, nthreads=1):
        """Multi-source columns for l = lmin..lmin+nl-1 at the current frequency:
        gf[js, n, isp, il] for the len(jsnod) sources and the force at jre (last)."""
        n = self.nodes.nnod
        nspmax = 20 if istyp == 1 else 10
        jsn = np.ascontiguousarray(jsnod, dtype=np.int32)
        gf = np.zeros((len(jsn) + 1, n, nspmax, nl), np.complex128)
        status = np.zeros(nl, np.int32)
        rc = lib().orc_aski_degrees(self.h, omre, omim, self.nem.rearth, lmin, nl, len(jsn), jsn.ctypes.data_as(I),
                                    jre, istyp, eps, nthreads, _c(gf), status.ctypes.data_as(I),
                                    self.stats.ctypes.data_as(L))
        return rc, gf, status

    def sysmat_at(self, sys, omre, omim, dll1, r):
        """System matrix at r: sys 0 solid, 1 toroidal, 2 fluid, 3 solidt (adjoint)."""
        nv = 4 if sys in (0, 3) else 2
        a = np.zeros((nv, nv), np.complex128)
        lib().orc_sysmat_at(self.h, sys, omre, omim, dll1, r, _c(a))
        return a


def propagate_const(A, x1, x2, htry, eps, y0):
    A = np.ascontiguousarray(A, dtype=np.complex128)
    y0 = np.ascontiguousarray(y0, dtype=np.complex128)
    y1 = np.zeros_like(y0)
    st = np.zeros(10, np.int64)
    rc = lib().orc_propagate_const(A.shape[0], _c(A), x1, x2, htry, eps, _c(y0), _c(y1), st.ctypes.data_as(L))
    return rc, y1, st


def sphancf(zx, el, kind):
    out = np.zeros(2)
    rc = lib().orc_sphancf(zx.real, zx.imag, el, kind, out.ctypes.data_as(D))
    return rc, complex(out[0], out[1])


def zetl(zx2, el):
    out = np.zeros(2)
    rc = lib().orc_zetl(zx2.real, zx2.imag, el, out.ctypes.data_as(D))
    return rc, complex(out[0], out[1])


def exp10(x):
    """the oracle's 10^x of the minors renormalisation"""
    return lib().orc_exp10(float(x))


def num_threads():
    return lib().orc_num_threads()


def reciprocity_sph(elp1, re, r, state_node, bgin):
    bgin = np.asfortranarray(bgin, dtype=np.complex128)
    flat = np.ascontiguousarray(bgin.T.reshape(-1))        # column-major order
    out = np.zeros(16, np.complex128)
    lib().orc_reciprocity_sph(elp1, re, r, state_node, flat.ctypes.data_as(D), out.ctypes.data_as(D))
    return out.reshape(4, 4).T                               # back to [comp, jump]


def vti(fratio, ro, alfv, alfh, betv, beth, eta, qkinv, qminv):
    z = np.zeros(7, np.complex128)
    lib().orc_vti(fratio, ro, alfv, alfh, betv, beth, eta, qkinv, qminv, z.ctypes.data_as(D))
    return z


# ---- synthetics half: Legendre functions, degree sums, frequency-time transform ----

def legendre(lmax, theta):
    """dict of plm, dpdt, pdst, d2pdt2, dtpdst, each [3, lmax+1] (m = 0..2; pdst/dtpdst row 0 unused)."""
    out = {k: np.zeros((3, lmax + 1)) for k in ("plm", "dpdt", "pdst", "d2pdt2", "dtpdst")}
    rc = lib().orc_legendre(int(lmax), C.c_double(theta), *[_d(out[k]) for k in ("plm", "dpdt", "pdst", "d2pdt2",
                                                                                   "dtpdst")])
    assert rc == 0
    return out


def strain_spectra(istyp, nf1, nf2, nwn, gr, gi, theta, phi, re, fomt, elcon=None, tapfrac=0.2, nthreads=1,
                   dwn=0.0, rearth=0.0):
    """Strain spectra of one node (spectra with radial derivatives, nsp 20 moment / 10 force):
    dict hdsum [nrcv, nf, nsum], zsp [nrcv, 3, nf2], strain [nrcv, 3, 3, nf2], stress (with elcon).
    dwn > 0: the regional branch, wavenumber integrals at x = theta * rearth (theta in radians)."""
    theta = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
    phi = np.ascontiguousarray(np.broadcast_to(phi, theta.shape), dtype=np.float64)
    nrcv, nf, nsum = theta.size, nf2 - nf1 + 1, 42 if istyp == 1 else 21
    gr = np.ascontiguousarray(gr, dtype=np.float32)
    gi = np.ascontiguousarray(gi, dtype=np.float32)
    nkf = gr.shape[1]
    nw = np.ascontiguousarray(nwn, dtype=np.int32)
    out = dict(hdsum=np.zeros((nrcv, nf, nsum), np.complex128), zsp=np.zeros((nrcv, 3, nf2), np.complex128),
               strain=np.zeros((nrcv, 3, 3, nf2), np.complex128),
               stress=np.zeros((nrcv, 3, 3, nf2), np.complex128) if elcon is not None else None)
    el = None if elcon is None else np.ascontiguousarray(elcon, dtype=np.complex128)
    rc = lib().orc_strain_spectra_wn(int(istyp), int(nf1), int(nf2), nw.ctypes.data_as(I), C.c_long(nkf),
                                     gr.ctypes.data_as(C.POINTER(C.c_float)), gi.ctypes.data_as(C.POINTER(C.c_float)),
                                     nrcv, _d(theta), _d(phi), C.c_double(re), _d(fomt),
                                     None if el is None else _c(el), C.c_double(tapfrac), C.c_double(dwn),
                                     C.c_double(rearth), _c(out["hdsum"]), _c(out["zsp"]), _c(out["strain"]),
                                     None if el is None else _c(out["stress"]), int(nthreads))
    assert rc == 0
    return out


def spherical_strain(u, r, delta):
    """sphericalStrainSingleFrequencyDisplacement: u[4, 3] complex (u, d/dr u, 1/r d/dtheta u,
    1/(r sin theta) d/dphi u) -> strain [3, 3] complex."""
    u = np.ascontiguousarray(u, dtype=np.complex128)
    e = np.zeros((3, 3), np.complex128)
    lib().orc_spherical_strain(_c(u), C.c_double(r), C.c_double(delta), _c(e))
    return e


def taper(lmax, tapfrac=0.2):
    w = np.zeros(lmax + 1)
    lib().orc_taper(int(lmax), C.c_double(tapfrac), _d(w))
    return w


def fft(x, y, is_=1):
    """In-place-style radix-2 FFT of the reference (no normalisation); returns (x, y)."""
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    y = np.ascontiguousarray(y, dtype=np.float64).copy()
    n = int(round(np.log2(x.size)))
    assert 1 << n == x.size
    lib().orc_fft(n, int(is_), _d(x), _d(y))
    return x, y


def _degree_sums(fn, nsum, nf1, nf2, nwn, gr, gi, theta, tapfrac, nthreads):
    nwn = np.ascontiguousarray(nwn, dtype=np.int32)
    gr = np.ascontiguousarray(gr, dtype=np.float32)
    gi = np.ascontiguousarray(gi, dtype=np.float32)
    theta = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
    nf = nf2 - nf1 + 1
    h = np.zeros((theta.size, nf, nsum), np.complex128)
    F = C.POINTER(C.c_float)
    fn(int(nf1), int(nf2), nwn.ctypes.data_as(I), C.c_long(gr.shape[1]), gr.ctypes.data_as(F),
       gi.ctypes.data_as(F), theta.size, _d(theta), C.c_double(tapfrac), h.ctypes.data_as(D), int(nthreads))
    return h


def moment_sums(nf1, nf2, nwn, gr, gi, theta, tapfrac=0.2, nthreads=1):
    """14 moment harmonic degree sums: complex [nrcv, nf, 14].  gr, gi: FP32 [nsp = 10, nkf] of one node."""
    return _degree_sums(lib().orc_moment_sums, 14, nf1, nf2, nwn, gr, gi, theta, tapfrac, nthreads)


def force_sums(nf1, nf2, nwn, gr, gi, theta, tapfrac=0.2, nthreads=1):
    """7 force harmonic degree sums: complex [nrcv, nf, 7].  gr, gi: FP32 [nsp = 5, nkf] of one node
    (U1, U2, V1, V2, W1)."""
    return _degree_sums(lib().orc_force_sums, 7, nf1, nf2, nwn, gr, gi, theta, tapfrac, nthreads)


def besselj(order, x):
    lib().orc_besselj.restype = C.c_double
    return lib().orc_besselj(int(order), C.c_double(x))


def bessely(order, x):
    lib().orc_bessely.restype = C.c_double
    return lib().orc_bessely(int(order), C.c_double(x))


def bessel_sums(nf1, nf2, nwn, gr, gi, dwn, rearth, x, tapfrac=0.2, nthreads=1, istyp=1):
    """Wavenumber integrals: complex [nrcv, nf, 14] (istyp 1, moment) or [nrcv, nf, 7] (istyp 0, force);
    x = epicentral distances in m."""
    nwn = np.ascontiguousarray(nwn, dtype=np.int32)
    gr = np.ascontiguousarray(gr, dtype=np.float32)
    gi = np.ascontiguousarray(gi, dtype=np.float32)
    x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
    h = np.zeros((x.size, nf2 - nf1 + 1, 14 if istyp == 1 else 7), np.complex128)
    F = C.POINTER(C.c_float)
    lib().orc_bessel_sums_src(int(istyp), int(nf1), int(nf2), nwn.ctypes.data_as(I), C.c_long(gr.shape[1]),
                              gr.ctypes.data_as(F), gi.ctypes.data_as(F), C.c_double(dwn), C.c_double(rearth),
                              x.size, _d(x), C.c_double(tapfrac), h.ctypes.data_as(D), int(nthreads))
    return h


def synthetics(nf1, nf2, df, sigma, nwn, gr, gi, theta, phi, mt, nsamp, dt, propdir=None, tred=None, velocity=False,
               nthreads=1, dwn=0.0, rearth=0.0, evfilter=None, stafilter=None):
    """(ts float32 [nrcv, 3, nsamp], zsp complex [nrcv, 3, nf2]) of one source node; dwn > 0 selects the
    regional (wavenumber integral) branch with distances theta * rearth.  The source type follows the
    number of spectra: 10 (moment, mt = Mrr, Mtt, Mff, Mrt, Mrf, Mtf) or 5 (force, mt = Fr, Ft, Ff).
    evfilter: complex [nf2], stafilter: complex [nrcv, 3, nf2] (None: unity)."""
    nwn = np.ascontiguousarray(nwn, dtype=np.int32)
    gr = np.ascontiguousarray(gr, dtype=np.float32)
    gi = np.ascontiguousarray(gi, dtype=np.float32)
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    mt = np.ascontiguousarray(mt, dtype=np.float64)
    nr = theta.size
    ts = np.zeros((nr, 3, nsamp), np.float32)
    zsp = np.zeros((nr, 3, nf2), np.complex128)
    F = C.POINTER(C.c_float)
    # one value per receiver (a scalar applies to all): the library reads nr values
    pd = None if propdir is None else np.ascontiguousarray(np.broadcast_to(np.asarray(propdir, np.float64), (nr,)))
    tr = None if tred is None else np.ascontiguousarray(np.broadcast_to(np.asarray(tred, np.float64), (nr,)))
    ev = None if evfilter is None else np.ascontiguousarray(evfilter, dtype=np.complex128)
    st = None if stafilter is None else np.ascontiguousarray(stafilter, dtype=np.complex128)
    istyp = {10: 1, 5: 0}[gr.shape[0]]
    assert mt.size == (6 if istyp == 1 else 3)
    rc = lib().orc_synthetics_src(istyp, int(nf1), int(nf2), C.c_double(df), C.c_double(sigma), nwn.ctypes.data_as(I),
                              C.c_long(gr.shape[1]), gr.ctypes.data_as(F), gi.ctypes.data_as(F), C.c_double(dwn),
                              C.c_double(rearth), nr, _d(theta),
                              _d(phi), None if pd is None else _d(pd), None if tr is None else _d(tr),
                              None if ev is None else ev.ctypes.data_as(D), None if st is None else st.ctypes.data_as(D),
                              _d(mt),
                              int(bool(velocity)), int(nsamp), C.c_double(dt), ts.ctypes.data_as(F),
                              zsp.ctypes.data_as(D), int(nthreads))
    assert rc == 0, rc
    return ts, zsp


def butterworth_bandpass(filspecs, nf1, nf2, df, sigma):
    """computeEventfilter('BUTTERWORTH_BANDPASS', filspecs, ...): complex [nf2], zero below nf1."""
    spec = np.ascontiguousarray(filspecs, dtype=np.float64)
    h = np.zeros(nf2, np.complex128)
    lib().orc_butterworth_bandpass(_d(spec), int(nf1), int(nf2), C.c_double(df), C.c_double(sigma),
                                   h.ctypes.data_as(D))
    return h


def freq_time(nf1, nf2, df, sigma, sp, nsamp, dt):
    """transformFrequencyTime of one spectrum sp (complex [nf2]); returns (status, ts float32 [nsamp])."""
    sp = np.ascontiguousarray(sp, dtype=np.complex128)
    ts = np.zeros(nsamp, np.float32)
    rc = lib().orc_freq_time(int(nf1), int(nf2), C.c_double(df), C.c_double(sigma), sp.ctypes.data_as(D), int(nsamp),
                             C.c_double(dt), ts.ctypes.data_as(C.POINTER(C.c_float)))
    return rc, ts


def travel_times(table, rs, re, delta):
    """getTravelTimes (f90/travelTimes.f90:373-388) for receivers at radii re[r]
    and epicentral distances delta[r] (radians) from a source at radius rs.
    table: dict with rnod [nnod], rtp [ntp], slowness [ntp], delta and ttime [nnod, ntp]
    (FP32, the reference's raytable_travel_times) and istele.  Returns (tt, slowness, status)."""
    f32 = lambda a: np.asfortranarray(a, dtype=np.float32)
    rnod, rtp, slow = f32(table["rnod"]), f32(table["rtp"]), f32(table["slowness"])
    dl, tm = f32(table["delta"]), f32(table["ttime"])
    re = np.ascontiguousarray(re, dtype=np.float64)
    dn = np.ascontiguousarray(delta, dtype=np.float64)
    n = re.size
    tt = np.zeros(n)
    p = np.zeros(n)
    st = np.zeros(n, np.int32)
    F = C.POINTER(C.c_float)
    lib().orc_travel_times(int(rnod.size), int(rtp.size), int(bool(table.get("istele", False))), rnod.ctypes.data_as(F),
                           rtp.ctypes.data_as(F), slow.ctypes.data_as(F), dl.ctypes.data_as(F), tm.ctypes.data_as(F),
                           C.c_double(rs), int(n), _d(re), _d(dn), _d(tt), _d(p), st.ctypes.data_as(C.POINTER(C.c_int)))
    return tt, p, st


def homogeneous_ray_table(radius, v, nnod, rmin_frac=0.3, istele=False):
    """Ray table of a homogeneous sphere (straight rays) on nnod radii from rmin_frac * radius to
    radius, with turning points at all radii but the top one: a ray turning at rtp reaches radius
    r >= rtp at angular distance acos(rtp / r) after sqrt(r^2 - rtp^2) / v seconds; its angular
    slowness is rtp / v.  Entries with r < rtp are zero."""
    rnod = np.linspace(rmin_frac * radius, radius, nnod)
    rtp = rnod[:-1]
    R, P = np.meshgrid(rnod, rtp, indexing="ij")
    reach = R >= P
    dl = np.where(reach, np.arccos(np.minimum(P / R, 1.0)), 0.0)
    tm = np.where(reach, np.sqrt(np.maximum(R * R - P * P, 0.0)) / v, 0.0)
    return {"rnod": rnod, "rtp": rtp, "slowness": rtp / v, "delta": dl, "ttime": tm, "istele": istele}
