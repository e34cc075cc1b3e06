"""Synthetics half of the Green-FK path on the GPU: Legendre functions and
harmonic degree sums (global branch) or Bessel wavenumber integrals (regional
branch), displacement spectra and the frequency-to-time transform
(f90/computeSyntheticSeismograms.f90:311-388).

The Green FK spectra of one source node come in the layout the reference's
synthetics stage reads from the GFK file (readDataNodeGreenFKSpectra,
f90/greenFKSpectra.f90:171-212): FP32 real and imaginary parts [nsp][nkf], k
running over the frequencies nf1..nf2 with the degrees of each frequency
fastest.  That is also one node's slice of GreenFKEngine.gfk()'s FP32 output.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import D, I, VP, check

MOMENT_SUMS = 14
FORCE_SUMS = 7
# spectra per node in the GFK file for dsvmask (U, V, W): numberGreenFKSpectra
# (f90/greenFKSpectra.f90:216-220) for a single force (istyp 0) and a moment tensor (istyp 1)
NSP = {0: 5, 1: 10}


# spectrum layouts of a GFK file and their flags (GM_FLAG_STRESS_SPECTRA, GM_FLAG_DERIV_SPECTRA,
# GM_FLAG_OCEAN_SPECTRA): "ocean" and "ocean_stress" are the files of a receiver in the ocean
# (U, V or U, R, V; no W)
LAYOUT_FLAGS = {"uvw": 0, "stress": 2, "deriv": 4, "ocean": 8, "ocean_stress": 10}
# spectra per node of a moment source in each layout (half of it for a single force)
LAYOUT_NSP = {"uvw": 10, "stress": 20, "deriv": 20, "ocean": 8, "ocean_stress": 12}


class SyntheticsEngine:
    def __init__(self, device: int = 0):
        self.device = device
        h = VP()
        rc = _lib.lib().gm_ctx_create(int(device), C.byref(h))
        if rc != 0:
            raise _lib.GeminiFKError(rc, _lib.last_error(None))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().gm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def timing(self):
        """ms of the last call: Legendre basis, spectra preparation, contraction, combination, transform, total."""
        ms = np.zeros(6)
        check(self.h, _lib.lib().gm_last_synth_timing(self.h, ms.ctypes.data_as(D)))
        return ms

    def travel_times(self, table, rs, re, delta, raise_on_error=False):
        """getTravelTimes (f90/travelTimes.f90:373-388) on the GPU for a source at radius rs and
        receivers at radii re[r] and epicentral distances delta[r] (radians).  table: dict with the
        FP32 arrays of raytable_travel_times (rnod [nnod], rtp [ntp], slowness [ntp], delta and
        ttime [nnod, ntp]) and istele.  Returns (tt, slowness, status); with raise_on_error a failing
        receiver raises GeminiFKError with the reference's message instead."""
        f32 = lambda a: np.asfortranarray(a, dtype=np.float32)
        rnod, rtp, slow = f32(table["rnod"]), f32(table["rtp"]), f32(table["slowness"])
        dl, tm = f32(table["delta"]), f32(table["ttime"])
        if dl.shape != (rnod.size, rtp.size) or tm.shape != dl.shape or slow.size != rtp.size:
            raise ValueError("delta and ttime must be [nnod, ntp] and slowness [ntp]")
        re = np.ascontiguousarray(np.broadcast_to(re, np.shape(delta)), dtype=np.float64).ravel()
        dn = np.ascontiguousarray(delta, dtype=np.float64).ravel()
        tt, p, st = np.zeros(dn.size), np.zeros(dn.size), np.zeros(dn.size, np.int32)
        Fp = C.POINTER(C.c_float)
        check(self.h, _lib.lib().gm_travel_times(
            self.h, rnod.size, rtp.size, int(bool(table.get("istele", False))), rnod.ctypes.data_as(Fp),
            rtp.ctypes.data_as(Fp), slow.ctypes.data_as(Fp), dl.ctypes.data_as(Fp), tm.ctypes.data_as(Fp), float(rs),
            dn.size, re.ctypes.data_as(D), dn.ctypes.data_as(D), tt.ctypes.data_as(D), p.ctypes.data_as(D),
            None if raise_on_error else st.ctypes.data_as(I)))
        return tt, p, st

    def auto_timeshift(self, table, rs, re, delta, tsbuf, twinlen, tlen, dt, nsamp):
        """Reduction time and true sample count per station for autoTimeshift
        (f90/computeSyntheticSeismograms.f90:344-352): tred = tt - tsbuf, the window ends at
        tt + twinlen but at most tlen after tred, nstrue = min(int((tend - tred) / dt) + 1, nsamp).
        A failing station raises GeminiFKError with the reference's message, as there."""
        tt, _, _ = self.travel_times(table, rs, re, delta, raise_on_error=True)
        tred = tt - tsbuf
        tend = np.minimum(tt + twinlen, tred + tlen)
        nstrue = np.minimum(((tend - tred) / dt).astype(np.int64) + 1, nsamp)
        return tred, nstrue

    def legendre(self, lmax, theta):
        """computeLegendreFunctions (f90/legendreFunctions.f90:48) at each theta: dict of [nrcv, 3, lmax+1]."""
        theta = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
        out = {k: np.zeros((theta.size, 3, lmax + 1)) for k in ("plm", "dpdt", "pdst", "d2pdt2", "dtpdst")}
        check(self.h, _lib.lib().gm_legendre_batch(self.h, theta.size, theta.ctypes.data_as(D), int(lmax),
                                                   *[out[k].ctypes.data_as(D) for k in ("plm", "dpdt", "pdst",
                                                                                        "d2pdt2", "dtpdst")]))
        return out

    @staticmethod
    def _spectra(gr, gi):
        gr = np.ascontiguousarray(gr, dtype=np.float32)
        gi = np.ascontiguousarray(gi, dtype=np.float32)
        if gr.shape != gi.shape or gr.ndim != 2:
            raise ValueError("gr, gi must be [nsp, nkf] arrays of one shape")
        return gr, gi

    def _sums(self, fn, nsum, nf1, nf2, nwn, gr, gi, theta, layout="uvw"):
        nwn = np.ascontiguousarray(nwn, dtype=np.int32)
        gr, gi = self._spectra(gr, gi)
        theta = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
        out = np.zeros((nsum, theta.size, nf2 - nf1 + 1), np.complex128)
        check(self.h, fn(self.h, int(nf1), int(nf2), nwn.ctypes.data_as(I), gr.shape[0], gr.shape[1], gr.ctypes.data,
                         gi.ctypes.data, theta.size, theta.ctypes.data_as(D), LAYOUT_FLAGS[layout], out.ctypes.data))
        return out

    def _wn_sums(self, fn, nsum, nf1, nf2, nwn, gr, gi, dwn, rearth, x):
        nwn = np.ascontiguousarray(nwn, dtype=np.int32)
        gr, gi = self._spectra(gr, gi)
        x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
        out = np.zeros((nsum, x.size, nf2 - nf1 + 1), np.complex128)
        check(self.h, fn(self.h, int(nf1), int(nf2), nwn.ctypes.data_as(I), gr.shape[0], gr.shape[1], gr.ctypes.data,
                         gi.ctypes.data, float(dwn), float(rearth), x.size, x.ctypes.data_as(D), 0, out.ctypes.data))
        return out

    def moment_sums(self, nf1, nf2, nwn, gr, gi, theta, layout="uvw"):
        """14 moment harmonic degree sums, complex [14, nrcv, nf2-nf1+1] (momentHarmonicDegreeSums);
        gr, gi: [LAYOUT_NSP[layout], nkf], e.g. [10, nkf] (layout "uvw") or [20, nkf] ("stress", "deriv")."""
        return self._sums(_lib.lib().gm_moment_sums, MOMENT_SUMS, nf1, nf2, nwn, gr, gi, theta, layout)

    def force_sums(self, nf1, nf2, nwn, gr, gi, theta):
        """7 single-force harmonic degree sums, complex [7, nrcv, nf2-nf1+1] (forceHarmonicDegreeSums);
        gr, gi: [5, nkf] (U1, U2, V1, V2, W1)."""
        return self._sums(_lib.lib().gm_force_sums, FORCE_SUMS, nf1, nf2, nwn, gr, gi, theta)

    def strain_spectra(self, nf1, nf2, nwn, gr, gi, theta, phi, re, fomt, istyp=1, elcon=None, want_sums=False,
                       dwn=None, rearth=None):
        """Strain (and, with elcon = 7 complex moduli at the receiver, stress) spectra from spectra with radial
        derivatives (gr, gi: [20, nkf] moment or [10, nkf] force, U V dU dV W dW): dict strain and stress
        [nrcv, 3, 3, nf2], zsp [nrcv, 3, nf2] (components r, theta, phi), hdsum [42 or 21, nrcv, nf]
        (sphericalHarmonicsDisplacementSpectra with strainsp, displacementSpectra.f90:49-106).
        Regional branch (dwn given): theta = epicentral distances in metres, wavenumber integrals with
        k = (j-1) dwn and rearth (besselDisplacementSpectra, :125-184)."""
        nwn = np.ascontiguousarray(nwn, dtype=np.int32)
        gr, gi = self._spectra(gr, gi)
        theta = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
        phi = np.ascontiguousarray(np.broadcast_to(phi, theta.shape), dtype=np.float64)
        fm = np.ascontiguousarray(fomt, dtype=np.float64)
        if fm.size != (6 if istyp == 1 else 3):
            raise ValueError("fomt needs 6 moment-tensor (istyp 1) or 3 force (istyp 0) components")
        R, nf = theta.size, nf2 - nf1 + 1
        out = dict(zsp=np.zeros((R, 3, nf2), np.complex128), strain=np.zeros((R, 3, 3, nf2), np.complex128),
                   stress=np.zeros((R, 3, 3, nf2), np.complex128) if elcon is not None else None,
                   hdsum=np.zeros((42 if istyp == 1 else 21, R, nf), np.complex128) if want_sums else None)
        el = None if elcon is None else np.ascontiguousarray(elcon, dtype=np.complex128)
        vp = lambda a: None if a is None else a.ctypes.data
        eld = None if el is None else el.ctypes.data_as(D)
        if dwn:
            rc = _lib.lib().gm_strain_spectra_regional(self.h, int(nf1), int(nf2), nwn.ctypes.data_as(I), gr.shape[0],
                                                       gr.shape[1], gr.ctypes.data, gi.ctypes.data, R,
                                                       theta.ctypes.data_as(D), phi.ctypes.data_as(D), float(re),
                                                       float(dwn), float(rearth if rearth else re), int(istyp),
                                                       fm.ctypes.data_as(D), eld, 0, vp(out["hdsum"]), vp(out["zsp"]),
                                                       vp(out["strain"]), vp(out["stress"]))
        else:
            rc = _lib.lib().gm_strain_spectra(self.h, int(nf1), int(nf2), nwn.ctypes.data_as(I), gr.shape[0],
                                              gr.shape[1], gr.ctypes.data, gi.ctypes.data, R,
                                              theta.ctypes.data_as(D), phi.ctypes.data_as(D), float(re),
                                              int(istyp), fm.ctypes.data_as(D), eld, 0, vp(out["hdsum"]),
                                              vp(out["zsp"]), vp(out["strain"]), vp(out["stress"]))
        check(self.h, rc)
        return out

    def bessel_sums(self, nf1, nf2, nwn, gr, gi, dwn, rearth, x):
        """14 moment wavenumber integrals, complex [14, nrcv, nf2-nf1+1] (momentWavenumberIntegrals);
        x = epicentral distances in metres, wavenumbers k = (j-1) dwn."""
        return self._wn_sums(_lib.lib().gm_bessel_sums, MOMENT_SUMS, nf1, nf2, nwn, gr, gi, dwn, rearth, x)

    def force_bessel_sums(self, nf1, nf2, nwn, gr, gi, dwn, rearth, x):
        """7 single-force wavenumber integrals, complex [7, nrcv, nf2-nf1+1] (forceWavenumberIntegrals)."""
        return self._wn_sums(_lib.lib().gm_force_bessel_sums, FORCE_SUMS, nf1, nf2, nwn, gr, gi, dwn, rearth, x)

    def synthetics(self, nf1, nf2, df, sigma, nwn, gr, gi, theta, phi, mt, nsamp, dt, propdir=None, tred=None,
                   velocity=False, want_spectra=False, device_ptrs=None, dwn=None, rearth=None, evfilter=None,
                   stafilter=None, istyp=None, layout="uvw"):
        """Seismograms [nrcv, 3, nsamp] FP32 (and displacement spectra [nrcv, 3, nf2] if asked).

        Source type istyp (the GFK file's): 1, moment tensor, nsp = 10, mt = (Mrr, Mtt, Mff, Mrt, Mrf, Mtf);
        0, single force, nsp = 5, mt = (Fr, Ft, Ff) (getForceMomentSeismicEvent's fomt).  None: the type
        those spectrum counts imply.
        evfilter: complex [nf2] event filter (e.g. event_filter()), stafilter: complex [nrcv, 3, nf2]
        station filters of the r, theta, phi components; None: unity.
        Global branch: theta = epicentral distances in radians.  Regional branch (dwn given):
        theta = epicentral distances in metres, wavenumber integrals with k = (j-1) dwn.
        device_ptrs = (gr_ptr, gi_ptr, nsp, nkf, ts_ptr): spectra and output already in HBM
        (integers, e.g. torch tensors' data_ptr()); gr, gi are then ignored and ts is written in place.
        layout: the file's spectra per node: "uvw" (U, V, W), "stress" (U, R, V, S, W, T: STRESSFLAG
        files), "deriv" (U, V, dU/dr, dV/dr, W, dW/dr: the multi-source driver's files), "ocean"
        (U, V: a receiver in the ocean) or "ocean_stress" (U, R, V).
        """
        lflag = LAYOUT_FLAGS[layout]
        nwn = np.ascontiguousarray(nwn, dtype=np.int32)
        theta = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
        phi = np.ascontiguousarray(np.atleast_1d(phi), dtype=np.float64)
        mt = np.ascontiguousarray(mt, dtype=np.float64)
        nsp_in = device_ptrs[2] if device_ptrs is not None else np.shape(gr)[0]
        if istyp is None:
            istyp = 0 if int(nsp_in) == LAYOUT_NSP[layout] // 2 else 1
        if phi.size != theta.size or mt.size != (6 if istyp == 1 else 3):
            raise ValueError("phi must match theta and mt must hold 6 (moment) or 3 (force) entries")
        ev = None if evfilter is None else np.ascontiguousarray(evfilter, dtype=np.complex128)
        st = None if stafilter is None else np.ascontiguousarray(stafilter, dtype=np.complex128)
        if (ev is not None and ev.shape != (nf2,)) or (st is not None and st.shape != (theta.size, 3, nf2)):
            raise ValueError("evfilter must be [nf2] and stafilter [nrcv, 3, nf2]")
        # one value per receiver: a scalar (e.g. the event's reduction time of the reference's non-auto
        # branch, getTimeshiftSeismicEvent) applies to every receiver; any other size is an error, since
        # the library reads nrcv values
        pd = None if propdir is None else per_receiver(propdir, theta.size, "propdir")
        tr = None if tred is None else per_receiver(tred, theta.size, "tred")
        zsp = np.zeros((theta.size, 3, nf2), np.complex128) if want_spectra else None
        if device_ptrs is None:
            gr, gi = self._spectra(gr, gi)
            nsp, nkf, pgr, pgi, flags = gr.shape[0], gr.shape[1], gr.ctypes.data, gi.ctypes.data, lflag
            ts = np.zeros((theta.size, 3, nsamp), np.float32)
            pts = ts.ctypes.data
        else:
            if want_spectra:
                raise ValueError("want_spectra needs host outputs")
            pgr, pgi, nsp, nkf, pts = device_ptrs
            ts, flags = None, 1 | lflag
        tail = (theta.size, theta.ctypes.data_as(D), phi.ctypes.data_as(D), None if pd is None else pd.ctypes.data_as(D),
                None if tr is None else tr.ctypes.data_as(D), None if ev is None else ev.ctypes.data_as(D),
                None if st is None else st.ctypes.data_as(D), int(istyp), mt.ctypes.data_as(D), int(bool(velocity)),
                int(nsamp),
                float(dt), flags, pts, None if zsp is None else zsp.ctypes.data)
        L = _lib.lib()
        if dwn:
            rc = L.gm_synthetics_regional(self.h, int(nf1), int(nf2), float(df), float(sigma), nwn.ctypes.data_as(I),
                                          int(nsp), int(nkf), pgr, pgi, float(dwn), float(rearth), *tail)
        else:
            rc = L.gm_synthetics(self.h, int(nf1), int(nf2), float(df), float(sigma), nwn.ctypes.data_as(I), int(nsp),
                                 int(nkf), pgr, pgi, *tail)
        check(self.h, rc)
        return (ts, zsp) if want_spectra else ts


GM_FILTER_BUTTERWORTH_BANDPASS = 1


def per_receiver(a, nrcv, name):
    """a as nrcv contiguous doubles: a scalar is broadcast, an array must hold exactly nrcv values."""
    a = np.asarray(a, dtype=np.float64)
    if a.size == 1:
        return np.full(nrcv, float(a.reshape(-1)[0]))
    if a.size != nrcv:
        raise ValueError(f"{name} holds {a.size} values for {nrcv} receivers")
    return np.ascontiguousarray(a.reshape(-1))


def event_filter(filspecs, nf1, nf2, df, sigma, filtype=GM_FILTER_BUTTERWORTH_BANDPASS):
    """computeEventfilter (f90/eventFilter.f90:40-83): complex [nf2], zero below nf1.
    filspecs = (high-pass order, corner Hz, low-pass order, corner Hz)."""
    spec = np.ascontiguousarray(filspecs, dtype=np.float64)
    if spec.size != 4:
        raise ValueError("filspecs holds 4 numbers")
    h = np.zeros(nf2, np.complex128)
    check(None, _lib.lib().gm_event_filter(int(filtype), spec.ctypes.data_as(D), int(nf1), int(nf2), float(df),
                                           float(sigma), h.ctypes.data_as(D)))
    return h
