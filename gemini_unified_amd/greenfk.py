"""Host-side mirror of the reference's Green-FK driver interface.

``GreenFKEngine`` owns one C-ABI context (one HIP device, one stream) and
replaces, for batches of (omega, l):

* ``computeToroidalMinors`` / ``computeSpheroidalMinors``
  (f90/toroidalMinors.f90:71, f90/spheroidalMinors.f90:72) -> :meth:`minors`,
* the degree-loop body of ``computeGreenFKSpectraForSynthetics``
  (f90/computeGreenFKSpectraForSynthetics.f90:410-494) -> :meth:`gfk`,
* the whole frequency loop writing ``gfre/gfim(nkf, nsp, nnod)``
  -> :meth:`compute_green_fk_spectra_for_synthetics`.

Errors follow the reference's level-2 convention: a failing call raises
``GeminiFKError`` with the code and message (the reference prints and calls
MPI_Abort).  All numerics run in libgemini_fk.so on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import D, F, I, LL, check
from .earthmodel import ATTMODES, ExternalRadialNodes, FrequencyGrid, NodeEarthmodel

STAT_NAMES = ["propagations", "accepted_steps", "rejected_steps", "mmp_sequences", "matvecs",
              "sysmat_evals", "neville_cells", "flops", "zero_det", "subintervals"]
VARIANTS = {"exact": 0, "fast": 1}


def _dp(a):
    return None if a is None else a.ctypes.data_as(D)


class GreenFKEngine:
    def __init__(self, nem: NodeEarthmodel, nodes: ExternalRadialNodes, device: int = 0, variant: str = "fast"):
        L = _lib.lib()
        h = C.c_void_p()
        rc = L.gm_ctx_create(int(device), C.byref(h))
        if rc != 0:
            raise _lib.GeminiFKError(rc, _lib.last_error(None))
        self.h = h
        self.nem, self.nodes = nem, nodes
        self.set_variant(variant)
        k = [np.ascontiguousarray(x, dtype=np.float64) for x in
             (nem.rk, nem.ro, nem.vpv, nem.vph, nem.vsv, nem.vsh, nem.eta, nem.qkinv, nem.qminv)]
        iktop = np.ascontiguousarray(nem.iktop, dtype=np.int32)
        rb = np.ascontiguousarray(nem.rb, dtype=np.float64)
        iflso = np.ascontiguousarray(nem.iflso, dtype=np.int32)
        check(h, L.gm_model_set(h, nem.nk, *[x.ctypes.data_as(D) for x in k], nem.nlay, iktop.ctypes.data_as(I),
                                rb.ctypes.data_as(D), iflso.ctypes.data_as(I), nem.nic, nem.noc, nem.fsflag,
                                nem.equidis, nem.rearth, nem.ric, nem.roc, nem.fref))
        rn = np.ascontiguousarray(nodes.rnod, dtype=np.float64)
        st = np.ascontiguousarray(nodes.state, dtype=np.int32)
        ly = np.ascontiguousarray(nodes.layer, dtype=np.int32)
        check(h, L.gm_nodes_set(h, nodes.nnod, rn.ctypes.data_as(D), st.ctypes.data_as(I), ly.ctypes.data_as(I)))
        self.stats = np.zeros(10, dtype=np.int64)

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().gm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_variant(self, variant: str):
        check(self.h, _lib.lib().gm_set_option(self.h, 1, VARIANTS[variant]))
        self.variant = variant

    def configure(self, persistent=None, unit=None, waves=None, propose_every=None, sph_split=None,
                  order_target=None, graph=None, adjoint_pair=None):
        """Kernel schedule options (all result-preserving): persistent substep-scheduled kernels or
        one lane per job, work-unit size, waves per SIMD of the persistent kernels; how often the
        learned proposer is consulted (every N accepted steps, plus sub-interval starts); and the
        target try count of the step rule without a network (6: the reference's; changes results
        at the tolerance level); graph: the kernel stages of plain-control batches as one hipGraph,
        captured once and replayed while shapes and buffers are unchanged; adjoint_pair: two adjoint
        integrations per leg give the four unit jumps (half the adjoint work, eps-level differences)."""
        L = _lib.lib()
        if graph is not None:
            check(self.h, L.gm_set_option(self.h, 8, int(bool(graph))))
        if adjoint_pair is not None:
            check(self.h, L.gm_set_option(self.h, 9, int(bool(adjoint_pair))))
        if order_target is not None:
            check(self.h, L.gm_set_option(self.h, 7, int(order_target)))
        if propose_every is not None:
            check(self.h, L.gm_set_option(self.h, 5, int(propose_every)))
        if sph_split is not None:
            check(self.h, L.gm_set_option(self.h, 6, int(sph_split)))
        if persistent is not None:
            check(self.h, L.gm_set_option(self.h, 2, int(bool(persistent))))
        if unit is not None:
            check(self.h, L.gm_set_option(self.h, 3, int(unit)))
        if waves is not None:
            check(self.h, L.gm_set_option(self.h, 4, int(waves)))

    def timing(self):
        ms = np.zeros(6)
        _lib.lib().gm_last_timing(self.h, ms.ctypes.data_as(D))
        return ms

    def system_stats(self, sys: int):
        st = np.zeros(10, np.int64)
        check(self.h, _lib.lib().gm_last_system_stats(self.h, int(sys), st.ctypes.data_as(LL)))
        return st

    def run_minors_resident(self, freqs, lmin, nl, sigma, eps, jsl, jsr, attmode="ATTENUATION_AND_DISPERSION",
                            dwn=None):
        """Minors of all (f, l) pairs left resident in the engine's HBM buffers (no copy-out)."""
        freqs, lmin, nl, nk = self._plan(freqs, lmin, nl)
        st = np.zeros(10, np.int64)
        am = ATTMODES[attmode] if isinstance(attmode, str) else int(attmode)
        if dwn:
            rc = _lib.lib().gm_minors_batch_wn(self.h, freqs.size, freqs.ctypes.data_as(D), lmin.ctypes.data_as(I),
                                               nl.ctypes.data_as(I), float(dwn), float(sigma), float(eps), am,
                                               int(jsl), int(jsr), 0, *([None] * 8), None, st.ctypes.data_as(LL))
        else:
            rc = _lib.lib().gm_minors_batch(self.h, freqs.size, freqs.ctypes.data_as(D), lmin.ctypes.data_as(I),
                                            nl.ctypes.data_as(I), float(sigma), float(eps), am, int(jsl), int(jsr), 0,
                                            *([None] * 8), None, st.ctypes.data_as(LL))
        check(self.h, rc)
        return st

    def run_gfk_resident(self, freqs, lmin, nl, sigma, eps, jre, istyp=1, attmode="ATTENUATION_AND_DISPERSION",
                         dwn=None):
        """Green-FK columns of all (f, l) pairs left resident in HBM (no copy-out)."""
        freqs, lmin, nl, nk = self._plan(freqs, lmin, nl)
        st = np.zeros(10, np.int64)
        am = ATTMODES[attmode] if isinstance(attmode, str) else int(attmode)
        if dwn:
            rc = _lib.lib().gm_gfk_batch_wn(self.h, freqs.size, freqs.ctypes.data_as(D), lmin.ctypes.data_as(I),
                                            nl.ctypes.data_as(I), float(dwn), float(sigma), float(eps), am, int(jre),
                                            int(istyp), 0, None, None, None, None, st.ctypes.data_as(LL))
        else:
            rc = _lib.lib().gm_gfk_batch(self.h, freqs.size, freqs.ctypes.data_as(D), lmin.ctypes.data_as(I),
                                         nl.ctypes.data_as(I), float(sigma), float(eps), am, int(jre), int(istyp), 0,
                                         None, None, None, None, st.ctypes.data_as(LL))
        check(self.h, rc)
        return st

    def set_proposer(self, params=None, norm=None, safety=0.9, kcheck=False, enable=True, mode="step"):
        """Install the learned step proposer (or remove it with enable=False).  mode "step": the network
        proposes during the integration (sequential step engine); "policy": it chooses, once per
        (omega, l), the target try count and first step, and the kernels keep the fused sweep."""
        if not enable:
            check(self.h, _lib.lib().gm_set_proposer(self.h, None, None, 1.0, 0, 0))
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        nv = np.ascontiguousarray(norm, dtype=np.float32)
        check(self.h, _lib.lib().gm_set_proposer(self.h, p.ctypes.data_as(F), nv.ctypes.data_as(F), float(safety),
                                                 int(bool(kcheck)), {"step": 1, "policy": 2}[mode]))

    def collect_steplog(self, freqs, lmin, nl, sigma, eps, jsl, jsr, maxrec=2_000_000,
                        attmode="ATTENUATION_AND_DISPERSION", dwn=None):
        """Step records of the plain spheroidal-minors integration, float32 [n, 20] (layout in include/gemini_fk.h)."""
        freqs, lmin, nl, nk = self._plan(freqs, lmin, nl)
        rec = np.zeros((maxrec, 20), np.float32)
        nrec = C.c_int()
        st = np.zeros(10, np.int64)
        am = ATTMODES[attmode] if isinstance(attmode, str) else int(attmode)
        check(self.h, _lib.lib().gm_collect_steplog(self.h, freqs.size, freqs.ctypes.data_as(D),
                                                    lmin.ctypes.data_as(I), nl.ctypes.data_as(I),
                                                    float(dwn or 0.0), float(sigma),
                                                    float(eps), am, int(jsl), int(jsr), int(maxrec),
                                                    rec.ctypes.data_as(F), C.byref(nrec), st.ctypes.data_as(LL)))
        return rec[:min(nrec.value, maxrec)], nrec.value, st

    @staticmethod
    def _plan(freqs, lmin, nl):
        freqs = np.ascontiguousarray(np.atleast_1d(freqs), dtype=np.float64)
        nf = freqs.size
        lmin = np.ascontiguousarray(np.broadcast_to(lmin, (nf,)), dtype=np.int32)
        nl = np.ascontiguousarray(np.broadcast_to(nl, (nf,)), dtype=np.int32)
        return freqs, lmin, nl, int(nl.sum())

    def minors(self, freqs, lmin, nl, sigma, eps, jsl, jsr, attmode="ATTENUATION_AND_DISPERSION", raise_on_error=True,
               dwn=None):
        """Toroidal + spheroidal minors of all (f, l) pairs; arrays indexed [node, comp, k].
        With dwn, lmin/nl are 1-based wavenumber indices (regional mode, k = (iwn-1) dwn)."""
        freqs, lmin, nl, nk = self._plan(freqs, lmin, nl)
        nn = self.nodes.nnod
        out = dict(u_tor=np.zeros((nn, 2, nk), np.complex128), d_tor=np.zeros((nn, 2, nk), np.complex128),
                   det_tor=np.zeros((nn, nk), np.complex128), u_sph=np.zeros((nn, 6, nk), np.complex128),
                   d_sph=np.zeros((nn, 6, nk), np.complex128), ulg=np.zeros((nn, nk)), dlg=np.zeros((nn, nk)),
                   det_sph=np.zeros((nn, nk), np.complex128), status=np.zeros(nk, np.int32))
        st = np.zeros(10, np.int64)
        am = ATTMODES[attmode] if isinstance(attmode, str) else int(attmode)
        outs = [out[k].ctypes.data_as(D) for k in ("u_tor", "d_tor", "det_tor", "u_sph", "d_sph", "ulg", "dlg",
                                                    "det_sph")]
        if dwn:
            rc = _lib.lib().gm_minors_batch_wn(self.h, freqs.size, freqs.ctypes.data_as(D), lmin.ctypes.data_as(I),
                                               nl.ctypes.data_as(I), float(dwn), float(sigma), float(eps), am,
                                               int(jsl), int(jsr), 0, *outs, out["status"].ctypes.data_as(I),
                                               st.ctypes.data_as(LL))
        else:
            rc = _lib.lib().gm_minors_batch(self.h, freqs.size, freqs.ctypes.data_as(D), lmin.ctypes.data_as(I),
                                            nl.ctypes.data_as(I), float(sigma), float(eps), am, int(jsl), int(jsr), 0,
                                            *outs, out["status"].ctypes.data_as(I), st.ctypes.data_as(LL))
        self.stats += st
        out["stats"] = st
        if raise_on_error:
            check(self.h, rc)
        return out

    def nsp(self, jre, istyp=1, stress=False):
        """spectra per node of the columns for a receiver at node jre: U, V (x 4 / 2 jumps) and W
        (x 2 / 1), 10 or 5; 8 or 4 for a receiver in the ocean, which has no W
        (computeGreenFKSpectraForSynthetics.f90:293-313).  With stresses (STRESSFLAG, :209-212)
        U, R, V, S and W, T: 20 or 10; 16 or 8 in the ocean (U, R, V and zero S)."""
        njsph, njtor = (4, 2) if istyp == 1 else (2, 1)
        nmask = 4 if stress else 2
        return nmask * njsph + (0 if self.nodes.state[int(jre) - 1] == 1 else nmask // 2 * njtor)

    def gfk(self, freqs, lmin, nl, sigma, eps, jre, istyp=1, attmode="ATTENUATION_AND_DISPERSION",
            want_fp32=True, raise_on_error=True, dwn=None, want_gf=True, stress=False):
        """Green-FK columns; gf[node, isp, k] complex128 (and FP32 re/im as written to the file),
        with nsp = self.nsp(jre, istyp, stress) spectra per node; stress: the columns with
        stresses (U, R, V, S, W, T) of STRESSFLAG."""
        freqs, lmin, nl, nk = self._plan(freqs, lmin, nl)
        nn = self.nodes.nnod
        if not 1 <= int(jre) <= nn:
            raise _lib.GeminiFKError(3, f"receiver node {jre} outside 1..{nn}")
        nsp = self.nsp(jre, istyp, stress)
        fl = _lib.GM_FLAG_STRESS_SPECTRA if stress else 0
        gf = np.zeros((nn, nsp, nk), np.complex128) if want_gf else None
        gfre = np.zeros((nn, nsp, nk), np.float32) if want_fp32 else None
        gfim = np.zeros((nn, nsp, nk), np.float32) if want_fp32 else None
        status = np.zeros(nk, np.int32)
        st = np.zeros(10, np.int64)
        am = ATTMODES[attmode] if isinstance(attmode, str) else int(attmode)
        fre = None if gfre is None else gfre.ctypes.data_as(F)
        fim = None if gfim is None else gfim.ctypes.data_as(F)
        if dwn:
            rc = _lib.lib().gm_gfk_batch_wn(self.h, freqs.size, freqs.ctypes.data_as(D), lmin.ctypes.data_as(I),
                                            nl.ctypes.data_as(I), float(dwn), float(sigma), float(eps), am, int(jre),
                                            int(istyp), fl, _dp(gf), fre, fim, status.ctypes.data_as(I),
                                            st.ctypes.data_as(LL))
        else:
            rc = _lib.lib().gm_gfk_batch(self.h, freqs.size, freqs.ctypes.data_as(D), lmin.ctypes.data_as(I),
                                         nl.ctypes.data_as(I), float(sigma), float(eps), am, int(jre), int(istyp), fl,
                                         _dp(gf), fre, fim, status.ctypes.data_as(I),
                                         st.ctypes.data_as(LL))
        self.stats += st
        if raise_on_error:
            check(self.h, rc)
        return dict(gf=gf, gfre=gfre, gfim=gfim, status=status, stats=st)

    def ujgf(self, freqs, lmin, nl, sigma, eps, jre, attmode="ATTENUATION_AND_DISPERSION", raise_on_error=True,
             dwn=None):
        """Unit-jump Green functions of a jump at node jre (gm_ujgf_batch, computeUnitJumpGreenFunctions):
        ujs[node, jump, comp, k] (4 x 4) and ujt[node, jump, comp, k] (2 x 2) complex128."""
        freqs, lmin, nl, nk = self._plan(freqs, lmin, nl)
        nn = self.nodes.nnod
        ujs = np.zeros((nn, 4, 4, nk), np.complex128)
        ujt = np.zeros((nn, 2, 2, nk), np.complex128)
        status = np.zeros(nk, np.int32)
        st = np.zeros(10, np.int64)
        am = ATTMODES[attmode] if isinstance(attmode, str) else int(attmode)
        if dwn:
            rc = _lib.lib().gm_ujgf_batch_wn(self.h, freqs.size, freqs.ctypes.data_as(D), lmin.ctypes.data_as(I),
                                             nl.ctypes.data_as(I), float(dwn), float(sigma), float(eps), am, int(jre),
                                             0, _dp(ujs), _dp(ujt), status.ctypes.data_as(I), st.ctypes.data_as(LL))
        else:
            rc = _lib.lib().gm_ujgf_batch(self.h, freqs.size, freqs.ctypes.data_as(D), lmin.ctypes.data_as(I),
                                          nl.ctypes.data_as(I), float(sigma), float(eps), am, int(jre), 0, _dp(ujs),
                                          _dp(ujt), status.ctypes.data_as(I), st.ctypes.data_as(LL))
        self.stats += st
        if raise_on_error:
            check(self.h, rc)
        return dict(ujs=ujs, ujt=ujt, status=status, stats=st)

    def aski(self, freqs, lmin, nl, sigma, eps, jsnod, jre, istyp=1, attmode="ATTENUATION_AND_DISPERSION",
             want_fp32=True, raise_on_error=True, want_gf=True):
        """Multi-source columns of computeGreenFKSpectraForASKI (gm_aski_batch): for the sources at the
        1-based nodes jsnod (MOMENT istyp = 1 or FORCE 0) and a FORCE at the receiver node jre (last),
        gf[js, node, isp, k] with isp = U(j) V(j) dU/dr(j) dV/dr(j) W(j) dW/dr(j), 20 (moment) or 10
        (force) spectra per node, zero beyond a source's own count."""
        freqs, lmin, nl, nk = self._plan(freqs, lmin, nl)
        nn = self.nodes.nnod
        js = np.ascontiguousarray(jsnod, dtype=np.int32)
        nsrc, nsp = js.size + 1, (20 if istyp == 1 else 10)
        gf = np.zeros((nsrc, nn, nsp, nk), np.complex128) if want_gf else None
        gfre = np.zeros((nsrc, nn, nsp, nk), np.float32) if want_fp32 else None
        gfim = np.zeros((nsrc, nn, nsp, nk), np.float32) if want_fp32 else None
        status = np.zeros(nk, np.int32)
        st = np.zeros(10, np.int64)
        am = ATTMODES[attmode] if isinstance(attmode, str) else int(attmode)
        rc = _lib.lib().gm_aski_batch(self.h, freqs.size, freqs.ctypes.data_as(D), lmin.ctypes.data_as(I),
                                      nl.ctypes.data_as(I), float(sigma), float(eps), am, int(js.size),
                                      js.ctypes.data_as(I), int(jre), int(istyp), 0, _dp(gf),
                                      None if gfre is None else gfre.ctypes.data_as(F),
                                      None if gfim is None else gfim.ctypes.data_as(F), status.ctypes.data_as(I),
                                      st.ctypes.data_as(LL))
        self.stats += st
        if raise_on_error:
            check(self.h, rc)
        return dict(gf=gf, gfre=gfre, gfim=gfim, status=status, stats=st)

    def node_properties(self, f=None, attmode="ELASTIC"):
        """elasticConstantsExternalRadialNodes (externalRadialNodes.f90:422-433): density ro[nnod],
        complex moduli elcon[nnod, 7] (A, C, F, L, N, kappa, mu) and inverse Q qinv[nnod, 2] at the
        nodes, for frequency f (default: the model's reference frequency) and attenuation mode."""
        n = self.nodes.nnod
        ro, el, q = np.zeros(n), np.zeros((n, 7), np.complex128), np.zeros((n, 2))
        f = self.nem.fref if f is None else f
        rc = _lib.lib().gm_node_properties(self.h, float(f), ATTMODES[attmode] if isinstance(attmode, str) else
                                           int(attmode), ro.ctypes.data_as(D), el.ctypes.data_as(D),
                                           q.ctypes.data_as(D))
        check(self.h, rc)
        return ro, el, q

    def compute_green_fk_spectra_for_synthetics(self, grid: FrequencyGrid, jre: int, eps: float, istyp: int = 1,
                                                attmode="ATTENUATION_AND_DISPERSION", freq_block: int = 64):
        """The frequency loop of computeGreenFKSpectraForSynthetics: returns gfre, gfim with the
        file layout (nkf, nsp, nnod) in Fortran order, i.e. C arrays [nnod][nsp][nkf]."""
        nn = self.nodes.nnod
        nsp = self.nsp(jre, istyp)
        gfre = np.zeros((nn, nsp, grid.nkf), np.float32)
        gfim = np.zeros((nn, nsp, grid.nkf), np.float32)
        freqs = grid.freqs
        off = 0
        for b in range(0, freqs.size, freq_block):
            fb = freqs[b:b + freq_block]
            nwn = grid.nwn[b:b + freq_block]
            r = self.gfk(fb, 0, nwn, grid.sigma, eps, jre, istyp, attmode)
            n = int(nwn.sum())
            gfre[:, :, off:off + n] = r["gfre"]
            gfim[:, :, off:off + n] = r["gfim"]
            off += n
        return gfre, gfim
