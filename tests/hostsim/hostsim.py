"""CPU build of the engine's kernels (libgmsim.so, built by the Makefile beside
this file): the library's own device code, compiled as host C++ and run one
block at a time with one thread per block.  Test infrastructure only.

HostSim mirrors GreenFKEngine's minors / gfk / step-log / learned-control
calls, with outputs in the same [node][component][k] layouts.
"""
from __future__ import annotations

import ctypes as C
import fcntl
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libgmsim.so")
D = C.POINTER(C.c_double)
F = C.POINTER(C.c_float)
I = C.POINTER(C.c_int)
U = C.POINTER(C.c_ulonglong)
VARIANTS = ("exact", "fast", "exact_fused", "fast_fused")
_lib = None


def lib():
    global _lib
    if _lib is None:
        # one build at a time when several test processes start together
        with open(os.path.join(HERE, ".build.lock"), "w") as lk:
            fcntl.flock(lk, fcntl.LOCK_EX)
            subprocess.run(["make", "-s", "-C", HERE, "-j4"], check=True)
            fcntl.flock(lk, fcntl.LOCK_UN)
        L = C.CDLL(LIB)
        L.sim_model_new.restype = C.c_void_p
        L.sim_model_new.argtypes = [C.c_int] + [D] * 9 + [C.c_int, I, D, I, C.c_int, C.c_int, C.c_int, C.c_int,
                                                          C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                                          D, I, I]
        L.sim_model_free.argtypes = [C.c_void_p]
        L.sim_proposer_new.restype = C.c_void_p
        L.sim_proposer_new.argtypes = [F, C.c_float, C.c_int, C.c_int]
        L.sim_proposer_free.argtypes = [C.c_void_p]
        L.sim_set_adjoint_pair.argtypes = [C.c_int]
        L.sim_set_stress.argtypes = [C.c_int]
        L.sim_set_sources.argtypes = [C.c_int, I]
        for v in VARIANTS:
            fn = getattr(L, "sim_run_" + v)
            fn.argtypes = [C.c_void_p, C.c_int, D, I, I, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                           C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, F, C.c_void_p, C.c_int, C.c_int,
                           C.c_int] + [D] * 9 + [I, U, F, I, F]
            getattr(L, "sim_run_" + v + "_node_props").argtypes = [C.c_void_p, C.c_double, C.c_int, D, D, D]
        _lib = L
    return _lib


class HostSim:
    def __init__(self, nem, nodes, variant="exact"):
        assert variant in VARIANTS
        L = lib()
        self.nem, self.nodes, self.variant = nem, nodes, variant
        self._k = [np.ascontiguousarray(x, dtype=np.float64) for x in
                   (nem.rk, nem.ro, nem.vpv, nem.vph, nem.vsv, nem.vsh, nem.eta, nem.qkinv, nem.qminv, nem.rb,
                    nodes.rnod)]
        self._i = [np.ascontiguousarray(x, dtype=np.int32) for x in (nem.iktop, nem.iflso, nodes.state, nodes.layer)]
        k, i = self._k, self._i
        self.m = L.sim_model_new(nem.nk, *[x.ctypes.data_as(D) for x in k[:9]], nem.nlay, i[0].ctypes.data_as(I),
                                 k[9].ctypes.data_as(D), i[1].ctypes.data_as(I), nem.nic, nem.noc, nem.fsflag,
                                 nem.equidis, nem.rearth, nem.ric, nem.roc, nem.fref, nodes.nnod,
                                 k[10].ctypes.data_as(D), i[2].ctypes.data_as(I), i[3].ctypes.data_as(I))
        assert self.m
        self.prop, self.mlp = None, None
        self._sources = None
        self.policy_net, self.order_target = False, 6

    def __del__(self):
        try:
            lib().sim_model_free(self.m)
            if self.prop:
                lib().sim_proposer_free(self.prop)
        except Exception:
            pass

    def node_properties(self, f, attmode=0):
        """k_spline + k_node_props on the CPU: ro[nnod], elcon[nnod, 7] complex, qinv[nnod, 2]."""
        n = self.nodes.nnod
        ro, el, q = np.zeros(n), np.zeros((n, 7), np.complex128), np.zeros((n, 2))
        rc = getattr(lib(), "sim_run_" + self.variant + "_node_props")(
            self.m, float(f), int(attmode), ro.ctypes.data_as(D), el.ctypes.data_as(D), q.ctypes.data_as(D))
        assert rc == 0
        return ro, el, q

    def set_proposer(self, params, norm, safety=0.8, kcheck=False, every=1, mode="step"):
        """mode "step": learned control during the integration (runs with learned=True); "policy": the
        network's per-instance target try count and first step for plain runs."""
        self.policy_net = mode == "policy"
        self.mlp = np.ascontiguousarray(params, dtype=np.float32)
        nv = np.ascontiguousarray(norm, dtype=np.float32)
        if self.prop:
            lib().sim_proposer_free(self.prop)
        self.prop = lib().sim_proposer_new(nv.ctypes.data_as(F), float(safety), int(bool(kcheck)), int(every))

    def _run(self, freqs, lmin, nl, sigma, eps, jsl, jsr, jre, istyp, full, mode, dwn=0.0, attmode=3, logmax=0):
        freqs = np.ascontiguousarray(np.atleast_1d(freqs), dtype=np.float64)
        nf = freqs.size
        lmin = np.ascontiguousarray(np.broadcast_to(lmin, (nf,)), dtype=np.int32)
        nl = np.ascontiguousarray(np.broadcast_to(nl, (nf,)), dtype=np.int32)
        n, nn = int(nl.sum()), self.nodes.nnod
        nmask = 4 if (full and self._stress) else 2
        njs, njt = (4, 2) if istyp == 1 else (2, 1)
        nsp = nmask * njs + (0 if (full and self.nodes.state[jre - 1] == 1) else nmask // 2 * njt)
        nsrc = 1
        if self._sources is not None:
            nsp, nsrc = (20 if istyp == 1 else 10), len(self._sources) + 1
        o = dict(u_tor=np.zeros((nn, 2, n), complex), d_tor=np.zeros((nn, 2, n), complex),
                 det_tor=np.zeros((nn, n), complex), u_sph=np.zeros((nn, 6, n), complex),
                 d_sph=np.zeros((nn, 6, n), complex), ulg=np.zeros((nn, n)), dlg=np.zeros((nn, n)),
                 det_sph=np.zeros((nn, n), complex),
                 gf=(np.zeros((nn, nsp, n), complex) if nsrc == 1 else np.zeros((nsrc, nn, nsp, n), complex))
                 if full else None,
                 status=np.zeros(n, np.int32), hist=np.zeros(48, np.uint64))
        log = np.zeros((max(logmax, 1), 20), np.float32)
        nlog = C.c_int(0)
        o["policy"] = np.zeros((n, 2), np.float32)
        dp = lambda a: None if a is None else a.ctypes.data_as(D)
        lib().sim_set_stress(1 if (full and self._stress and self._sources is None) else 0)
        fn = getattr(lib(), "sim_run_" + self.variant)
        rc = fn(self.m, nf, freqs.ctypes.data_as(D), lmin.ctypes.data_as(I), nl.ctypes.data_as(I), float(dwn),
                float(sigma), float(eps), int(attmode), int(jsl), int(jsr), int(jre), int(istyp), int(full), int(mode),
                None if self.mlp is None else self.mlp.ctypes.data_as(F), self.prop, int(logmax),
                int(self.policy_net), int(self.order_target), dp(o["u_tor"]), dp(o["d_tor"]), dp(o["det_tor"]), dp(o["u_sph"]), dp(o["d_sph"]), dp(o["ulg"]),
                dp(o["dlg"]), dp(o["det_sph"]), dp(o["gf"]), o["status"].ctypes.data_as(I),
                o["hist"].ctypes.data_as(U), log.ctypes.data_as(F), C.byref(nlog), o["policy"].ctypes.data_as(F))
        assert rc == 0, rc
        o["stats"] = derive_stats(o["hist"])
        if mode == 1:
            o["log"], o["nlog"] = log[:min(nlog.value, logmax)], nlog.value
        return o

    def minors(self, freqs, lmin, nl, sigma, eps, jsl, jsr, dwn=0.0, learned=False):
        return self._run(freqs, lmin, nl, sigma, eps, jsl, jsr, jsr, 1, False, 2 if learned else 0, dwn)

    _stress = False

    def gfk(self, freqs, lmin, nl, sigma, eps, jre, istyp=1, dwn=0.0, learned=False, stress=False):
        self._stress = stress
        try:
            return self._run(freqs, lmin, nl, sigma, eps, jre, jre, jre, istyp, True, 2 if learned else 0, dwn)
        finally:
            self._stress = False

    def aski(self, freqs, lmin, nl, sigma, eps, jsnod, jre, istyp=1):
        """Multi-source columns (gm_aski_batch): gf[js, node, isp, instance]."""
        js = np.ascontiguousarray(jsnod, dtype=np.int32)
        self._sources = js
        lib().sim_set_sources(len(js), js.ctypes.data_as(I))
        try:
            return self._run(freqs, lmin, nl, sigma, eps, int(js.min()), int(js.max()), jre, istyp, True, 0)
        finally:
            self._sources = None
            lib().sim_set_sources(-1, js.ctypes.data_as(I))

    def steplog(self, freqs, lmin, nl, sigma, eps, jsl, jsr, maxrec, dwn=0.0):
        return self._run(freqs, lmin, nl, sigma, eps, jsl, jsr, jsr, 1, False, 1, dwn, logmax=maxrec)


# the library's derivation of the work counters from the per-system histogram (gm_capi.cpp derive_stats)
_N, _NNZ, _ESYS = (5, 2, 2, 2, 4), (19, 4, 3, 4, 14), (219, 148, 156, 196, 206)


def derive_stats(hist):
    h = np.asarray(hist, dtype=np.int64)
    seq, nx = [0, 2, 4, 6, 8, 12, 16, 24], [0, 3, 2, 4, 4, 4, 8, 8]
    MV, SE, NC = [0] * 8, [0] * 8, [0] * 8
    for k in range(1, 8):
        MV[k] = MV[k - 1] + seq[k] + 1
        SE[k] = SE[k - 1] + nx[k]
        NC[k] = NC[k - 1] + (k * (k - 1) // 2 if k >= 3 else 0)
    st = np.zeros(10, np.int64)
    for s in range(5):
        hs = h[s * 8:(s + 1) * 8]
        for k in range(3, 9):
            cnt = int(hs[k - 3] if k <= 7 else hs[5])
            kk = min(k, 7)
            if not cnt:
                continue
            st[3] += cnt * kk
            st[4] += cnt * MV[kk]
            st[5] += cnt * SE[kk]
            st[6] += cnt * NC[kk]
            st[7] += cnt * (SE[kk] * _ESYS[s] + MV[kk] * 8 * _NNZ[s] + 4 * _N[s] * MV[kk] + 8 * _N[s] * NC[kk])
            if k <= 7:
                st[1] += cnt
            else:
                st[2] += cnt
        st[0] += int(hs[6])
        st[9] += int(hs[7])
    st[8] += int(h[40])
    return st
