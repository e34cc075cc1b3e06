"""GFK file (HDF5) writer and reader: the file computeGreenFKSpectraForSynthetics
writes and computeSyntheticSeismograms reads, over libgemini_gfkfile.so.

* :func:`write_green_fk_spectra_for_synthetics` — the file output of
  computeGreenFKSpectraForSynthetics (f90/computeGreenFKSpectraForSynthetics.f90:236-345,
  500-519): meta data, external nodes with their properties at the reference
  frequency, then the spectra of the GPU batches, one hyperslab per batch of
  consecutive frequencies.
* :class:`GreenFKSpectra` — the green_fk_spectra object
  (f90/greenFKSpectra.f90:38-63) with readMetaGreenFKSpectra,
  readDataNodeGreenFKSpectra and getNodeIndexFromRadiusGreenFKSpectra.
* :class:`GfkFileWriter` — the individual writes, for other layouts (the ASKI
  variant's .meta file with sourceNodeRadii and one data file per source node).

Errors raise :class:`GfkFileError` with the library's code and message.
"""
from __future__ import annotations

import ctypes as C
import os
import time
from dataclasses import dataclass, field

import numpy as np

from . import build as _build
from .earthmodel import ExternalRadialNodes, FrequencyGrid

D = C.POINTER(C.c_double)
F = C.POINTER(C.c_float)
I = C.POINTER(C.c_int)
VP = C.c_void_p

IDENTITY, EARTH_MODEL_NAME, ATTENUATION_MODE = 0, 1, 2
# components of the spectra (greenFKSpectra.f90:64)
COMPONENT_NAMES = ("U", "R", "V", "S", "UP", "VP", "W", "T", "WP")

# every symbol declared in include/gemini_gfkfile.h
EXPORTS = ["gm_gfkfile_create", "gm_gfkfile_open", "gm_gfkfile_close", "gm_gfkfile_last_error",
           "gm_gfkfile_write_identity", "gm_gfkfile_write_meta", "gm_gfkfile_write_nodes",
           "gm_gfkfile_write_source_nodes", "gm_gfkfile_create_spectra", "gm_gfkfile_write_frequencies",
           "gm_gfkfile_read_string", "gm_gfkfile_read_meta", "gm_gfkfile_read_nwn", "gm_gfkfile_read_dsi",
           "gm_gfkfile_node_count", "gm_gfkfile_read_nodes", "gm_gfkfile_read_source_nodes", "gm_gfkfile_read_node",
           "gm_gfk_number_of_spectra"]

_lib = None


class GfkFileError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gemini_gfkfile error {code}: {msg}")
        self.code = code


def lib():
    """libgemini_gfkfile.so (built by __graft_entry__.build() where the HDF5 C library exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_build.FILE_LIB):
            raise RuntimeError(f"{_build.FILE_LIB} is not built (needs the HDF5 C library under {_build.HDF5_DIR})")
        L = C.CDLL(_build.FILE_LIB)
        L.gm_gfkfile_create.argtypes = [C.c_char_p, C.POINTER(VP)]
        L.gm_gfkfile_open.argtypes = [C.c_char_p, C.POINTER(VP)]
        L.gm_gfkfile_close.argtypes = [VP]
        L.gm_gfkfile_last_error.argtypes = [VP, C.c_char_p, C.c_int]
        L.gm_gfkfile_write_identity.argtypes = [VP, C.c_char_p]
        L.gm_gfkfile_write_meta.argtypes = [VP, C.c_char_p, C.c_char_p, D, I, I]
        L.gm_gfkfile_write_nodes.argtypes = [VP, C.c_double, C.c_int, D, I, I, D, D, D]
        L.gm_gfkfile_write_source_nodes.argtypes = [VP, C.c_int, D]
        L.gm_gfkfile_create_spectra.argtypes = [VP, I, C.c_int, C.c_int, C.c_int, C.c_int, I]
        L.gm_gfkfile_write_frequencies.argtypes = [VP, C.c_int, C.c_int, F, F]
        L.gm_gfkfile_read_string.argtypes = [VP, C.c_int, C.c_char_p, C.c_int]
        L.gm_gfkfile_read_meta.argtypes = [VP, D, I]
        L.gm_gfkfile_read_nwn.argtypes = [VP, C.c_int, I]
        L.gm_gfkfile_read_dsi.argtypes = [VP, I]
        L.gm_gfkfile_node_count.argtypes = [VP, I]
        L.gm_gfkfile_read_nodes.argtypes = [VP, D, D, I, I, D, D, D]
        L.gm_gfkfile_read_source_nodes.argtypes = [VP, I, D]
        L.gm_gfkfile_read_node.argtypes = [VP, C.c_int, C.c_int, C.c_longlong, C.c_longlong, F, F]
        L.gm_gfk_number_of_spectra.argtypes = [I, I, I]
        _lib = L
    return _lib


def _check(h, rc):
    if rc:
        buf = C.create_string_buffer(512)
        lib().gm_gfkfile_last_error(h, buf, 512)
        raise GfkFileError(rc, buf.value.decode(errors="replace"))


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def fdate() -> str:
    """The identity string of a run: the date in fdate's (ctime's) format."""
    return time.ctime()


def number_of_spectra(integers, dsi) -> int:
    """numberGreenFKSpectra (greenFKSpectra.f90:210-215) from integers[8] and dataSpecificIntegers[5]."""
    n = C.c_int()
    _check(None, lib().gm_gfk_number_of_spectra(_i(integers).ctypes.data_as(I), _i(dsi).ctypes.data_as(I),
                                                C.byref(n)))
    return n.value


def writer_number_of_spectra(dsvstep: int, ncsph: int, njsph: int, njtor: int) -> int:
    """nsp of the writer (computeGreenFKSpectraForSynthetics.f90:207-212, 289-312): spheroidal mask
    (U, V) or all four with stresses (dsvstep 1), toroidal (W) or (W, T), none at a receiver in the ocean."""
    sph = 4 if dsvstep == 1 else 2
    tor = 0 if ncsph == 2 else (2 if dsvstep == 1 else 1)
    return sph * njsph + tor * njtor


class GfkFileWriter:
    """One GFK file open for writing (createFileParallelAccessHDFWrapper and the writes after it)."""

    def __init__(self, path: str):
        h = VP()
        _check(None, lib().gm_gfkfile_create(os.fsencode(path), C.byref(h)))
        self.h = h
        self.path = path
        self.nwn = None

    def close(self):
        if self.h:
            h, self.h = self.h, None
            _check(None, lib().gm_gfkfile_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def write_identity(self, identity: str):
        _check(self.h, lib().gm_gfkfile_write_identity(self.h, identity.encode()))

    def write_meta(self, earth_model_name, attenuation_mode, reals, integers, nwn):
        """earthModelName, attenuationMode, reals (rearth, re, sigma, df, dwn, fref), nwn[nf2] and
        integers (nf1, nf2, nkfmax, nwnmax, dsvstep, derivflag, numtasks, global)."""
        r, ii, nw = _d(reals), _i(integers), _i(nwn)
        _check(self.h, lib().gm_gfkfile_write_meta(self.h, earth_model_name.encode(), attenuation_mode.encode(),
                                                   r.ctypes.data_as(D), ii.ctypes.data_as(I), nw.ctypes.data_as(I)))

    def write_nodes(self, nodes: ExternalRadialNodes, ro, qinv, elcon):
        """externalNodes group: radii, state, layer and the node properties ro[nnod], qinv[nnod, 2],
        elcon[nnod, 7] complex."""
        rn, st, la = _d(nodes.rnod), _i(nodes.state), _i(nodes.layer)
        ro, q = _d(ro), _d(qinv)
        el = np.ascontiguousarray(elcon, dtype=np.complex128)
        _check(self.h, lib().gm_gfkfile_write_nodes(self.h, float(nodes.rearth), nodes.nnod, rn.ctypes.data_as(D),
                                                    st.ctypes.data_as(I), la.ctypes.data_as(I), ro.ctypes.data_as(D),
                                                    q.ctypes.data_as(D), el.ctypes.data_as(D)))

    def write_source_nodes(self, rsnod):
        r = _d(rsnod)
        _check(self.h, lib().gm_gfkfile_write_source_nodes(self.h, r.size, r.ctypes.data_as(D)))

    def create_spectra(self, dsi, nsp, nnod, nf1, nf2, nwn):
        """dataSpecificIntegers (istyp, ncsph, nctor, njsph, njtor) and the (nkfmax, nsp, nnod) datasets."""
        self.nf1, self.nf2 = int(nf1), int(nf2)
        self.nwn = _i(nwn)
        self.nsp, self.nnod = int(nsp), int(nnod)
        _check(self.h, lib().gm_gfkfile_create_spectra(self.h, _i(dsi).ctypes.data_as(I), self.nsp, self.nnod,
                                                       self.nf1, self.nf2, self.nwn.ctypes.data_as(I)))

    def write_frequencies(self, jf1, jf2, gfre, gfim):
        """Frequencies jf1..jf2 (1-based): gfre/gfim [nnod, nsp, sum(nwn(jf1:jf2))] float32."""
        if self.nwn is None:                # no datasets yet: the library reports it
            z = np.zeros(1, np.float32)
            _check(self.h, lib().gm_gfkfile_write_frequencies(self.h, int(jf1), int(jf2), z.ctypes.data_as(F),
                                                              z.ctypes.data_as(F)))
        cnt = int(self.nwn[jf1 - 1:jf2].sum())
        shape = (self.nnod, self.nsp, cnt)
        a = np.ascontiguousarray(gfre, dtype=np.float32)
        b = np.ascontiguousarray(gfim, dtype=np.float32)
        if a.shape != shape or b.shape != shape:
            raise ValueError(f"spectra of frequencies {jf1}..{jf2} must have shape {shape}, got {a.shape}, {b.shape}")
        _check(self.h, lib().gm_gfkfile_write_frequencies(self.h, int(jf1), int(jf2), a.ctypes.data_as(F),
                                                          b.ctypes.data_as(F)))


@dataclass
class GreenFKSpectra:
    """green_fk_spectra (greenFKSpectra.f90:38-63) of one open GFK file."""
    path: str
    nf1: int = 0
    nf2: int = 0
    nkfmax: int = 0
    nwnmax: int = 0
    istyp: int = 0
    ncsph: int = 0
    nctor: int = 0
    njsph: int = 0
    njtor: int = 0
    numtasks: int = 0
    global_: int = 0
    dsvmask: np.ndarray = field(default_factory=lambda: np.zeros(9, np.int32))
    nwn: np.ndarray = field(default=None, repr=False)
    df: float = 0.0
    dwn: float = 0.0
    sigma: float = 0.0
    rearth: float = 0.0
    rse: float = 0.0
    fref: float = 0.0
    rsnod: np.ndarray = field(default=None, repr=False)
    earth_model_name: str = ""
    attnmode: str = ""
    identity: str = ""
    exnod: ExternalRadialNodes = field(default=None, repr=False)
    node_ro: np.ndarray = field(default=None, repr=False)
    node_qinv: np.ndarray = field(default=None, repr=False)
    node_elcon: np.ndarray = field(default=None, repr=False)

    def __post_init__(self):
        h = VP()
        _check(None, lib().gm_gfkfile_open(os.fsencode(self.path), C.byref(h)))
        self.h = h
        self._integers = None

    @classmethod
    def open(cls, path: str, paths=("identity", "reals", "integers", "numberOfWavenumbersForFrequency",
                                    "externalNodes", "dataSpecificIntegers")):
        """Opens path and reads the meta data computeSyntheticSeismograms reads
        (computeSyntheticSeismograms.f90:190-203)."""
        g = cls(path)
        for p in paths:
            g.read_meta(p)
        return g

    def close(self):
        if getattr(self, "h", None):
            h, self.h = self.h, None
            _check(None, lib().gm_gfkfile_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _string(self, which):
        buf = C.create_string_buffer(4096)
        _check(self.h, lib().gm_gfkfile_read_string(self.h, which, buf, 4096))
        return buf.value.decode()

    def read_meta(self, path: str):
        """readMetaGreenFKSpectra (greenFKSpectra.f90:71-137) for one path."""
        L = lib()
        if path == "identity":
            self.identity = self._string(IDENTITY)
        elif path == "earthModelName":
            self.earth_model_name = self._string(EARTH_MODEL_NAME)
        elif path == "attenuationMode":
            self.attnmode = self._string(ATTENUATION_MODE)
        elif path == "reals":
            r = np.zeros(6)
            _check(self.h, L.gm_gfkfile_read_meta(self.h, r.ctypes.data_as(D), None))
            self.rearth, self.rse, self.sigma, self.df, self.dwn, self.fref = (float(x) for x in r)
        elif path == "integers":
            ii = np.zeros(8, np.int32)
            _check(self.h, L.gm_gfkfile_read_meta(self.h, None, ii.ctypes.data_as(I)))
            self._integers = ii
            self.nf1, self.nf2, self.nkfmax, self.nwnmax = (int(x) for x in ii[:4])
            dsvstep, derivflag = int(ii[4]), int(ii[5])
            self.numtasks, self.global_ = int(ii[6]), int(ii[7])
            m = np.array([1, 0, 1, 0, 0, 0, 1, 0, 0], np.int32)
            if derivflag == 1:
                m[[4, 5, 8]] = 1
            if dsvstep == 1:
                m[[1, 3, 7]] = 1
            self.dsvmask = m
        elif path == "numberOfWavenumbersForFrequency":
            ii = np.zeros(8, np.int32)               # the array holds nf2 entries
            _check(self.h, L.gm_gfkfile_read_meta(self.h, None, ii.ctypes.data_as(I)))
            buf = np.zeros(max(int(ii[1]), 1), np.int32)
            _check(self.h, L.gm_gfkfile_read_nwn(self.h, buf.size, buf.ctypes.data_as(I)))
            self.nwn = buf
        elif path == "externalNodes":
            n = C.c_int()
            _check(self.h, L.gm_gfkfile_node_count(self.h, C.byref(n)))
            nn = n.value
            re = C.c_double()
            rn, ro, q = np.zeros(nn), np.zeros(nn), np.zeros((nn, 2))
            st, la = np.zeros(nn, np.int32), np.zeros(nn, np.int32)
            el = np.zeros((nn, 7), np.complex128)
            _check(self.h, L.gm_gfkfile_read_nodes(self.h, C.byref(re), rn.ctypes.data_as(D), st.ctypes.data_as(I),
                                                   la.ctypes.data_as(I), ro.ctypes.data_as(D), q.ctypes.data_as(D),
                                                   el.ctypes.data_as(D)))
            self.exnod = ExternalRadialNodes(rearth=re.value, rnod=rn, state=st, layer=la)
            self.node_ro, self.node_qinv, self.node_elcon = ro, q, el
        elif path == "sourceNodeRadii":
            n = C.c_int()
            _check(self.h, L.gm_gfkfile_read_source_nodes(self.h, C.byref(n), None))
            r = np.zeros(n.value)
            _check(self.h, L.gm_gfkfile_read_source_nodes(self.h, C.byref(n), r.ctypes.data_as(D)))
            self.rsnod = r
        elif path == "dataSpecificIntegers":
            d = np.zeros(5, np.int32)
            _check(self.h, L.gm_gfkfile_read_dsi(self.h, d.ctypes.data_as(I)))
            self.istyp, self.ncsph, self.nctor, self.njsph, self.njtor = (int(x) for x in d)
            if self.ncsph == 2:
                self.dsvmask[3] = 0
            if self.nctor == 0:
                self.dsvmask[6:9] = 0
        else:
            raise GfkFileError(9, f"Path_{path} does not exist in HDF file")

    def number_of_spectra(self) -> int:
        """numberGreenFKSpectra (greenFKSpectra.f90:210-215)."""
        return int(self.dsvmask[:6].sum()) * self.njsph + int(self.dsvmask[6:].sum()) * self.njtor

    @property
    def layout(self) -> str:
        """the synthetics engine's name of this file's spectrum layout (dsvmask, greenFKSpectra.f90:
        107-133): "uvw", "stress" (STRESSFLAG), "deriv" (radial derivatives), "ocean" or
        "ocean_stress" (a receiver in the ocean, no W)"""
        m = self.dsvmask
        if self.nctor == 0:
            return "ocean_stress" if m[1] else "ocean"
        if m[4]:
            return "deriv"
        return "stress" if m[1] else "uvw"

    def position(self, isp: int, jsp: int) -> int:
        """positionDataGreenFKSpectra (greenFKSpectra.f90:197-205): 1-based spectrum index of
        component isp (1..9) and jump jsp."""
        m = self.dsvmask
        if isp <= 6:
            return int(m[:isp - 1].sum()) * self.njsph + jsp
        return int(m[:6].sum()) * self.njsph + int(m[6:isp - 1].sum()) * self.njtor + jsp

    def read_data_node(self, jr: int, k0: int = 0, nk: int | None = None, nsp: int | None = None):
        """readDataNodeGreenFKSpectra (greenFKSpectra.f90:146-167): gfkr, gfki [nsp, nkfmax] float32
        of node jr (1-based); k0/nk select a wavenumber range.  nsp: the spectra per node where the
        integers do not tell them (a file written with STRESSFLAG carries dsvstep = 2 there)."""
        nsp = self.number_of_spectra() if nsp is None else int(nsp)
        nk = self.nkfmax - k0 if nk is None else nk
        gr, gi = np.zeros((nsp, nk), np.float32), np.zeros((nsp, nk), np.float32)
        _check(self.h, lib().gm_gfkfile_read_node(self.h, int(jr), nsp, int(k0), int(nk), gr.ctypes.data_as(F),
                                                  gi.ctypes.data_as(F)))
        return gr, gi

    def node_index(self, rs: float) -> int:
        """getNodeIndexFromRadiusGreenFKSpectra (greenFKSpectra.f90:239-252): the node closest to
        rs among the file's (single precision) radii, 1-based."""
        return self.exnod.node_index(rs)

    def check_identity(self, other: "GreenFKSpectra") -> bool:
        """checkIdentityGreenFKSpectra (greenFKSpectra.f90:171-185): other's identity attribute
        equals this one's."""
        return other._string(IDENTITY) == self.identity


def write_green_fk_spectra_for_synthetics(engine, path: str, grid: FrequencyGrid, jre: int, eps: float,
                                          istyp: int = 1, attmode: str = "ATTENUATION_AND_DISPERSION",
                                          earth_model_name: str = "", identity: str | None = None,
                                          global_: int = 1, dwn: float | None = None, nwnmax: int | None = None,
                                          freq_block: int = 64, stress: bool = False):
    """Computes the Green-FK spectra of every frequency of grid on the GPU and writes the GFK file
    as computeGreenFKSpectraForSynthetics does (computeGreenFKSpectraForSynthetics.f90:236-345,
    500-519), receiver node jre in a solid layer or the ocean (ncsph = 2, nctor = 0: U and V of the
    spheroidal jumps, :293-300).  Global mode: degrees l = 0 .. nwn-1; regional
    mode (global_ = 0): wavenumbers k = (iwn-1) dwn, iwn = 1 .. nwn, with k = 0 written as zero.
    stress (STRESSFLAG, :209-212): the datasets hold U, R, V, S and W, T (20 or 10 spectra; 16 or
    8 in the ocean), while the integers carry dsvstep = 2 as the reference's writer puts there
    (:283-285), so a reader of such a file is given the stress layout by its caller.
    Returns the identity string."""
    nodes = engine.nodes
    nem = engine.nem
    nf1, nf2 = int(grid.nf1), int(grid.nf2)
    nwn = np.zeros(nf2, np.int32)
    nwn[nf1 - 1:] = grid.nwn
    nkfmax = int(nwn.sum())
    dwn = float(getattr(grid, "dwn", 0.0) if dwn is None else dwn)
    if global_ == 0 and not dwn > 0.0:
        raise ValueError("regional mode needs the wavenumber spacing dwn")
    njsph, njtor = (4, 2) if istyp == 1 else (2, 1)
    ncsph, nctor = (2, 0) if nodes.state[jre - 1] == 1 else (4, 2)   # receiver in the ocean or a solid
    nsp = writer_number_of_spectra(1 if stress else 2, ncsph, njsph, njtor)
    identity = fdate() if identity is None else identity
    ro, elcon, qinv = engine.node_properties(nem.fref, "ELASTIC")
    reals = (nem.rearth, nodes.rnod[jre - 1], grid.sigma, grid.df, dwn, nem.fref)
    integers = (nf1, nf2, nkfmax, int(grid.nwn.max()) if nwnmax is None else int(nwnmax), 2, 0, 1, int(global_))
    with GfkFileWriter(path) as w:
        w.write_identity(identity)
        w.write_meta(earth_model_name, attmode, reals, integers, nwn)
        w.write_nodes(nodes, ro, qinv, elcon)
        w.create_spectra((istyp, ncsph, nctor, njsph, njtor), nsp, nodes.nnod, nf1, nf2, nwn)
        freqs = grid.freqs
        for b in range(0, freqs.size, freq_block):
            fb = freqs[b:b + freq_block]
            nb = nwn[nf1 - 1 + b:nf1 - 1 + b + fb.size]
            if global_ == 1:
                r = engine.gfk(fb, 0, nb, grid.sigma, eps, jre, istyp, attmode, want_gf=False, stress=stress)
                gre, gim = r["gfre"], r["gfim"]
            else:                          # iwn = 2 .. nwn per frequency, column iwn = 1 stays zero
                r = engine.gfk(fb, 2, nb - 1, grid.sigma, eps, jre, istyp, attmode, dwn=dwn, want_gf=False,
                               stress=stress)
                cols = np.concatenate([o + 1 + np.arange(n - 1) for o, n in zip(np.cumsum(nb) - nb, nb)])
                gre = np.zeros((nodes.nnod, nsp, int(nb.sum())), np.float32)
                gim = np.zeros_like(gre)
                gre[:, :, cols], gim[:, :, cols] = r["gfre"], r["gfim"]
            w.write_frequencies(nf1 + b, nf1 + b + fb.size - 1, gre, gim)
    return identity
