"""gm_gfk_batch_noalloc (the Green-FK batch that allocates nothing) against
gm_gfk_batch, through the C ABI: two calls in a row, the second with a smaller
input, each bitwise equal to the allocating entry point's columns in the exact
variant; oversized batches are rejected before any launch."""
import ctypes as C

import numpy as np
import pytest

from gemini_unified_amd import ExternalRadialNodes, GreenFKEngine, global_model, _lib

pytestmark = pytest.mark.gpu

SIGMA, EPS = 5.0 / 2048.0, 1e-6
GM_ERR_ARG = 9


@pytest.fixture(scope="module")
def case():
    nem = global_model(1, 200)
    nodes = ExternalRadialNodes.from_blocks(nem, [3], [100000.0])
    eng = GreenFKEngine(nem, nodes, device=0, variant="exact")
    yield eng, nodes.node_index(nem.rearth)
    eng.close()


class DeviceArray:
    """Device memory of the caller, from the HIP runtime the library itself is linked against (a second
    runtime in the process, such as torch's own, would not see the same device state)."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            cls._hip = C.CDLL("libamdhip64.so.7")
            cls._hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            cls._hip.hipFree.argtypes = [C.c_void_p]
            cls._hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            cls._hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
            cls._hip.hipDeviceSynchronize.argtypes = []
        return cls._hip

    def __init__(self, shape, dtype, fill_byte=0xFF):
        self.shape, self.dtype = shape, np.dtype(dtype)
        self.nbytes = int(np.prod(shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)) == 0
        assert self.hip().hipMemset(self.ptr, fill_byte, self.nbytes) == 0     # NaN floats, -1 ints
        assert self.hip().hipDeviceSynchronize() == 0

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        assert self.hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, 2) == 0   # D2H
        return out

    def __del__(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)


def _noalloc(eng, freqs, lmin, nl, jre, istyp=1):
    freqs = np.ascontiguousarray(freqs, dtype=np.float64)
    lmin = np.ascontiguousarray(lmin, dtype=np.int32)
    nl = np.ascontiguousarray(nl, dtype=np.int32)
    n, nn, nsp = int(nl.sum()), eng.nodes.nnod, 10 if istyp == 1 else 5
    m = max(n, 1)          # full-size outputs even for the batches that must be rejected
    gf = DeviceArray((nn, nsp, m, 2), np.float64)
    gfre, gfim = DeviceArray((nn, nsp, m), np.float32), DeviceArray((nn, nsp, m), np.float32)
    status = DeviceArray((m,), np.int32)
    st = np.zeros(10, np.int64)
    rc = _lib.lib().gm_gfk_batch_noalloc(eng.h, freqs.size, freqs.ctypes.data_as(_lib.D), lmin.ctypes.data_as(_lib.I),
                                         nl.ctypes.data_as(_lib.I), SIGMA, EPS, 3, int(jre), int(istyp), gf.ptr,
                                         gfre.ptr, gfim.ptr, status.ptr, st.ctypes.data_as(_lib.LL))
    out = gf.numpy()
    return rc, out[..., 0] + 1j * out[..., 1], gfre.numpy(), gfim.numpy(), status.numpy(), st


def test_noalloc_matches_gfk_batch_twice(case):
    eng, jre = case
    for freqs, lmin, nl in (([40 / 2048.0, 90 / 2048.0, 13 / 2048.0], [0, 120, 400], [60, 48, 36]),
                            ([70 / 2048.0], [7], [20])):
        print(f"no-alloc batch: {len(freqs)} frequencies, {sum(nl)} columns", flush=True)
        rc, gf, gr, gi, status, st = _noalloc(eng, freqs, lmin, nl, jre)
        assert rc == 0, _lib.last_error(eng.h)
        ref = eng.gfk(freqs, lmin, nl, SIGMA, EPS, jre, 1)
        assert np.all(status == 0)
        both = np.isnan(gf) & np.isnan(ref["gf"])
        assert np.array_equal(np.where(both, 0, gf), np.where(both, 0, ref["gf"]))
        np.testing.assert_array_equal(gr, ref["gfre"])
        np.testing.assert_array_equal(gi, ref["gfim"])
        np.testing.assert_array_equal(st, ref["stats"])


def test_noalloc_rejects_oversized_batches(case):
    eng, jre = case
    # 1 frequency x 100000 degrees: over 2 KB of buffers per column at 4 or more nodes, more than the
    # 128 MiB array (rejected before anything is launched)
    assert eng.nodes.nnod >= 4
    rc, *_ = _noalloc(eng, [0.01], [0], [100000], jre)
    assert rc == GM_ERR_ARG and "without allocation" in _lib.last_error(eng.h)
    freqs = np.full(1025, 0.01)
    rc, *_ = _noalloc(eng, freqs, np.zeros(1025, np.int32), np.ones(1025, np.int32), jre)
    assert rc == GM_ERR_ARG
    # the array still serves a normal batch after a rejection
    rc, gf, *_ = _noalloc(eng, [0.02], [3], [8], jre)
    assert rc == 0
