"""Malformed models and nodes are rejected by the C ABI before anything indexes them
(gm_model_set, gm_nodes_set and the batch calls).  A context needs a device; no
kernel is launched by these calls."""
import ctypes as C

import numpy as np
import pytest

from gemini_unified_amd import ExternalRadialNodes, GreenFKEngine, global_model, _lib
from gemini_unified_amd._lib import D, I, GeminiFKError

pytestmark = pytest.mark.gpu

GM_ERR_ARG = 9


@pytest.fixture(scope="module")
def eng():
    nem = global_model(1, 200)
    nodes = ExternalRadialNodes.from_blocks(nem, [1], [30000.0])
    e = GreenFKEngine(nem, nodes, device=0, variant="exact")
    yield e
    e.close()


def _model_args(nem, **over):
    a = {k: np.array(getattr(nem, k), dtype=np.float64) for k in
         ("rk", "ro", "vpv", "vph", "vsv", "vsh", "eta", "qkinv", "qminv", "rb")}
    a["iktop"] = np.array(nem.iktop, dtype=np.int32)
    a["iflso"] = np.array(nem.iflso, dtype=np.int32)
    s = {"nk": nem.nk, "nlay": nem.nlay, "nic": nem.nic, "noc": nem.noc, "fsflag": nem.fsflag,
         "equidis": nem.equidis, "rearth": nem.rearth}
    for k, v in over.items():
        if k in a:
            a[k] = v
        else:
            s[k] = v
    return a, s


def _set(eng, a, s):
    L = _lib.lib()
    return L.gm_model_set(eng.h, s["nk"], *[a[k].ctypes.data_as(D) for k in
                                            ("rk", "ro", "vpv", "vph", "vsv", "vsh", "eta", "qkinv", "qminv")],
                          s["nlay"], a["iktop"].ctypes.data_as(I), a["rb"].ctypes.data_as(D),
                          a["iflso"].ctypes.data_as(I), s["nic"], s["noc"], s["fsflag"], s["equidis"], s["rearth"],
                          eng.nem.ric, eng.nem.roc, eng.nem.fref)


def _bad_cases(nem):
    ik = np.array(nem.iktop, np.int32)
    yield "iktop not increasing", dict(iktop=np.array([ik[0], ik[2], ik[1], ik[3]], np.int32))
    yield "layer with one knot", dict(iktop=np.array([ik[0], ik[0] + 1, ik[2], ik[3]], np.int32))
    yield "iktop(1) below 1", dict(iktop=np.array([0, ik[1], ik[2], ik[3]], np.int32))
    yield "iktop(nlay) != nk", dict(iktop=np.array([ik[0], ik[1], ik[2], ik[3] - 1], np.int32))
    yield "noc beyond nlay", dict(noc=nem.nlay + 1)
    yield "nic above noc", dict(nic=nem.noc + 1)
    yield "negative nic", dict(nic=-1)
    yield "isfluid not 0/1", dict(iflso=np.array([0, 0, 2, 0], np.int32))
    rk = np.array(nem.rk, np.float64)
    rk[30], rk[31] = rk[31], rk[30]
    yield "knot radii decrease", dict(rk=rk)
    rb = np.array(nem.rb, np.float64)
    rb[1] = rb[2] + 1.0
    yield "layer tops decrease", dict(rb=rb)
    rb = np.array(nem.rb, np.float64)
    rb[2] += 1000.0
    yield "layer top off its top knot", dict(rb=rb)
    ro = np.array(nem.ro, np.float64)
    ro[5] = np.nan
    yield "non-finite knot value", dict(ro=ro)
    yield "equidistant flag not 0/1", dict(equidis=3)


def test_malformed_models_are_rejected(eng):
    nem = eng.nem
    for name, over in _bad_cases(nem):
        a, s = _model_args(nem, **over)
        rc = _set(eng, a, s)
        print(f"{name}: rc {rc}, {_lib.last_error(eng.h)}", flush=True)
        assert rc == GM_ERR_ARG, name
    # a rejected model leaves the context without a model: batch calls fail cleanly
    with pytest.raises(GeminiFKError) as e:
        eng.minors([0.01], 0, 4, 5.0 / 2048, 1e-6, 2, 2)
    assert e.value.code == GM_ERR_ARG
    # the unmodified model is accepted again
    a, s = _model_args(nem)
    assert _set(eng, a, s) == 0


def test_malformed_nodes_are_rejected(eng):
    L = _lib.lib()
    nodes = eng.nodes
    base = (np.array(nodes.rnod, np.float64), np.array(nodes.state, np.int32), np.array(nodes.layer, np.int32))

    def setn(rn, st, ly):
        return L.gm_nodes_set(eng.h, rn.size, rn.ctypes.data_as(D), st.ctypes.data_as(I), ly.ctypes.data_as(I))

    rn, st, ly = (x.copy() for x in base)
    rn[0] = rn[1] + 1.0
    assert setn(rn, st, ly) == GM_ERR_ARG                  # not ascending
    rn, st, ly = (x.copy() for x in base)
    st[0] = 5
    assert setn(rn, st, ly) == GM_ERR_ARG                  # state not 0/1
    rn, st, ly = (x.copy() for x in base)
    ly[0] = 0
    assert setn(rn, st, ly) == GM_ERR_ARG                  # layer below 1
    # a layer index beyond the model's layers passes gm_nodes_set (the model may come later) but
    # stops the batch call before any launch
    rn, st, ly = (x.copy() for x in base)
    ly[0] = eng.nem.nlay + 1
    assert setn(rn, st, ly) == 0
    with pytest.raises(GeminiFKError) as e:
        eng.minors([0.01], 0, 4, 5.0 / 2048, 1e-6, 2, 2)
    assert e.value.code == GM_ERR_ARG
    rn, st, ly = base
    assert setn(rn, st, ly) == 0
    out = eng.minors([0.01], 0, 4, 5.0 / 2048, 1e-6, 2, 2)
    assert np.all(out["status"] == 0)
