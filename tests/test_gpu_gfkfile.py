"""GFK file written from the GPU batches (write_green_fk_spectra_for_synthetics)
and read back per node as computeSyntheticSeismograms reads it, against the
oracle: exact variant, so the FP32 file values are the oracle's FP64 columns
rounded, bit for bit; node properties at the reference frequency equal the
oracle's spline evaluation bit for bit."""
import os

import numpy as np
import pytest

import oracle
from gemini_unified_amd import ExternalRadialNodes, FrequencyGrid, GreenFKEngine, MC_PI, global_model, regional_model
from gemini_unified_amd import build as gbuild

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(gbuild.FILE_LIB), reason="libgemini_gfkfile.so not built")]

from gemini_unified_amd import gfkfile  # noqa: E402


def test_node_properties_match_oracle():
    nem = global_model(1, 200)
    prn = np.concatenate(([50000.0], nem.rearth - 150000.0 * np.arange(40, -1, -1)))
    nodes = ExternalRadialNodes.split_at_discos(nem, prn, nem.rearth)
    orc = oracle.Oracle(nem, nodes)
    for variant in ("exact", "fast"):
        eng = GreenFKEngine(nem, nodes, variant=variant)
        for f, am in ((nem.fref, 0), (0.037, 3)):
            print(f"node properties variant={variant} f={f} attmode={am}", flush=True)
            ro, el, q = eng.node_properties(f, am)
            orc.set_frequency(f, am)
            ref = [orc.eval_elcon(r, nl) for r, nl in zip(nodes.rnod, nodes.layer)]
            if variant == "exact" and am == 0:
                # elastic moduli at fref (the file's case) take no transcendental function
                np.testing.assert_array_equal(ro, [x[0] for x in ref])
                np.testing.assert_array_equal(el, [x[1] for x in ref])
                np.testing.assert_array_equal(q, [x[2] for x in ref])
            elif variant == "exact":
                # attenuation and dispersion go through the device's pow/atan/sin/cos, which may differ
                # from the host's libm in the last bit; the spline arithmetic after them is the oracle's
                np.testing.assert_array_equal(ro, [x[0] for x in ref])
                np.testing.assert_allclose(el, [x[1] for x in ref], rtol=1e-15)
                np.testing.assert_array_equal(q, [x[2] for x in ref])
            else:
                np.testing.assert_allclose(ro, [x[0] for x in ref], rtol=1e-13)
                np.testing.assert_allclose(q, [x[2] for x in ref], rtol=1e-12, atol=1e-15)
        eng.close()


def test_global_file_matches_oracle(tmp_path):
    nem = global_model(1, 200)
    nodes = ExternalRadialNodes.from_blocks(nem, [2], [30000.0])
    jre = nodes.node_index(nem.rearth)
    tlen = 2048.0
    grid = FrequencyGrid(df=1 / tlen, nf1=20, nf2=24, sigma=5 / tlen, nwn=np.array([30, 31, 32, 33, 34]))
    eng = GreenFKEngine(nem, nodes, variant="exact")
    path = str(tmp_path / "gfk_glob.MOMENT")
    print("global GFK file, 5 frequencies", flush=True)
    ident = gfkfile.write_green_fk_spectra_for_synthetics(eng, path, grid, jre, 1e-6, istyp=1,
                                                          earth_model_name="test", freq_block=2)
    g = gfkfile.GreenFKSpectra.open(path)
    assert g.identity == ident and g.nkfmax == int(grid.nwn.sum()) and g.number_of_spectra() == 10
    assert g.rse == float(np.float32(nodes.rnod[jre - 1]))
    orc = oracle.Oracle(nem, nodes)
    ref = {}
    for i, f in enumerate(grid.freqs):
        orc.set_frequency(f, 3)
        rc, gf, _ = orc.gfk_degrees(2.0 * MC_PI * f, -grid.sigma, 0, int(grid.nwn[i]), jre, 1, 1e-6)
        assert rc == 0
        ref[i] = gf
    for jr in range(1, nodes.nnod + 1):
        gr, gi = g.read_data_node(jr)
        want_r = np.concatenate([ref[i][jr - 1].real for i in range(grid.freqs.size)], axis=1).astype(np.float32)
        want_i = np.concatenate([ref[i][jr - 1].imag for i in range(grid.freqs.size)], axis=1).astype(np.float32)
        np.testing.assert_array_equal(gr, want_r)
        np.testing.assert_array_equal(gi, want_i)
    # the external nodes' properties at the reference frequency, elastic
    orc.set_frequency(nem.fref, 0)
    ro_ref = np.array([orc.eval_elcon(r, nl)[0] for r, nl in zip(nodes.rnod, nodes.layer)])
    np.testing.assert_array_equal(g.node_ro, ro_ref.astype(np.float32))
    g.close()
    eng.close()


def test_regional_file_matches_oracle(tmp_path):
    nem = regional_model(6)
    nodes = ExternalRadialNodes.from_blocks(nem, [4], [5000.0])
    jre = nodes.node_index(nem.rearth)
    grid = FrequencyGrid.regional_fixed(tlen=64.0, fmax=2.0, nwn=14, xlen=8.0e6, fmin=1.5)
    eng = GreenFKEngine(nem, nodes, variant="exact")
    path = str(tmp_path / "gfk_reg.MOMENT")
    print(f"regional GFK file, {grid.freqs.size} frequencies", flush=True)
    gfkfile.write_green_fk_spectra_for_synthetics(eng, path, grid, jre, 1e-6, global_=0, freq_block=8)
    g = gfkfile.GreenFKSpectra.open(path)
    assert g.global_ == 0 and g.dwn == float(np.float32(grid.dwn))
    orc = oracle.Oracle(nem, nodes)
    gr, gi = g.read_data_node(1)
    off = 0
    for i, f in enumerate(grid.freqs):
        n = int(grid.nwn[i])
        assert not gr[:, off].any() and not gi[:, off].any()          # k = 0 written as zero
        orc.set_frequency(f, 3)
        rc, gf, _ = orc.gfk_wavenumbers(2.0 * MC_PI * f, -grid.sigma, 2, n - 1, grid.dwn, jre, 1, 1e-6)
        assert rc == 0
        np.testing.assert_array_equal(gr[:, off + 1:off + n], gf[0].real.astype(np.float32))
        np.testing.assert_array_equal(gi[:, off + 1:off + n], gf[0].imag.astype(np.float32))
        off += n
    g.close()
    eng.close()


def test_ocean_receiver_file_matches_oracle(tmp_path):
    """A receiver in the ocean: dataSpecificIntegers (istyp, 2, 0, njsph, njtor), 8 spectra per node
    (U, V), the file's values the oracle's columns rounded; the synthetics read it in its layout."""
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from ocean_model import ocean_model
    from test_oracle_ocean import ocean_receivers
    from gemini_unified_amd.synthetics import SyntheticsEngine
    nem = ocean_model()
    nodes = ocean_receivers(nem)
    jre = 8
    tlen = 2048.0
    grid = FrequencyGrid(df=1 / tlen, nf1=10, nf2=13, sigma=5 / tlen, nwn=np.array([10, 11, 11, 12]))
    eng = GreenFKEngine(nem, nodes, variant="exact")
    path = str(tmp_path / "gfk_ocean.MOMENT")
    print("ocean-receiver GFK file, 4 frequencies", flush=True)
    gfkfile.write_green_fk_spectra_for_synthetics(eng, path, grid, jre, 1e-6, istyp=1, freq_block=2)
    eng.close()
    g = gfkfile.GreenFKSpectra.open(path)
    assert (g.ncsph, g.nctor) == (2, 0) and g.number_of_spectra() == 8 and g.layout == "ocean"
    orc = oracle.Oracle(nem, nodes)
    ref = {}
    for i, f in enumerate(grid.freqs):
        orc.set_frequency(f, 3)
        rc, gf, _ = orc.gfk_degrees(2.0 * MC_PI * f, -grid.sigma, 0, int(grid.nwn[i]), jre, 1, 1e-6)
        assert rc == 0
        ref[i] = gf
    for jr in (1, 3, 6, 10):
        gr, gi = g.read_data_node(jr)
        want = np.concatenate([ref[i][jr - 1] for i in range(grid.freqs.size)], axis=1)
        np.testing.assert_array_equal(gr, want.real.astype(np.float32))
        np.testing.assert_array_equal(gi, want.imag.astype(np.float32))
    # synthetics of a source at node 3 (mantle) in the file's layout: those of U, V with W = 0
    gr, gi = g.read_data_node(3)
    g.close()
    nwn = np.zeros(grid.nf2, np.int32)
    nwn[grid.nf1 - 1:] = grid.nwn
    theta, phi, mt = np.array([0.3, 1.1]), np.array([0.2, 2.0]), np.array([1.0, -0.5, -0.5, 0.3, 0.1, -0.2])
    z = np.zeros((10, gr.shape[1]), np.float32)
    zr, zi = z.copy(), z.copy()
    zr[:8], zi[:8] = gr, gi
    syn = SyntheticsEngine(0)
    kw = dict(istyp=1, want_spectra=True)
    to, so = syn.synthetics(grid.nf1, grid.nf2, grid.df, grid.sigma, nwn, gr, gi, theta, phi, mt, 32, tlen / 32,
                            layout=g.layout, **kw)
    tu, su = syn.synthetics(grid.nf1, grid.nf2, grid.df, grid.sigma, nwn, zr, zi, theta, phi, mt, 32, tlen / 32, **kw)
    syn.close()
    assert np.array_equal(so, su) and np.array_equal(to, tu) and np.abs(so).max() > 0


@pytest.mark.parametrize("jre", [4, 8])
def test_stress_file_matches_oracle(tmp_path, jre):
    """A file written with STRESSFLAG (U, R, V, S and W, T per node; in the ocean U, R, V and zero S):
    the datasets hold the oracle's columns with stresses rounded, the integers carry dsvstep = 2 as
    the reference's writer puts there (computeGreenFKSpectraForSynthetics.f90:283-285)."""
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from ocean_model import ocean_model
    from test_oracle_ocean import ocean_receivers
    nem = ocean_model()
    nodes = ocean_receivers(nem)
    tlen = 2048.0
    grid = FrequencyGrid(df=1 / tlen, nf1=10, nf2=12, sigma=5 / tlen, nwn=np.array([10, 11, 12]))
    eng = GreenFKEngine(nem, nodes, variant="exact")
    path = str(tmp_path / "gfk_stress.MOMENT")
    print(f"stress GFK file, receiver node {jre}", flush=True)
    gfkfile.write_green_fk_spectra_for_synthetics(eng, path, grid, jre, 1e-6, istyp=1, freq_block=2, stress=True)
    eng.close()
    g = gfkfile.GreenFKSpectra.open(path)
    ocean = jre == 8
    nsp = 16 if ocean else 20
    assert g.ncsph == (2 if ocean else 4) and int(g._integers[4]) == 2
    orc = oracle.Oracle(nem, nodes)
    ref = {}
    for i, f in enumerate(grid.freqs):
        orc.set_frequency(f, 3)
        rc, gf, _ = orc.gfk_degrees_stress(2.0 * MC_PI * f, -grid.sigma, 0, int(grid.nwn[i]), jre, 1, 1e-6)
        assert rc == 0
        ref[i] = gf
    for jr in (1, 3, 6, 10):
        gr, gi = g.read_data_node(jr, nsp=nsp)
        want = np.concatenate([ref[i][jr - 1] for i in range(grid.freqs.size)], axis=1)
        np.testing.assert_array_equal(gr, want.real.astype(np.float32))
        np.testing.assert_array_equal(gi, want.imag.astype(np.float32))
    g.close()
