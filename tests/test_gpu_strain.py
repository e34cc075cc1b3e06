"""GPU parity of the strain and stress spectra (gm_strain_spectra: the 42 moment / 21 force
degree sums in three matrix-core contractions, displacement derivatives, spherical strain,
VTI stress) against the CPU oracle, on seeded random spectra with radial derivatives."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import oracle  # noqa: E402

pytestmark = pytest.mark.gpu

RE = 6371000.0


@pytest.fixture(scope="module")
def syn():
    from gemini_unified_amd.synthetics import SyntheticsEngine
    e = SyntheticsEngine(0)
    yield e
    e.close()


def _case(istyp, seed):
    rng = np.random.default_rng(seed)
    nf1, nf2 = 2, 11
    nwn = np.zeros(nf2, np.int32)
    nwn[nf1 - 1:] = np.linspace(120, 330, nf2 - nf1 + 1).astype(np.int32)
    nkf = int(nwn.sum())
    l = np.concatenate([np.arange(n) for n in nwn[nf1 - 1:]])
    env = np.exp(-l / 80.0) + 0.05
    nsp = 20 if istyp == 1 else 10
    gr = (rng.normal(size=(nsp, nkf)) * env).astype(np.float32)
    gi = (rng.normal(size=(nsp, nkf)) * env).astype(np.float32)
    theta = np.sort(rng.uniform(0.05, 3.09, size=45))
    phi = rng.uniform(0, 2 * np.pi, size=45)
    fomt = rng.normal(size=6 if istyp == 1 else 3)
    return nf1, nf2, nwn, gr, gi, theta, phi, fomt


@pytest.mark.parametrize("istyp", [1, 0])
def test_strain_spectra_match_oracle(syn, istyp):
    nf1, nf2, nwn, gr, gi, theta, phi, fomt = _case(istyp, 3 + istyp)
    el = np.array([2.1e11 + 1e8j, 1.9e11 + 2e8j, 0.8e11 + 1e7j, 0.6e11 + 3e8j, 0.7e11 + 2e8j, 1.3e11, 0.65e11])
    print(f"strain istyp={istyp}: {theta.size} receivers, {nf2 - nf1 + 1} frequencies", flush=True)
    g = syn.strain_spectra(nf1, nf2, nwn, gr, gi, theta, phi, RE, fomt, istyp=istyp, elcon=el, want_sums=True)
    o = oracle.strain_spectra(istyp, nf1, nf2, nwn, gr, gi, theta, phi, RE, fomt, elcon=el)
    # sums: error relative to the largest sum of each (kint, receiver), as for the displacement sums
    h = o["hdsum"].transpose(2, 0, 1)
    scale = np.abs(h).max(axis=2, keepdims=True) + 1e-300
    err = np.abs(g["hdsum"] - h) / scale
    print(f"  sums max rel {err.max():.2e}", flush=True)
    assert err.max() < 1e-5
    for key in ("zsp", "strain", "stress"):
        a, b = g[key], o[key]
        sc = np.abs(b).max(axis=-1, keepdims=True) + 1e-300
        e = (np.abs(a - b) / sc).max()
        print(f"  {key} max rel {e:.2e}", flush=True)
        assert e < 1e-5
        assert np.all(a[..., :nf1 - 1] == 0)


def test_strain_spectra_errors(syn):
    nf1, nf2, nwn, gr, gi, theta, phi, fomt = _case(1, 5)
    from gemini_unified_amd._lib import GeminiFKError
    with pytest.raises(GeminiFKError):          # displacement-only spectra have no radial derivatives
        syn.strain_spectra(nf1, nf2, nwn, gr[:10], gi[:10], theta, phi, RE, fomt, istyp=1)
    with pytest.raises(GeminiFKError):
        syn.strain_spectra(nf1, nf2, nwn, gr, gi, theta, phi, -1.0, fomt, istyp=1)


@pytest.mark.parametrize("istyp", [1, 0])
@pytest.mark.parametrize("layout", ["stress", "deriv"])
def test_file_layouts_give_the_displacement_of_their_uvw_spectra(syn, istyp, layout):
    """Files with stresses (STRESSFLAG: U, R, V, S, W, T) or radial derivatives (U, V, dU, dV, W, dW):
    the seismograms, global and regional, are those of the U, V, W spectra alone, bit for bit
    (positionDataGreenFKSpectra, greenFKSpectra.f90:204-212, with each file's dsvmask)."""
    nf1, nf2, nwn, gr, gi, theta, phi, fomt = _case(istyp, 9)
    nj, nt = (4, 2) if istyp == 1 else (2, 1)
    if layout == "stress":
        uvw = np.concatenate([np.arange(nj), 2 * nj + np.arange(nj), 4 * nj + np.arange(nt)])
    else:
        uvw = np.concatenate([np.arange(2 * nj), 4 * nj + np.arange(nt)])
    df, sigma = 1 / 2048.0, 5 / 2048.0
    kw = dict(istyp=istyp, want_spectra=True)
    ts20, z20 = syn.synthetics(nf1, nf2, df, sigma, nwn, gr, gi, theta, phi, fomt, 16, 1.0 / (16 * df),
                               layout=layout, **kw)
    ts10, z10 = syn.synthetics(nf1, nf2, df, sigma, nwn, gr[uvw], gi[uvw], theta, phi, fomt, 16, 1.0 / (16 * df), **kw)
    assert np.array_equal(z20, z10) and np.array_equal(ts20, ts10)
    x = theta * 1.0e5
    kwr = dict(dwn=2 * np.pi / 8.0e6, rearth=6371000.0, **kw)
    tr20, zr20 = syn.synthetics(nf1, nf2, df, sigma, nwn, gr, gi, x, phi, fomt, 16, 1.0 / (16 * df), layout=layout,
                                **kwr)
    tr10, zr10 = syn.synthetics(nf1, nf2, df, sigma, nwn, gr[uvw], gi[uvw], x, phi, fomt, 16, 1.0 / (16 * df), **kwr)
    assert np.array_equal(zr20, zr10) and np.array_equal(tr20, tr10)


@pytest.mark.parametrize("istyp", [1, 0])
@pytest.mark.parametrize("layout", ["ocean", "ocean_stress"])
def test_ocean_layouts_give_the_displacement_without_w(syn, istyp, layout):
    """Files of a receiver in the ocean (ncsph = 2, nctor = 0: U, V or U, R, V; greenFKSpectra.f90:
    129-133): the seismograms, global and regional, are those of U, V, W spectra with W = 0, bit for
    bit (the absent W's sums are zero, displacementSpectra.f90:84-87)."""
    nf1, nf2, nwn, gr, gi, theta, phi, fomt = _case(istyp, 9)
    nj, nt = (4, 2) if istyp == 1 else (2, 1)
    ocean = layout == "ocean"
    sel = np.arange(2 * nj if ocean else 3 * nj)                   # U, V or U, R, V
    vrows = (1 if ocean else 2) * nj + np.arange(nj)
    g10r = np.zeros((2 * nj + nt, gr.shape[1]), np.float32)        # U, V, W = 0
    g10i = np.zeros_like(g10r)
    g10r[:nj], g10i[:nj] = gr[:nj], gi[:nj]
    g10r[nj:2 * nj], g10i[nj:2 * nj] = gr[vrows], gi[vrows]
    df, sigma = 1 / 2048.0, 5 / 2048.0
    kw = dict(istyp=istyp, want_spectra=True)
    for kx in (dict(), dict(dwn=2 * np.pi / 8.0e6, rearth=6371000.0)):
        th = theta if not kx else theta * 1.0e5
        to, zo = syn.synthetics(nf1, nf2, df, sigma, nwn, gr[sel], gi[sel], th, phi, fomt, 16, 1.0 / (16 * df),
                                layout=layout, **kw, **kx)
        tu, zu = syn.synthetics(nf1, nf2, df, sigma, nwn, g10r, g10i, th, phi, fomt, 16, 1.0 / (16 * df), **kw, **kx)
        assert np.array_equal(zo, zu) and np.array_equal(to, tu)
        assert np.abs(zo).max() > 0


@pytest.mark.parametrize("istyp", [1, 0])
def test_regional_strain_spectra_match_oracle(syn, istyp):
    """The regional branch (besselDisplacementSpectra with strainsp): wavenumber integrals kint 1..42 /
    1..21 (displacement and radial-derivative spectra, and the x-derivative kernels) in three
    matrix-core contractions with their end terms, then the strain and stress of the angle x / R."""
    nf1, nf2, nwn, gr, gi, theta, phi, fomt = _case(istyp, 13 + istyp)
    x = 1.0e3 + theta / np.pi * 2.0e5                 # 1 .. 200 km
    dwn = 2 * np.pi / 8.0e6
    el = np.array([2.1e11 + 1e8j, 1.9e11 + 2e8j, 0.8e11 + 1e7j, 0.6e11 + 3e8j, 0.7e11 + 2e8j, 1.3e11, 0.65e11])
    print(f"regional strain istyp={istyp}: {x.size} receivers, {nf2 - nf1 + 1} frequencies", flush=True)
    g = syn.strain_spectra(nf1, nf2, nwn, gr, gi, x, phi, RE, fomt, istyp=istyp, elcon=el, want_sums=True,
                           dwn=dwn, rearth=RE)
    o = oracle.strain_spectra(istyp, nf1, nf2, nwn, gr, gi, x / RE, phi, RE, fomt, elcon=el, dwn=dwn, rearth=RE)
    h = o["hdsum"].transpose(2, 0, 1)
    scale = np.abs(h).max(axis=2, keepdims=True) + 1e-300
    err = np.abs(g["hdsum"] - h) / scale
    print(f"  integrals max rel {err.max():.2e}", flush=True)
    assert err.max() < 1e-5
    for key in ("zsp", "strain", "stress"):
        a, b = g[key], o[key]
        sc = np.abs(b).max(axis=-1, keepdims=True) + 1e-300
        e = (np.abs(a - b) / sc).max()
        print(f"  {key} max rel {e:.2e}", flush=True)
        assert e < 1e-5
        assert np.all(a[..., :nf1 - 1] == 0)
    from gemini_unified_amd._lib import GeminiFKError
    with pytest.raises(GeminiFKError):            # distances must be positive
        syn.strain_spectra(nf1, nf2, nwn, gr, gi, -x, phi, RE, fomt, istyp=istyp, dwn=dwn, rearth=RE)
