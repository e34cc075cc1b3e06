"""GPU parity of the synthetics half (Legendre functions, degree sums on the
matrix cores, seismograms) against the CPU oracle, through the C ABI."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import oracle  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def syn():
    from gemini_unified_amd.synthetics import SyntheticsEngine
    e = SyntheticsEngine(0)
    yield e
    e.close()


def test_legendre_matches_oracle(syn):
    thetas = np.array([0.0, 0.004, 0.3, 1.2, 2.5, 3.1])
    L = 2500
    g = syn.legendre(L, thetas)
    print(f"legendre: {thetas.size} receivers, lmax {L}", flush=True)
    for i, th in enumerate(thetas):
        o = oracle.legendre(L, th)
        for k in ("plm", "dpdt", "pdst", "d2pdt2", "dtpdst"):
            for m in range(3):
                ref = o[k][m]
                scale = max(np.abs(ref).max(), 1e-300)
                err = np.abs(g[k][i, m] - ref).max() / scale
                assert err < 1e-10, (th, k, m, err)


def _random_spectra(rng, nwn):
    nkf = int(nwn.sum())
    # spectra decaying with degree like evanescent FK spectra, FP32 as in the GFK file
    l = np.concatenate([np.arange(n) for n in nwn])
    env = np.exp(-l / 80.0) + 0.05
    gr = (rng.normal(size=(10, nkf)) * env).astype(np.float32)
    gi = (rng.normal(size=(10, nkf)) * env).astype(np.float32)
    return gr, gi


def test_moment_sums_match_oracle(syn):
    rng = np.random.default_rng(11)
    nf1, nf2 = 2, 9
    nwn_all = np.zeros(nf2, np.int32)
    nwn_all[nf1 - 1:] = np.linspace(150, 360, nf2 - nf1 + 1).astype(np.int32)
    gr, gi = _random_spectra(rng, nwn_all[nf1 - 1:])
    # l = 0 of the V spectra is 0/0 in real FK spectra; no sum reads it
    off = np.concatenate([[0], np.cumsum(nwn_all[nf1 - 1:])[:-1]])
    gr[4:8, off] = np.nan
    theta = np.sort(rng.uniform(0.05, 3.1, size=37))
    print(f"moment sums: {theta.size} receivers, {nf2 - nf1 + 1} frequencies", flush=True)
    g = syn.moment_sums(nf1, nf2, nwn_all, gr, gi, theta)                     # [14, R, nf]
    o = oracle.moment_sums(nf1, nf2, nwn_all, gr, gi, theta).transpose(2, 0, 1)
    scale = np.abs(o).max(axis=2, keepdims=True) + 1e-300
    assert np.all(np.isfinite(g)) and np.all(np.isfinite(o))
    err = np.abs(g - o) / scale
    print("max error relative to the largest sum of each (kint, receiver):", err.max(), flush=True)
    assert err.max() < 1e-5


def test_moment_sums_generating_function(syn):
    """sum_l P_l t^l = (1 - 2 x t + t^2)^(-1/2) through the GPU contraction."""
    L, t = 400, 0.6
    nwn = np.array([L], np.int32)
    gr = np.zeros((10, L), np.float32)
    gi = np.zeros((10, L), np.float32)
    l = np.arange(L)
    gr[0] = t ** l * 4 * float(np.float32(np.pi)) / (2 * l + 1)      # 4 pi of the reference (FP32 pi)
    theta = np.linspace(0.1, 3.0, 129)
    h = syn.moment_sums(1, 1, nwn, gr, gi, theta)[0, :, 0]
    ref = (1 - 2 * np.cos(theta) * t + t * t) ** -0.5
    err = np.abs(h - ref) / ref
    print("generating function max rel error", err.max(), flush=True)
    assert err.max() < 1e-6


def test_synthetics_match_oracle(syn):
    rng = np.random.default_rng(12)
    nf1, nf2, nsamp = 2, 64, 256
    T = 256.0
    df, dt = 1.0 / T, T / nsamp
    sigma = 5 * df
    nwn_all = np.zeros(nf2, np.int32)
    nwn_all[nf1 - 1:] = np.linspace(40, 120, nf2 - nf1 + 1).astype(np.int32)
    gr, gi = _random_spectra(rng, nwn_all[nf1 - 1:])
    R = 6
    theta = rng.uniform(0.1, 3.0, R)
    phi = rng.uniform(0, 2 * np.pi, R)
    propdir = rng.uniform(0, 2 * np.pi, R)
    tred = rng.uniform(0, 20, R)
    mt = rng.normal(size=6) * 1e18
    print(f"synthetics: {R} receivers, nf2 {nf2}, nsamp {nsamp}", flush=True)
    # the third case is the reference's non-automatic time shift: one reduction time for the event
    # (getTimeshiftSeismicEvent), given as a scalar and applied to every receiver
    for vel, pd, tr in [(False, None, None), (True, propdir, tred), (False, 0.0, 7.5)]:
        ts, zsp = syn.synthetics(nf1, nf2, df, sigma, nwn_all, gr, gi, theta, phi, mt, nsamp, dt, propdir=pd,
                                 tred=tr, velocity=vel, want_spectra=True)
        opd = None if pd is None else np.broadcast_to(np.asarray(pd, float), (R,)).copy()
        otr = None if tr is None else np.broadcast_to(np.asarray(tr, float), (R,)).copy()
        ots, ozsp = oracle.synthetics(nf1, nf2, df, sigma, nwn_all, gr, gi, theta, phi, mt, nsamp, dt, propdir=opd,
                                      tred=otr, velocity=vel)
        zerr = np.abs(zsp - ozsp).max() / np.abs(ozsp).max()
        ts64, ots64 = ts.astype(np.float64), ots.astype(np.float64)       # FP32 squares overflow at 1e18 N m
        terr = np.linalg.norm(ts64 - ots64, axis=2) / np.linalg.norm(ots64, axis=2)
        print(f"velocity={vel}: spectra max rel {zerr:.2e}, seismogram rel-L2 max {terr.max():.2e}", flush=True)
        assert zerr < 1e-5
        assert terr.max() < 1e-5


def test_synthetics_errors(syn):
    from gemini_unified_amd._lib import GeminiFKError
    nwn = np.array([10, 10], np.int32)
    g = np.zeros((3, 20), np.float32)
    with pytest.raises(GeminiFKError) as e:
        syn.synthetics(1, 2, 1 / 64, 0.0, nwn, g, g, [1.0], [0.0], np.ones(6), 64, 1.0)
    assert e.value.code == 8                        # neither 10 (moment) nor 5 (force) spectra
    g = np.zeros((10, 20), np.float32)
    with pytest.raises(GeminiFKError) as e:         # a force with stresses (STRESSFLAG) also has 10 spectra
        syn.synthetics(1, 2, 1 / 64, 0.0, nwn, g, g, [1.0], [0.0], np.ones(3), 64, 1.0, istyp=0)
    assert e.value.code == 8
    with pytest.raises(GeminiFKError) as e:
        syn.moment_sums(1, 2, nwn, np.zeros((5, 20), np.float32), np.zeros((5, 20), np.float32), [1.0])
    assert e.value.code == 9                        # force spectra given to the moment sums
    with pytest.raises(GeminiFKError) as e:
        syn.force_sums(1, 2, nwn, np.zeros((10, 20), np.float32), np.zeros((10, 20), np.float32), [1.0])
    assert e.value.code == 9
    g = np.zeros((10, 20), np.float32)
    with pytest.raises(GeminiFKError) as e:
        syn.synthetics(1, 2, 1 / 64, 0.0, nwn, g, g, [1.0], [0.0], np.ones(6), 128, 1.0)
    assert e.value.code == 9
    with pytest.raises(ValueError):                 # one reduction time per receiver, or one for all
        syn.synthetics(1, 2, 1 / 64, 0.0, nwn, g, g, [1.0, 2.0], [0.0, 0.0], np.ones(6), 64, 1.0, tred=[1.0, 2.0, 3.0])


def test_bessel_sums_match_oracle(syn):
    """Regional wavenumber integrals: receivers on both sides of the trapezoid/Filon switch
    (j3 = 3/(x dwn) + 2 beyond, at and below the last wavenumber of each frequency)."""
    rng = np.random.default_rng(21)
    nf1, nf2 = 2, 7
    nwn_all = np.zeros(nf2, np.int32)
    nwn_all[nf1 - 1:] = [800, 900, 1000, 1100, 1200, 1300]
    nkf = int(nwn_all.sum())
    dwn, re = 1e-5, 6371000.0
    k = np.concatenate([np.arange(n) * dwn for n in nwn_all[nf1 - 1:]])
    env = np.exp(-k / 4e-3) + 0.05
    gr = (rng.normal(size=(10, nkf)) * env).astype(np.float32)
    gi = (rng.normal(size=(10, nkf)) * env).astype(np.float32)
    x = np.array([150.0, 250.0, 375.0, 3.0 / (998.5 * dwn), 700.0, 1500.0, 5e3, 2e4, 1e5, 2e5])  # 4th: j3 = 1000
    print(f"bessel sums: {x.size} receivers, {nf2 - nf1 + 1} frequencies", flush=True)
    g = syn.bessel_sums(nf1, nf2, nwn_all, gr, gi, dwn, re, x)
    o = oracle.bessel_sums(nf1, nf2, nwn_all, gr, gi, dwn, re, x).transpose(2, 0, 1)
    assert np.all(np.isfinite(g))
    scale = np.abs(o).max(axis=2, keepdims=True) + 1e-300
    err = np.abs(g - o) / scale
    print("max error relative to the largest integral of each (kint, receiver):", err.max(), flush=True)
    assert err.max() < 1e-5


def test_bessel_sums_closed_form(syn):
    """(2/x) int J1(kx) exp(-ak) dk = (2/x)(1 - a/sqrt(a^2+x^2))/x through the GPU path."""
    pi32, re, a, dwn, nwn = float(np.float32(np.pi)), 6371000.0, 1000.0, 5e-6, 8000
    k = np.arange(nwn) * dwn
    gr = np.zeros((10, nwn), np.float32)
    gi = np.zeros((10, nwn), np.float32)
    gr[8] = (np.exp(-a * k) * 2 * pi32 / re ** 2 * 1e20).astype(np.float32)
    x = np.array([1000.0, 5000.0])
    h = syn.bessel_sums(1, 1, np.array([nwn], np.int32), gr, gi, dwn, re, x)[8, :, 0].real / 1e20
    ref = 2 / x * (1 - a / np.sqrt(a * a + x * x)) / x
    err = np.abs(h - ref) / ref
    print("closed form rel error", err, flush=True)
    assert np.all(err < 1e-4)


def test_regional_synthetics_match_oracle(syn):
    rng = np.random.default_rng(22)
    nf1, nf2, nsamp = 2, 40, 128
    T = 51.2
    df, dt = 1.0 / T, T / nsamp
    sigma = 5 * df
    dwn, re = 2e-5, 6371000.0
    nwn_all = np.zeros(nf2, np.int32)
    nwn_all[nf1 - 1:] = np.linspace(300, 600, nf2 - nf1 + 1).astype(np.int32)
    nkf = int(nwn_all.sum())
    k = np.concatenate([np.arange(n) * dwn for n in nwn_all[nf1 - 1:]])
    env = np.exp(-k / 3e-3) + 0.02
    gr = (rng.normal(size=(10, nkf)) * env).astype(np.float32)
    gi = (rng.normal(size=(10, nkf)) * env).astype(np.float32)
    R = 5
    x = rng.uniform(1e3, 2e5, R)
    phi = rng.uniform(0, 2 * np.pi, R)
    mt = rng.normal(size=6) * 1e18
    ts, zsp = syn.synthetics(nf1, nf2, df, sigma, nwn_all, gr, gi, x, phi, mt, nsamp, dt, velocity=True,
                             want_spectra=True, dwn=dwn, rearth=re)
    ots, ozsp = oracle.synthetics(nf1, nf2, df, sigma, nwn_all, gr, gi, x / re, phi, mt, nsamp, dt, velocity=True,
                                  dwn=dwn, rearth=re)
    zerr = np.abs(zsp - ozsp).max() / np.abs(ozsp).max()
    ts64, ots64 = ts.astype(np.float64), ots.astype(np.float64)
    terr = np.linalg.norm(ts64 - ots64, axis=2) / np.linalg.norm(ots64, axis=2)
    print(f"regional: spectra max rel {zerr:.2e}, seismogram rel-L2 max {terr.max():.2e}", flush=True)
    assert zerr < 1e-5 and terr.max() < 1e-5


def test_regional_gfk_columns_match_oracle():
    """Wavenumber-mode FK columns of the regional model, exact variant, against the oracle."""
    from gemini_unified_amd import ExternalRadialNodes, FrequencyGrid, GreenFKEngine, MC_PI, regional_model
    nem = regional_model(6)
    nodes = ExternalRadialNodes.from_blocks(nem, [8], [30000.0 / 8])
    jre = nodes.node_index(nem.rearth)
    grid = FrequencyGrid.regional_fixed(102.4, 5.0, 20000, 8.0e6)
    eng = GreenFKEngine(nem, nodes, device=0, variant="exact")
    orc = oracle.Oracle(nem, nodes)
    for f, iw0 in [(1.0, 2), (3.0, 4000), (5.0, 12000)]:
        n = 12
        r = eng.gfk([f], [iw0], [n], grid.sigma, 1e-6, jre, 1, want_fp32=False, dwn=grid.dwn)
        orc.set_frequency(f, 3)
        rc, ref, st = orc.gfk_wavenumbers(2.0 * MC_PI * f, -grid.sigma, iw0, n, grid.dwn, jre, 1, 1e-6)
        assert rc == 0 and np.all(r["status"] == 0)
        d = np.linalg.norm((r["gf"] - ref).reshape(-1, n), axis=0) / np.linalg.norm(ref.reshape(-1, n), axis=0)
        print(f"regional f={f} iwn={iw0}..{iw0 + n - 1}: max rel-L2 {d.max():.2e}", flush=True)
        assert d.max() < 1e-12
    eng.close()


# ---- single-force source (5 spectra U1 U2 V1 V2 W1, 7 sums) ----

def _random_force_spectra(rng, nwn):
    nkf = int(nwn.sum())
    l = np.concatenate([np.arange(n) for n in nwn])
    env = np.exp(-l / 80.0) + 0.05
    gr = (rng.normal(size=(5, nkf)) * env).astype(np.float32)
    gi = (rng.normal(size=(5, nkf)) * env).astype(np.float32)
    return gr, gi


def test_force_sums_match_oracle(syn):
    rng = np.random.default_rng(31)
    nf1, nf2 = 3, 10
    nwn_all = np.zeros(nf2, np.int32)
    nwn_all[nf1 - 1:] = np.linspace(130, 420, nf2 - nf1 + 1).astype(np.int32)
    gr, gi = _random_force_spectra(rng, nwn_all[nf1 - 1:])
    # every sum but kint 1 starts at l = 1: l = 0 of spectra U2, V1, V2, W1 is never read
    off = np.concatenate([[0], np.cumsum(nwn_all[nf1 - 1:])[:-1]])
    gr[1:, off] = np.nan
    theta = np.sort(rng.uniform(0.02, 3.12, size=41))
    print(f"force sums: {theta.size} receivers, {nf2 - nf1 + 1} frequencies", flush=True)
    g = syn.force_sums(nf1, nf2, nwn_all, gr, gi, theta)                      # [7, R, nf]
    o = oracle.force_sums(nf1, nf2, nwn_all, gr, gi, theta).transpose(2, 0, 1)
    assert g.shape == (7, theta.size, nf2 - nf1 + 1)
    assert np.all(np.isfinite(g)) and np.all(np.isfinite(o))
    err = np.abs(g - o) / (np.abs(o).max(axis=2, keepdims=True) + 1e-300)
    print("max error relative to the largest sum of each (kint, receiver):", err.max(), flush=True)
    assert err.max() < 1e-5


def test_force_sums_generating_function(syn):
    """kint 5 = -sum_l P_l^1/sin(theta) t^l = -t (1 - 2 x t + t^2)^(-3/2) through the GPU contraction."""
    L, t = 400, 0.6
    nwn = np.array([L], np.int32)
    gr = np.zeros((5, L), np.float32)
    gi = np.zeros((5, L), np.float32)
    l = np.arange(L)
    gi[4] = t ** l * 4 * float(np.float32(np.pi)) / (2 * l + 1) * l * (l + 1) / 2
    theta = np.linspace(0.1, 3.0, 129)
    h = syn.force_sums(1, 1, nwn, gr, gi, theta)[4, :, 0]
    ref = -1j * t * (1 - 2 * np.cos(theta) * t + t * t) ** -1.5
    err = np.abs(h - ref) / np.abs(ref)
    print("generating function max rel error", err.max(), flush=True)
    assert err.max() < 1e-6


def test_force_synthetics_match_oracle(syn):
    rng = np.random.default_rng(32)
    nf1, nf2, nsamp = 2, 64, 256
    T = 256.0
    df, dt = 1.0 / T, T / nsamp
    sigma = 5 * df
    nwn_all = np.zeros(nf2, np.int32)
    nwn_all[nf1 - 1:] = np.linspace(40, 120, nf2 - nf1 + 1).astype(np.int32)
    gr, gi = _random_force_spectra(rng, nwn_all[nf1 - 1:])
    R = 7
    theta = rng.uniform(0.1, 3.0, R)
    phi = rng.uniform(0, 2 * np.pi, R)
    propdir = rng.uniform(0, 2 * np.pi, R)
    tred = rng.uniform(0, 20, R)
    force = rng.normal(size=3) * 1e15
    print(f"force synthetics: {R} receivers, nf2 {nf2}, nsamp {nsamp}", flush=True)
    for vel, pd, tr in [(False, None, None), (True, propdir, tred)]:
        ts, zsp = syn.synthetics(nf1, nf2, df, sigma, nwn_all, gr, gi, theta, phi, force, nsamp, dt, propdir=pd,
                                 tred=tr, velocity=vel, want_spectra=True)
        ots, ozsp = oracle.synthetics(nf1, nf2, df, sigma, nwn_all, gr, gi, theta, phi, force, nsamp, dt,
                                      propdir=pd, tred=tr, velocity=vel)
        zerr = np.abs(zsp - ozsp).max() / np.abs(ozsp).max()
        ts64, ots64 = ts.astype(np.float64), ots.astype(np.float64)
        terr = np.linalg.norm(ts64 - ots64, axis=2) / np.linalg.norm(ots64, axis=2)
        print(f"velocity={vel}: spectra max rel {zerr:.2e}, seismogram rel-L2 max {terr.max():.2e}", flush=True)
        assert zerr < 1e-5
        assert terr.max() < 1e-5


def test_force_bessel_sums_match_oracle(syn):
    """Regional force wavenumber integrals on both sides of the trapezoid/Filon switch."""
    rng = np.random.default_rng(33)
    nf1, nf2 = 2, 7
    nwn_all = np.zeros(nf2, np.int32)
    nwn_all[nf1 - 1:] = [800, 900, 1000, 1100, 1200, 1300]
    nkf = int(nwn_all.sum())
    dwn, re = 1e-5, 6371000.0
    k = np.concatenate([np.arange(n) * dwn for n in nwn_all[nf1 - 1:]])
    env = np.exp(-k / 4e-3) + 0.05
    gr = (rng.normal(size=(5, nkf)) * env).astype(np.float32)
    gi = (rng.normal(size=(5, nkf)) * env).astype(np.float32)
    x = np.array([150.0, 250.0, 375.0, 3.0 / (998.5 * dwn), 700.0, 1500.0, 5e3, 2e4, 1e5, 2e5])  # 4th: j3 = 1000
    print(f"force bessel sums: {x.size} receivers, {nf2 - nf1 + 1} frequencies", flush=True)
    g = syn.force_bessel_sums(nf1, nf2, nwn_all, gr, gi, dwn, re, x)
    o = oracle.bessel_sums(nf1, nf2, nwn_all, gr, gi, dwn, re, x, istyp=0).transpose(2, 0, 1)
    assert g.shape == (7, x.size, nf2 - nf1 + 1) and np.all(np.isfinite(g))
    err = np.abs(g - o) / (np.abs(o).max(axis=2, keepdims=True) + 1e-300)
    print("max error relative to the largest integral of each (kint, receiver):", err.max(), flush=True)
    assert err.max() < 1e-5


def test_force_regional_synthetics_match_oracle(syn):
    rng = np.random.default_rng(34)
    nf1, nf2, nsamp = 2, 40, 128
    T = 51.2
    df, dt = 1.0 / T, T / nsamp
    sigma = 5 * df
    dwn, re = 2e-5, 6371000.0
    nwn_all = np.zeros(nf2, np.int32)
    nwn_all[nf1 - 1:] = np.linspace(300, 600, nf2 - nf1 + 1).astype(np.int32)
    nkf = int(nwn_all.sum())
    k = np.concatenate([np.arange(n) * dwn for n in nwn_all[nf1 - 1:]])
    env = np.exp(-k / 3e-3) + 0.02
    gr = (rng.normal(size=(5, nkf)) * env).astype(np.float32)
    gi = (rng.normal(size=(5, nkf)) * env).astype(np.float32)
    R = 5
    x = rng.uniform(1e3, 2e5, R)
    phi = rng.uniform(0, 2 * np.pi, R)
    force = rng.normal(size=3) * 1e15
    ts, zsp = syn.synthetics(nf1, nf2, df, sigma, nwn_all, gr, gi, x, phi, force, nsamp, dt, velocity=True,
                             want_spectra=True, dwn=dwn, rearth=re)
    ots, ozsp = oracle.synthetics(nf1, nf2, df, sigma, nwn_all, gr, gi, x / re, phi, force, nsamp, dt,
                                  velocity=True, dwn=dwn, rearth=re)
    zerr = np.abs(zsp - ozsp).max() / np.abs(ozsp).max()
    ts64, ots64 = ts.astype(np.float64), ots.astype(np.float64)
    terr = np.linalg.norm(ts64 - ots64, axis=2) / np.linalg.norm(ots64, axis=2)
    print(f"regional force: spectra max rel {zerr:.2e}, seismogram rel-L2 max {terr.max():.2e}", flush=True)
    assert zerr < 1e-5 and terr.max() < 1e-5


# ---- event and station filters ----

def test_event_filter_matches_oracle():
    from gemini_unified_amd.synthetics import event_filter
    from gemini_unified_amd._lib import GeminiFKError
    for spec in ([2, 0.01, 4, 0.08], [0, 0.01, 6, 0.15], [3, 0.004, 0, 1.0]):
        h = event_filter(spec, 3, 300, 1.0 / 512, 5.0 / 512)
        o = oracle.butterworth_bandpass(spec, 3, 300, 1.0 / 512, 5.0 / 512)
        assert np.all(h[:2] == 0)
        np.testing.assert_allclose(h, o, rtol=1e-12, atol=1e-15)
    with pytest.raises(GeminiFKError) as e:
        event_filter([2, 0.01, 4, 0.08], 3, 300, 1.0 / 512, 5.0 / 512, filtype=7)
    assert e.value.code == 9


def test_filtered_synthetics_match_oracle(syn):
    from gemini_unified_amd.synthetics import event_filter
    rng = np.random.default_rng(41)
    nf1, nf2, nsamp = 2, 64, 256
    T = 256.0
    df, dt = 1.0 / T, T / nsamp
    sigma = 5 * df
    nwn_all = np.zeros(nf2, np.int32)
    nwn_all[nf1 - 1:] = np.linspace(40, 120, nf2 - nf1 + 1).astype(np.int32)
    gr, gi = _random_spectra(rng, nwn_all[nf1 - 1:])
    R = 6
    theta = rng.uniform(0.1, 3.0, R)
    phi = rng.uniform(0, 2 * np.pi, R)
    propdir = rng.uniform(0, 2 * np.pi, R)
    tred = rng.uniform(0, 20, R)
    mt = rng.normal(size=6) * 1e18
    evf = event_filter([2, 0.01, 4, 0.15], nf1, nf2, df, sigma)
    stf = rng.normal(size=(R, 3, nf2)) + 1j * rng.normal(size=(R, 3, nf2))
    print(f"filtered synthetics: {R} receivers, nf2 {nf2}", flush=True)
    ts, zsp = syn.synthetics(nf1, nf2, df, sigma, nwn_all, gr, gi, theta, phi, mt, nsamp, dt, propdir=propdir,
                             tred=tred, velocity=True, want_spectra=True, evfilter=evf, stafilter=stf)
    ots, ozsp = oracle.synthetics(nf1, nf2, df, sigma, nwn_all, gr, gi, theta, phi, mt, nsamp, dt, propdir=propdir,
                                  tred=tred, velocity=True, evfilter=evf, stafilter=stf)
    zerr = np.abs(zsp - ozsp).max() / np.abs(ozsp).max()
    ts64, ots64 = ts.astype(np.float64), ots.astype(np.float64)
    terr = np.linalg.norm(ts64 - ots64, axis=2) / np.linalg.norm(ots64, axis=2)
    print(f"spectra max rel {zerr:.2e}, seismogram rel-L2 max {terr.max():.2e}", flush=True)
    assert zerr < 1e-5 and terr.max() < 1e-5


def test_frequency_time_every_length(syn):
    """The transform for every nsamp = 2^2 .. 2^13 (every mix of radix-16 and radix-2/4/8 passes),
    with nearly the most frequencies each length takes without decimation (transformFrequencyTime's
    test ceil(log2 nf2) + 1 <= log2 nsamp, in FP32 logarithms), against the oracle."""
    rng = np.random.default_rng(51)
    for lg in range(2, 14):
        nsamp = 1 << lg
        nf2 = max(nsamp // 2 - 1, 1)
        nf1 = min(2, nf2)
        df = 1.0 / 64.0
        dt = 1.0 / (df * nsamp)
        nwn = np.zeros(nf2, np.int32)
        nwn[nf1 - 1:] = 12
        gr, gi = _random_spectra(rng, nwn[nf1 - 1:])
        mt = rng.normal(size=6)
        ts, zsp = syn.synthetics(nf1, nf2, df, 5 * df, nwn, gr, gi, [1.1], [0.4], mt, nsamp, dt, want_spectra=True)
        ots, ozsp = oracle.synthetics(nf1, nf2, df, 5 * df, nwn, gr, gi, [1.1], [0.4], mt, nsamp, dt)
        ts64, ots64 = ts.astype(np.float64), ots.astype(np.float64)
        terr = (np.linalg.norm(ts64 - ots64, axis=2) / np.linalg.norm(ots64, axis=2)).max()
        print(f"nsamp {nsamp}: seismogram rel-L2 {terr:.2e}", flush=True)
        assert terr < 1e-5, (nsamp, terr)


@pytest.mark.parametrize("nsamp,ndec", [(128, 4), (64, 8), (300, 1), (100, 2)])
def test_decimated_synthetics_match_oracle(syn, nsamp, ndec):
    """Fewer samples than the frequencies need: the transform runs at the full length and keeps every
    ndec-th sample, as transformFrequencyTime's decimation branch does, also for ndec = 1 (the first
    nsamp samples at the transform's own step) and for nsamp that is no power of two; a dt that is no
    multiple of the transform's step is an error."""
    from gemini_unified_amd._lib import GeminiFKError
    rng = np.random.default_rng(61)
    nf1, nf2, T = 2, 200, 256.0                      # 200 frequencies need 512 samples
    df, dt = 1.0 / T, ndec * T / 512
    sigma = 5 * df
    nwn = np.zeros(nf2, np.int32)
    nwn[nf1 - 1:] = 20
    gr, gi = _random_spectra(rng, nwn[nf1 - 1:])
    theta, phi = rng.uniform(0.2, 3.0, 3), rng.uniform(0, 6.2, 3)
    mt = rng.normal(size=6)
    print(f"decimated synthetics: nsamp {nsamp} of 512, every {ndec}", flush=True)
    ts = syn.synthetics(nf1, nf2, df, sigma, nwn, gr, gi, theta, phi, mt, nsamp, dt, velocity=True)
    ots, _ = oracle.synthetics(nf1, nf2, df, sigma, nwn, gr, gi, theta, phi, mt, nsamp, dt, velocity=True)
    ts64, ots64 = ts.astype(np.float64), ots.astype(np.float64)
    terr = (np.linalg.norm(ts64 - ots64, axis=2) / np.linalg.norm(ots64, axis=2)).max()
    print(f"seismogram rel-L2 {terr:.2e}", flush=True)
    assert ts.shape == (3, 3, nsamp) and terr < 1e-5
    with pytest.raises(GeminiFKError) as e:
        syn.synthetics(nf1, nf2, df, sigma, nwn, gr, gi, theta, phi, mt, nsamp, 1.3 * dt)
    assert e.value.code == 9


def _ray_tables():
    rng = np.random.default_rng(11)
    R, v = 6371e3, 8000.0
    for n, tele in ((400, False), (400, True), (150, False)):
        yield oracle.homogeneous_ray_table(R, v, n, istele=tele)
    # a rough table: seeded perturbations make distance versus turning point
    # non-monotonic, so the search from the end meets several sign changes
    T = oracle.homogeneous_ray_table(R, v, 300)
    T["delta"] = T["delta"] * (1 + 0.02 * rng.standard_normal(T["delta"].shape))
    T["ttime"] = T["ttime"] * (1 + 0.02 * rng.standard_normal(T["ttime"].shape))
    yield T


def test_travel_times_match_oracle(syn):
    """Every branch (steep, turning, diffracted, out of range) against the
    oracle: same FP32 table, same unfused FP64 arithmetic in the same order,
    so the times agree to rounding (1e-13 relative) and the statuses exactly."""
    rng = np.random.default_rng(5)
    R = 6371e3
    for T in _ray_tables():
        for rs in (6000e3, R, 5700e3, 0.999 * R):
            nr = 3000
            re = rng.choice([R, 6200e3, 5900e3, 0.25 * R, 1.001 * R, R * (1 + 4e-6)], nr)
            d = rng.uniform(0.0, np.pi, nr)
            tt, p, st = syn.travel_times(T, rs, re, d)
            to, po, so = oracle.travel_times(T, rs, re, d)
            print(f"travel times: nnod {T['rnod'].size} istele {T['istele']} rs {rs:.0f}", flush=True)
            np.testing.assert_array_equal(st, so)
            assert (st == 0).sum() > nr // 2 and (st != 0).any()
            np.testing.assert_allclose(tt, to, rtol=1e-13, atol=1e-9)
            np.testing.assert_allclose(p, po, rtol=1e-13, atol=1e-12)
            assert (tt[st != 0] == -1.0).all()


def test_travel_times_errors_and_empty(syn):
    from gemini_unified_amd._lib import GeminiFKError
    T = oracle.homogeneous_ray_table(6371e3, 8000.0, 200)
    tt, p, st = syn.travel_times(T, 6000e3, np.zeros(0), np.zeros(0))
    assert tt.size == 0
    with pytest.raises(GeminiFKError) as e:
        syn.travel_times(T, 6000e3, np.array([6371e3, 1e6]), np.array([0.5, 0.5]), raise_on_error=True)
    assert e.value.code == 9 and "no turning point below receiver" in str(e.value)
    with pytest.raises(GeminiFKError):
        syn.travel_times(T, -1.0, np.array([6371e3]), np.array([0.5]))


def test_auto_timeshift_synthetics_match_oracle(syn):
    """autoTimeshift end to end: GPU travel times give the reduction times and true sample
    counts of f90/computeSyntheticSeismograms.f90:344-352, which then shift the GPU
    seismograms; the oracle does the same from its own travel times."""
    rng = np.random.default_rng(43)
    nf1, nf2, nsamp = 2, 64, 256
    T = 256.0
    df, dt = 1.0 / T, T / nsamp
    sigma = 5 * df
    nwn_all = np.zeros(nf2, np.int32)
    nwn_all[nf1 - 1:] = np.linspace(40, 120, nf2 - nf1 + 1).astype(np.int32)
    gr, gi = _random_spectra(rng, nwn_all[nf1 - 1:])
    R, rearth, rs = 8, 6371e3, 6371e3 - 35e3
    theta = rng.uniform(0.1, 3.0, R)
    phi = rng.uniform(0, 2 * np.pi, R)
    mt = rng.normal(size=6) * 1e18
    table = oracle.homogeneous_ray_table(rearth, 8000.0, 400)
    tsbuf, twinlen, tlen = 30.0, 120.0, T
    tred, nstrue = syn.auto_timeshift(table, rs, rearth, theta, tsbuf, twinlen, tlen, dt, nsamp)
    tto, _, so = oracle.travel_times(table, rs, np.full(R, rearth), theta)
    assert (so == 0).all()
    np.testing.assert_allclose(tred, tto - tsbuf, rtol=1e-13)
    assert (nstrue == np.minimum((np.minimum(tto + twinlen, tto - tsbuf + tlen) - (tto - tsbuf)) / dt + 1,
                                 nsamp).astype(np.int64)).all()
    assert (np.abs(nstrue - (int(150.0 / dt) + 1)) <= 1).all()          # window of tsbuf + twinlen
    print(f"auto timeshift: {R} receivers, tred {tred.min():.1f}..{tred.max():.1f} s", flush=True)
    ts = syn.synthetics(nf1, nf2, df, sigma, nwn_all, gr, gi, theta, phi, mt, nsamp, dt, tred=tred)
    ots, _ = oracle.synthetics(nf1, nf2, df, sigma, nwn_all, gr, gi, theta, phi, mt, nsamp, dt, tred=tto - tsbuf)
    ts64, ots64 = ts.astype(np.float64), ots.astype(np.float64)
    terr = np.linalg.norm(ts64 - ots64, axis=2) / np.linalg.norm(ots64, axis=2)
    assert terr.max() < 1e-5
