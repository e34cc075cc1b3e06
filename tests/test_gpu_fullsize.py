"""The C2 minors workload at full size on the GPU: 819 frequencies x l = 0..2500
(2,048,319 (omega, l) solves) at 16 mantle nodes 0-700 km deep.  The oracle
would need hours for this, so the run is checked through invariants of the
radial systems that hold for every instance whatever the grid size:

- every instance converges (status 0);
- toroidal: trace A = -2/r (toroidalOdeSystem.f90:97-100), so r^2 (u1 d2 - d1 u2)
  is the same at every solid node;
- spheroidal: the adjoint system's trace is +4/r (solidtOdeSystem.f90:97-112),
  so r^4 det is the same at every solid node;
- the minors of both legs satisfy the Pluecker relation m1 m6 - m2 m5 + m3 m4 = 0
  (m6 = -m1/dll1, spheroidalMinors.f90:137-142,299);
- a second run of one frequency block is bitwise identical: work claims change the
  order in which lanes run, never what a lane computes.

Bounds: the step control keeps each step's relative error below eps; over the
10^2-10^3 steps of a leg, and with the cancellation of two legs in the
determinant of evanescent pairs, the invariants hold to about 10^3 eps for
nearly all instances and 10^4 eps for the worst.  A wrong node radius,
normalisation or component breaks them by O(1)."""
import numpy as np
import pytest

from gemini_unified_amd import ExternalRadialNodes, FrequencyGrid, GreenFKEngine, global_model

pytestmark = pytest.mark.gpu

EPS = 1e-8


def test_c2_full_grid_invariants():
    nem = global_model(1, 200)
    nodes = ExternalRadialNodes.from_blocks(nem, [15], [700000.0 / 15])
    grid = FrequencyGrid.global_fixed_lmax(4096.0, 0.2, 2500)
    assert grid.freqs.size == 819 and int(grid.nwn.sum()) == 2_048_319
    eng = GreenFKEngine(nem, nodes, device=0, variant="fast")
    r = nodes.rnod
    F = grid.freqs
    dev_t, dev_s, plk = [], [], []
    first = None
    nsolve = 0
    for i, c in enumerate(np.array_split(np.arange(F.size), 7)):
        print(f"C2 block {i}: frequencies {c[0]}..{c[-1]}", flush=True)
        g = eng.minors(F[c], np.zeros(c.size, np.int32), grid.nwn[c], grid.sigma, EPS, 1, nodes.nnod)
        assert np.all(g["status"] == 0)
        nsolve += g["status"].size
        if i == 0:
            first = (c, g["u_sph"].copy(), g["det_tor"].copy())
        l = np.concatenate([np.arange(n) for n in grid.nwn[c]])
        sel = l >= 1                                   # l = 0: no toroidal motion, radial system
        for det, p, acc in ((g["det_tor"], 2, dev_t), (g["det_sph"], 4, dev_s)):
            d = det[:, sel]
            w = (r[:, None] ** p) * d
            ok = np.abs(d) > 0                         # nodes below the start radius hold zeros
            live = ok[-1]                              # start radius below the receiver node
            assert not ok[:, ~live].any()              # otherwise no node is reached
            ref = np.where(live, w[-1], 1.0)
            acc.append(np.where(ok, np.abs(w / ref - 1), 0).max(axis=0))
        for m in (g["u_sph"], g["d_sph"]):
            m = m[:, :, sel]
            q = m[:, 0] * m[:, 5] - m[:, 1] * m[:, 4] + m[:, 2] * m[:, 3]
            sc = np.abs(m).max(axis=1) ** 2
            okm = sc > 0
            plk.append(np.where(okm, np.abs(q) / np.where(okm, sc, 1), 0).max(axis=0))
    assert nsolve == 2_048_319
    for name, a in (("toroidal r^2 det", dev_t), ("spheroidal r^4 det", dev_s)):
        a = np.concatenate(a)
        print(f"{name}: max {a.max():.2e}, p99 {np.quantile(a, 0.99):.2e}", flush=True)
        assert a.max() < 2e4 * EPS and np.quantile(a, 0.99) < 1e3 * EPS
    plk = np.concatenate(plk)
    print(f"Pluecker relation: max {plk.max():.2e}", flush=True)
    assert plk.max() < 1e-10
    c, u0, t0 = first
    g = eng.minors(F[c], np.zeros(c.size, np.int32), grid.nwn[c], grid.sigma, EPS, 1, nodes.nnod)
    assert np.array_equal(g["u_sph"], u0, equal_nan=True) and np.array_equal(g["det_tor"], t0, equal_nan=True)
    eng.close()


def _rotate_mt(mt, a):
    c, s = np.cos(a), np.sin(a)
    R = np.array([[c, -s], [s, c]])
    H = R @ np.array([[mt[1], mt[5]], [mt[5], mt[2]]]) @ R.T
    v = R @ np.array([mt[3], mt[4]])
    return np.array([mt[0], H[0, 0], H[1, 1], v[0], v[1], H[0, 1]])


def test_c4_synthetics_invariants():
    """The C4 synthetics at full size (4096 receivers, 819 frequencies x l = 0..2500, 12,288 traces of
    8192 samples) with seeded FK-like spectra:
    - turning the moment tensor and every receiver azimuth by one angle leaves the (r, theta, phi)
      spectra unchanged (every azimuthal term of the moment combination);
    - the spectra are linear in the moment tensor;
    - an explosion (Mrr = Mtt = Mff) moves nothing transversely.
    The bounds are those of the split-bf16 contraction, whose sums agree with the FP64 oracle to
    about 1e-7 of the largest sum."""
    from gemini_unified_amd.synthetics import SyntheticsEngine
    rng = np.random.default_rng(44)
    T, fmax, lmax, nsamp = 4096.0, 0.2, 2500, 8192
    df = 1.0 / T
    nf1, nf2 = 2, int(round(fmax / df)) + 1
    nwn = np.zeros(nf2, np.int32)
    nwn[nf1 - 1:] = lmax + 1
    nkf = int(nwn.sum())
    l = np.tile(np.arange(lmax + 1), nf2 - nf1 + 1)
    env = (np.exp(-l / 300.0) + 0.02).astype(np.float32)
    gr = rng.standard_normal((10, nkf), dtype=np.float32) * env
    gi = rng.standard_normal((10, nkf), dtype=np.float32) * env
    R = 4096
    theta = np.deg2rad(rng.uniform(0.5, 179.5, R))
    phi = np.deg2rad(rng.uniform(0.0, 360.0, R))
    mt = rng.normal(size=6) * 1e18
    syn = SyntheticsEngine(0)
    args = (nf1, nf2, df, 5 * df, nwn, gr, gi, theta)
    print(f"C4 synthetics: {R} receivers, {nf2 - nf1 + 1} frequencies, {lmax + 1} degrees", flush=True)
    _, z0 = syn.synthetics(*args, phi, mt, nsamp, T / nsamp, want_spectra=True)
    _, z1 = syn.synthetics(*args, phi + 0.7, _rotate_mt(mt, 0.7), nsamp, T / nsamp, want_spectra=True)
    scale = np.abs(z0).max(axis=2, keepdims=True)
    rot = (np.abs(z1 - z0) / scale).max()
    print(f"rotation invariance: max {rot:.2e} of each trace's largest value", flush=True)
    assert rot < 1e-5
    m2 = rng.normal(size=6) * 1e18
    _, za = syn.synthetics(*args, phi, m2, nsamp, T / nsamp, want_spectra=True)
    _, zb = syn.synthetics(*args, phi, mt + m2, nsamp, T / nsamp, want_spectra=True)
    lin = (np.abs(zb - (z0 + za)) / np.abs(zb).max(axis=2, keepdims=True)).max()
    print(f"linearity: max {lin:.2e}", flush=True)
    assert lin < 1e-10
    _, ze = syn.synthetics(*args, phi, np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0]) * 1e18, nsamp, T / nsamp,
                           want_spectra=True)
    tr = (np.abs(ze[:, 2]).max(axis=1) / np.abs(ze).max(axis=(1, 2))).max()
    print(f"explosion: transverse / largest {tr:.2e}", flush=True)
    assert tr < 1e-12
    syn.close()
