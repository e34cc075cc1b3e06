"""Pins of the oracle's strain and stress spectra (harmonicDegreeSums.f90:95-351 for every
kint, singleFrequencyDisplacement.f90:35-278), on seeded random spectra of one node with
radial derivatives (nsp = 20 moment / 10 force, the multi-source driver's layout).

- The radial-derivative sums (moment kint 15..28, force 8..14) are the displacement sums
  taken on the derivative spectra, and the displacement sums (kint 1..14 / 1..7) are those
  of the displacement-only layout, which the generating-function pins fix.
- The theta-derivative sums (kint 29..42 / 15..21) are central differences in theta of the
  displacement sums.
- The strain is the symmetric part of the Cartesian gradient of the displacement field: the
  field from the displacement path at neighbouring (theta, phi), its radial derivative from
  the derivative spectra, differentiated numerically in Cartesian coordinates and rotated
  back to (r, theta, phi).  Nothing of the spherical strain formulas enters the check.
- The stress is Hooke's law of a transversely isotropic medium with a radial symmetry axis
  in Voigt form (C11 = C22 = A, C33 = C, C12 = A - 2N, C13 = C23 = F, C44 = C55 = L, C66 = N).
"""
import numpy as np
import pytest

import oracle

NF1, NF2 = 2, 5
NWN = np.array([0, 40, 45, 52, 60], np.int32)           # nwn[jf-1], jf = 1..5 (jf = 1 unused)
RE = 6371000.0


def spectra(istyp, seed=11):
    """random FP32 spectra [nsp][nkf] decaying in l, with V at l = 0 zero (0/0 in real files)"""
    nsp = 20 if istyp == 1 else 10
    rng = np.random.default_rng(seed)
    nkf = int(NWN[NF1 - 1:NF2].sum())
    l = np.concatenate([np.arange(n) for n in NWN[NF1 - 1:NF2]])
    amp = 1.0 / (1.0 + 0.05 * l)
    gr = (rng.standard_normal((nsp, nkf)) * amp).astype(np.float32)
    gi = (rng.standard_normal((nsp, nkf)) * amp).astype(np.float32)
    return gr, gi


def layout(istyp):
    """spectrum offsets of U, V, UP, VP, W, WP (positionDataGreenFKSpectra, derivflag = 1)"""
    nj, nt = (4, 2) if istyp == 1 else (2, 1)
    U, V, UP, VP = (np.arange(nj) + k * nj for k in range(4))
    W = 4 * nj + np.arange(nt)
    WP = 4 * nj + nt + np.arange(nt)
    return U, V, UP, VP, W, WP


FOMT = {1: np.array([1.0, -0.4, -0.6, 0.7, 0.2, -0.5]), 0: np.array([0.3, -1.0, 0.6])}


@pytest.mark.parametrize("istyp", [1, 0])
def test_displacement_and_radial_sums(istyp):
    gr, gi = spectra(istyp)
    U, V, UP, VP, W, WP = layout(istyp)
    theta = np.array([0.3, 1.2, 2.7])
    h = oracle.strain_spectra(istyp, NF1, NF2, NWN, gr, gi, theta, 0.4, RE, FOMT[istyp])["hdsum"]
    nd = 14 if istyp == 1 else 7
    disp = np.concatenate([U, V, W])
    ref = (oracle.moment_sums if istyp == 1 else oracle.force_sums)(NF1, NF2, NWN, gr[disp], gi[disp], theta)
    np.testing.assert_array_equal(h[..., :nd], ref)
    deriv = np.concatenate([UP, VP, WP])
    refd = (oracle.moment_sums if istyp == 1 else oracle.force_sums)(NF1, NF2, NWN, gr[deriv], gi[deriv], theta)
    np.testing.assert_array_equal(h[..., nd:2 * nd], refd)


@pytest.mark.parametrize("istyp", [1, 0])
def test_theta_derivative_sums_are_central_differences(istyp):
    gr, gi = spectra(istyp)
    nd = 14 if istyp == 1 else 7
    th, dh = 1.1, 1e-5
    f = lambda t: oracle.strain_spectra(istyp, NF1, NF2, NWN, gr, gi, np.array([t]), 0.4, RE, FOMT[istyp])["hdsum"][0]
    h0, hp, hm = f(th), f(th + dh), f(th - dh)
    fd = (hp[:, :nd] - hm[:, :nd]) / (2 * dh)
    got = h0[:, 2 * nd:3 * nd]
    assert np.abs(got - fd).max() < 1e-6 * np.abs(fd).max()


def rotation(theta, phi):
    """columns: r, theta, phi unit vectors in Cartesian coordinates"""
    st, ct, sp, cp = np.sin(theta), np.cos(theta), np.sin(phi), np.cos(phi)
    return np.array([[st * cp, ct * cp, -sp], [st * sp, ct * sp, cp], [ct, -st, 0.0]])


@pytest.mark.parametrize("istyp", [1, 0])
@pytest.mark.parametrize("theta,phi", [(0.7, 0.4), (2.1, 3.9)])
def test_strain_is_the_symmetric_cartesian_gradient(istyp, theta, phi):
    gr, gi = spectra(istyp, seed=5)
    U, V, UP, VP, W, WP = layout(istyp)
    fm = FOMT[istyp]
    run = lambda g1, g2, t, p: oracle.strain_spectra(istyp, NF1, NF2, NWN, g1, g2, np.array([t]), p, RE, fm)
    o = run(gr, gi, theta, phi)
    # the radial derivative of the field: the displacement path on the derivative spectra
    swap = np.arange(gr.shape[0])
    swap[np.concatenate([U, V, W])] = np.concatenate([UP, VP, WP])
    dur = run(gr[swap], gi[swap], theta, phi)["zsp"][0]
    h = 1e-5
    u = lambda t, p: rotation(t, p) @ run(gr, gi, t, p)["zsp"][0]          # Cartesian [3, nf2]
    du_t = (u(theta + h, phi) - u(theta - h, phi)) / (2 * h)
    du_p = (u(theta, phi + h) - u(theta, phi - h)) / (2 * h)
    Rm = rotation(theta, phi)
    du_r = Rm @ dur
    # d/dx_j = e_r,j d/dr + e_theta,j / r d/dtheta + e_phi,j / (r sin theta) d/dphi
    er, et, ep = Rm[:, 0], Rm[:, 1], Rm[:, 2]
    for jf in range(NF1 - 1, NF2):
        G = np.outer(du_r[:, jf], er) + np.outer(du_t[:, jf], et) / RE + np.outer(du_p[:, jf], ep) / (RE * np.sin(theta))
        eps_cart = 0.5 * (G + G.T)
        ref = Rm.T @ eps_cart @ Rm
        got = o["strain"][0, :, :, jf]
        assert np.abs(got - ref).max() < 1e-6 * np.abs(ref).max(), (jf, got, ref)
    # frequencies below nf1 are zero
    assert np.all(o["strain"][0, :, :, :NF1 - 1] == 0)


def test_stress_is_hooke_law_of_a_vti_medium():
    gr, gi = spectra(1, seed=3)
    A, Cc, F, L, N = 2.1e11 + 1e8j, 1.9e11 + 2e8j, 0.8e11 + 1e7j, 0.6e11 + 3e8j, 0.7e11 + 2e8j
    el = np.array([A, Cc, F, L, N, 1.3e11, 0.65e11], np.complex128)
    o = oracle.strain_spectra(1, NF1, NF2, NWN, gr, gi, np.array([0.9]), 1.3, RE, FOMT[1], elcon=el)
    # Voigt order (theta, phi, r) = (1, 2, 3): stiffness of a medium with symmetry axis r
    Cv = np.array([[A, A - 2 * N, F, 0, 0, 0], [A - 2 * N, A, F, 0, 0, 0], [F, F, Cc, 0, 0, 0],
                   [0, 0, 0, L, 0, 0], [0, 0, 0, 0, L, 0], [0, 0, 0, 0, 0, N]], np.complex128)
    for jf in range(NF1 - 1, NF2):
        e = o["strain"][0, :, :, jf]
        t = o["stress"][0, :, :, jf]
        ev = np.array([e[1, 1], e[2, 2], e[0, 0], 2 * e[2, 0], 2 * e[1, 0], 2 * e[1, 2]])
        sv = Cv @ ev
        ref = np.array([[sv[2], sv[4], sv[3]], [sv[4], sv[0], sv[5]], [sv[3], sv[5], sv[1]]])
        assert np.abs(t - ref).max() < 1e-13 * np.abs(ref).max()
        assert np.all(t == t.T) and np.all(e == e.T)


# ---- regional branch: wavenumber integrals kint 1..42 / 1..21 (wavenumberIntegrals.f90:122-477) ----

def _smooth_spectra(istyp, dwn, kmax=0.04):
    """spectra exp(-a_s k) with a rate a_s of their own per spectrum s (a spectrum read from the wrong
    position gives another function), in the units the integrals' R^2 / 2 pi expects"""
    nsp = 20 if istyp == 1 else 10
    n = int(round(kmax / dwn))
    k = np.arange(n) * dwn
    a = 800.0 + 150.0 * np.arange(nsp)
    c = (1.0 + 0.1 * np.arange(nsp)) * 1e20 * 2 * np.pi / RE ** 2
    gr = (np.exp(-a[:, None] * k) * c[:, None]).astype(np.float32)
    gi = (np.exp(-1.3 * a[:, None] * k) * 0.5 * c[:, None]).astype(np.float32)
    return np.array([n], np.int32), gr, gi


@pytest.mark.parametrize("istyp", [1, 0])
def test_regional_displacement_and_radial_integrals(istyp):
    """kint 1..14 / 1..7 are the displacement integrals of the U, V, W spectra and 15..28 / 8..14
    the same integrals of the radial-derivative spectra."""
    dwn = 1e-5
    nwn, gr, gi = _smooth_spectra(istyp, dwn)
    U, V, UP, VP, W, WP = layout(istyp)
    x = np.array([3000.0, 20000.0])
    nd = 14 if istyp == 1 else 7
    h = oracle.strain_spectra(istyp, 1, 1, nwn, gr, gi, x / RE, 0.3, RE, FOMT[istyp], dwn=dwn, rearth=RE)["hdsum"]
    disp, deriv = np.concatenate([U, V, W]), np.concatenate([UP, VP, WP])
    ref = oracle.bessel_sums(1, 1, nwn, gr[disp], gi[disp], dwn, RE, x, istyp=istyp)
    np.testing.assert_allclose(h[..., :nd], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    refd = oracle.bessel_sums(1, 1, nwn, gr[deriv], gi[deriv], dwn, RE, x, istyp=istyp)
    np.testing.assert_allclose(h[..., nd:2 * nd], refd, rtol=1e-12, atol=1e-12 * np.abs(refd).max())


@pytest.mark.parametrize("istyp", [1, 0])
def test_regional_x_derivative_integrals_converge_to_derivatives(istyp):
    """kint 29..42 (moment) and 15..21 (force) are R d/dx of kint 1..14 / 1..7: the difference
    between their quadrature and the central difference in x of the displacement quadrature falls
    at third order in dwn (a wrong kernel, factor or spectrum would leave an O(1) difference)."""
    nd = 14 if istyp == 1 else 7
    x, hx = 5000.0, 0.5
    errs = []
    for dwn in (1e-5, 5e-6, 2.5e-6):
        nwn, gr, gi = _smooth_spectra(istyp, dwn)
        f = lambda xx: oracle.strain_spectra(istyp, 1, 1, nwn, gr, gi, np.array([xx / RE]), 0.3, RE, FOMT[istyp],
                                             dwn=dwn, rearth=RE)["hdsum"][0, 0]
        h0, hp, hm = f(x), f(x + hx), f(x - hx)
        fd = RE * (hp[:nd] - hm[:nd]) / (2 * hx)
        errs.append(np.abs(h0[2 * nd:3 * nd] - fd) / np.abs(fd).max())
    assert errs[2].max() < 2e-3, errs[2]
    big = errs[0] > 1e-8
    assert big.sum() >= nd // 2
    assert np.all(errs[1][big] < errs[0][big] / 4) and np.all(errs[2][big] < errs[1][big] / 4), errs
