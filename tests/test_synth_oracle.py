"""Pins of the synthetics half of the CPU oracle (Legendre functions, degree
sums, displacement combination, FFT and frequency-time transform) against
closed forms, scipy/numpy and invariants.  No GPU needed."""
import os
import sys

import numpy as np
import pytest
from scipy.special import lpmv

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import oracle  # noqa: E402


@pytest.mark.parametrize("theta", [0.3, 1.1, 2.0, 2.9])
def test_legendre_vs_scipy(theta):
    """P_l^m without the Condon-Shortley phase: (-1)^m scipy.special.lpmv."""
    L = 60
    o = oracle.legendre(L, theta)
    x = np.cos(theta)
    for m in range(3):
        ref = np.array([(-1) ** m * lpmv(m, l, x) if l >= m else 0.0 for l in range(L + 1)])
        scale = np.maximum(np.abs(ref), 1e-300)
        np.testing.assert_allclose(o["plm"][m], ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())


def test_legendre_generating_function_high_degree():
    """sum_l t^l P_l(x) = (1 - 2xt + t^2)^(-1/2) with lmax = 2500 (recurrence stability)."""
    L, t = 2500, 0.99        # t^L ~ 1e-11: truncation below the tolerance
    for theta in [0.01, 0.7, 1.5707, 3.1]:
        x = np.cos(theta)
        p0 = oracle.legendre(L, theta)["plm"][0]
        s = np.sum(p0 * t ** np.arange(L + 1))
        ref = (1 - 2 * x * t + t * t) ** -0.5
        assert abs(s - ref) / abs(ref) < 1e-9, (theta, s, ref)


def test_legendre_theta_derivatives_by_finite_differences():
    L, th, h = 40, 0.9, 1e-5
    o, op, om = oracle.legendre(L, th), oracle.legendre(L, th + h), oracle.legendre(L, th - h)
    fd = (op["plm"] - om["plm"]) / (2 * h)
    np.testing.assert_allclose(o["dpdt"], fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max())
    fd2 = (op["dpdt"] - om["dpdt"]) / (2 * h)
    np.testing.assert_allclose(o["d2pdt2"], fd2, rtol=1e-5, atol=1e-5 * np.abs(fd2).max())
    st = np.sin(th)
    np.testing.assert_allclose(o["pdst"][1:], o["plm"][1:] / st, rtol=1e-14)
    fds = (op["pdst"] - om["pdst"]) / (2 * h)
    np.testing.assert_allclose(o["dtpdst"][1:], fds[1:], rtol=1e-5, atol=1e-5 * np.abs(fds).max())


@pytest.mark.parametrize("pole", [0.0])
def test_legendre_pole_limits(pole):
    """The pole formulas equal the limit theta -> 0 of the regular branch.  (theta = pi in FP64 has
    sin(theta) = 1.2e-16 > 0, so the reference's pole branch is reached at theta = 0 only.)"""
    L = 30
    at = oracle.legendre(L, pole)
    eps = 1e-4 if pole == 0.0 else -1e-4
    n1, n2 = oracle.legendre(L, pole + eps), oracle.legendre(L, pole + 2 * eps)
    for key in ("dpdt", "pdst", "d2pdt2", "dtpdst"):
        a, b = at[key], 2 * n1[key] - n2[key]      # linear extrapolation of the regular branch to the pole
        scale = np.abs(b).max()
        rows = [1] if key in ("pdst", "dtpdst") else [0, 1, 2]
        for m in rows:
            if key == "dtpdst" and m == 1:
                continue        # d/dtheta (P_l1/sin) -> 0 at the pole, the regular branch is O(theta)
            if key == "d2pdt2" and m == 0:
                # legendreFunctions.f90:146 sets +l(l+1)/2 sign(ct)^l; the limit is -l(l+1)/2 sign(ct)^l.
                # The oracle keeps the reference's value (only strain sums, kint >= 15, use it).
                np.testing.assert_allclose(a[m], -b[m], rtol=1e-4, atol=1e-4 * scale)
                continue
            np.testing.assert_allclose(a[m], b[m], rtol=1e-4, atol=1e-4 * scale, err_msg=f"{key} m={m}")
    np.testing.assert_allclose(at["dtpdst"][2], (2 * n1["dtpdst"] - n2["dtpdst"])[2], rtol=1e-4)


def test_taper_closed_form():
    lmax, tf = 1000, 0.2
    w = oracle.taper(lmax, tf)
    jt = int(round(lmax * (1 - tf)))
    assert np.all(w[:jt] == 1.0)
    assert abs(w[jt] - 1.0) < 1e-12            # cos^2(0) at the start of the taper
    assert abs(w[lmax]) < 1e-12                # cos^2(pi/2) at lmax
    assert np.all(np.diff(w[jt:]) <= 1e-15)    # monotone
    assert abs(w[(jt + lmax) // 2] - 0.5) < 1e-2


def test_fft_matches_numpy():
    """The reference's transform takes its twiddle angles from pi = 3.1415927410125732 (an FP32
    literal), so it differs from an exact DFT by ~1e-7 per twiddle and stage: agreement with numpy
    to 1e-5 of the largest value (a wrong sign, ordering or scale would be O(1))."""
    rng = np.random.default_rng(0)
    n = 2 ** 10
    x, y = rng.normal(size=n), rng.normal(size=n)
    for sgn in (+1, -1):
        a, b = oracle.fft(x, y, sgn)
        ref = np.fft.ifft(x + 1j * y) * n if sgn > 0 else np.fft.fft(x + 1j * y)
        err = np.abs(a + 1j * b - ref).max() / np.abs(ref).max()
        assert 1e-9 < err < 1e-5, err        # not exact: the reference's pi is visible


def test_freq_time_matches_irfft():
    """ts = df e^{sigma t} N irfft(X) for the Hermitian spectrum X_j, j = nf1-1..nf2-1."""
    rng = np.random.default_rng(1)
    nf1, nf2, nsamp = 3, 100, 512
    T = 256.0
    df, dt = 1.0 / T, T / nsamp
    sigma = 5 * df
    sp = np.zeros(nf2, np.complex128)
    sp[nf1 - 1:] = rng.normal(size=nf2 - nf1 + 1) + 1j * rng.normal(size=nf2 - nf1 + 1)
    rc, ts = oracle.freq_time(nf1, nf2, df, sigma, sp, nsamp, dt)
    assert rc == 0
    X = np.zeros(nsamp // 2 + 1, np.complex128)
    X[:nf2] = sp
    ref = np.fft.irfft(X, nsamp) * nsamp * df * np.exp(sigma * dt * np.arange(nsamp))
    np.testing.assert_allclose(ts, ref, rtol=0, atol=1e-5 * np.abs(ref).max())
    # 300 frequencies need 1024 samples: dt = 0.3 T / 512 is no multiple of T / 1024
    rc, _ = oracle.freq_time(nf1, 300, df, sigma, np.zeros(300, complex), 512, 0.3 * T / 512)
    assert rc == 2


@pytest.mark.parametrize("ndec,nkeep", [(2, None), (4, None), (1, 300), (2, 100)])
def test_freq_time_decimation(ndec, nkeep):
    """With fewer samples than the frequencies need, transformFrequencyTime transforms at the full
    length ns (step dts) and keeps every ndec-th sample; its damping factor advances by
    exp(sigma dts) per kept sample (frequencyTime.f90:104-109), so the kept samples are those of the
    full transform times exp(-sigma dts (ndec - 1) i)."""
    rng = np.random.default_rng(2)
    nf1, nf2 = 3, 200                               # needs ns = 512
    T = 256.0
    df, sigma = 1.0 / T, 5.0 / T
    sp = np.zeros(nf2, np.complex128)
    sp[nf1 - 1:] = rng.normal(size=nf2 - nf1 + 1) + 1j * rng.normal(size=nf2 - nf1 + 1)
    ns = 512
    dts = T / ns
    rc, full = oracle.freq_time(nf1, nf2, df, sigma, sp, ns, dts)
    assert rc == 0
    nsamp = ns // ndec if nkeep is None else nkeep   # ndec = 1 with nsamp < ns keeps the first nsamp
    rc, dec = oracle.freq_time(nf1, nf2, df, sigma, sp, nsamp, ndec * dts)
    assert rc == 0
    i = np.arange(nsamp)
    ref = full[::ndec][:nsamp].astype(np.float64) * np.exp(-sigma * dts * (ndec - 1) * i)
    np.testing.assert_allclose(dec, ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())
    if ndec > 1:
        # more kept samples than the transform holds: an error, not a read past the transform
        rc, _ = oracle.freq_time(nf1, nf2, df, sigma, sp, ns // ndec + 1, ndec * dts)
        assert rc == 4


def _gen_spectra(nf, L, t, kind):
    """Spectra G_s(l) that turn a degree sum into a generating function: G = t^l 4 pi/(2l+1) (x l(l+1)/2)."""
    l = np.arange(L)
    g = t ** l * 4 * float(np.float32(np.pi)) / (2 * l + 1)
    if kind == "m1":
        g = g * l * (l + 1) / 2.0
    return g


@pytest.mark.parametrize("theta", [0.4, 1.3, 2.6])
def test_moment_sums_generating_functions(theta):
    """sum_l P_l t^l, sum_l dP_l/dtheta t^l, sum_l P_l^1 t^l in closed form (taper region negligible)."""
    L, t = 240, 0.5
    pi32 = float(np.float32(np.pi))        # the sums' 4 pi uses the reference's pi
    nwn = np.array([L], np.int32)
    nsp = 10
    gr = np.zeros((nsp, L), np.float32)
    gi = np.zeros((nsp, L), np.float32)
    g0 = _gen_spectra(1, L, t, "m0")
    g1 = _gen_spectra(1, L, t, "m1")
    gr[0] = g0          # kint 1: P_l0 x U jump 1
    gi[4] = g0          # kint 5: dP_l0/dtheta x V jump 1 (imaginary part)
    gr[2] = g1          # kint 3: P_l1 2/(l(l+1)) x U jump 3
    h = oracle.moment_sums(1, 1, nwn, gr, gi, [theta])[0, 0]
    x, s = np.cos(theta), np.sin(theta)
    q = 1 - 2 * x * t + t * t
    # FP32 storage of the spectra limits the agreement to ~1e-7
    assert abs(h[0] - q ** -0.5) < 2e-7 * q ** -0.5
    assert abs(h[4] - 1j * (-t * s * q ** -1.5)) < 2e-7 * abs(t * s * q ** -1.5)
    assert abs(h[2] - s * t * q ** -1.5) < 2e-7 * abs(s * t * q ** -1.5)
    assert abs(h[1]) == 0.0 and abs(h[3]) == 0.0     # spectra 1 and 3 are zero


def test_displacement_isotropic_source_is_azimuth_independent():
    """An explosion (Mrr = Mtt = Mff) has no transverse motion and no azimuth dependence."""
    rng = np.random.default_rng(2)
    L = 50
    nsp = 10
    gr = rng.normal(size=(nsp, L)).astype(np.float32)
    gi = rng.normal(size=(nsp, L)).astype(np.float32)
    nwn = np.array([L], np.int32)
    mt = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0])
    _, zsp = oracle.synthetics(1, 1, 1.0 / 64, 0.0, nwn, gr, gi, [0.8] * 3, [0.1, 1.7, 4.0], mt, 4, 16.0)
    z = zsp[:, :, 0]
    assert np.all(np.abs(z[:, 2]) < 1e-12 * np.abs(z).max())
    np.testing.assert_allclose(z[1], z[0], rtol=1e-13)
    np.testing.assert_allclose(z[2], z[0], rtol=1e-13)


def test_bessel_functions_vs_scipy():
    from scipy.special import j0, j1, y0, y1
    for x in np.concatenate([np.linspace(0.05, 40.0, 97), [123.4, 2404.8, 9.9e4]]):
        assert abs(oracle.besselj(0, x) - j0(x)) < 1e-14
        assert abs(oracle.besselj(1, x) - j1(x)) < 1e-14
        assert abs(oracle.bessely(0, x) - y0(x)) < 1e-14 * max(1.0, abs(y0(x)))
        assert abs(oracle.bessely(1, x) - y1(x)) < 1e-14 * max(1.0, abs(y1(x)))


def _wn_closed_form(dwn, x, kint, a=1000.0, kmax=0.04):
    """Wavenumber integral of G(k) = exp(-a k) 2 pi / R^2 against the kint kernel, as computed by the
    oracle, and its closed form:  kint 1: int k J0(kx) e^{-ak} dk = a / (a^2 + x^2)^{3/2};
    kint 9: (2/x) int J1(kx) e^{-ak} dk = (2/x) (1 - a / sqrt(a^2 + x^2)) / x."""
    pi32, re = float(np.float32(np.pi)), 6371000.0
    nwn = int(round(kmax / dwn))
    k = np.arange(nwn) * dwn
    gr = np.zeros((10, nwn), np.float32)
    gi = np.zeros((10, nwn), np.float32)
    spec = {1: 0, 9: 8}[kint]
    scale = 1e20                                  # keep the FP32 spectra in range
    gr[spec] = (np.exp(-a * k) * 2 * pi32 / re ** 2 * scale).astype(np.float32)
    h = oracle.bessel_sums(1, 1, np.array([nwn], np.int32), gr, gi, dwn, re, [x])[0, 0, kint - 1].real / scale
    ref = a / (a * a + x * x) ** 1.5 if kint == 1 else 2 / x * (1 - a / np.sqrt(a * a + x * x)) / x
    return h, ref


@pytest.mark.parametrize("kint", [1, 9])
def test_wavenumber_integrals_converge_to_closed_form(kint):
    """Trapezoid below kx = 3 and Filon with Neumann end terms above: second-order convergence in dwn
    towards the closed-form integral (a wrong weight would stall or change the limit)."""
    x = 5000.0
    errs = []
    for dwn in (2e-5, 1e-5, 5e-6):
        h, ref = _wn_closed_form(dwn, x, kint)
        errs.append(abs(h - ref) / abs(ref))
    assert errs[2] < 1e-3, errs
    for a, b in zip(errs, errs[1:]):
        assert 3.5 < a / b < 4.5, errs          # O(dwn^2)



# ---- single-force source ----

@pytest.mark.parametrize("theta", [0.4, 1.3, 2.6])
def test_force_sums_generating_functions(theta):
    """Each of the 7 force sums against a closed-form generating function.  Spectrum s carries its own
    ratio t_s, so a sum that reads the wrong spectrum, basis row or sign misses its closed form:
      kint 1 (U1): sum P_l t^l = D^-1/2;  kint 2 (U2): sum P_l^1 t^l = s t D^-3/2;
      kint 3 (V1): sum dP_l/dtheta t^l = -s t D^-3/2;
      kint 4, 7 (V2): -sum dP_l^1/dtheta t^l, +sum P_l^1/sin t^l with d/dtheta (s t D^-3/2) = c t D^-3/2 - 3 t^2 s^2 D^-5/2;
      kint 5, 6 (W1): -sum P_l^1/sin t^l = -t D^-3/2 and +sum dP_l^1/dtheta t^l,
    with D = 1 - 2 t cos(theta) + t^2 (P_l^m without the Condon-Shortley sign)."""
    L = 240
    ts = [0.3, 0.4, 0.5, 0.55, 0.6]
    nwn = np.array([L], np.int32)
    gr = np.zeros((5, L), np.float32)
    gi = np.zeros((5, L), np.float32)
    for s_, t in enumerate(ts):
        g = _gen_spectra(1, L, t, "m0" if s_ in (0, 2) else "m1")
        if s_ in (1, 3):
            g = -g                                   # the -2/(l(l+1)) of kint 2 and 4 cancels
        (gi if s_ == 3 else gr)[s_] = g              # V2 in the imaginary part
    h = oracle.force_sums(1, 1, nwn, gr, gi, [theta])[0, 0]
    c, s = np.cos(theta), np.sin(theta)
    D = lambda t: 1 - 2 * t * c + t * t              # noqa: E731
    dP1 = lambda t: c * t * D(t) ** -1.5 - 3 * t * t * s * s * D(t) ** -2.5     # noqa: E731
    t0, t1, t2, t3, t4 = ts
    ref = [D(t0) ** -0.5, s * t1 * D(t1) ** -1.5, -s * t2 * D(t2) ** -1.5, 1j * dP1(t3),
           -t4 * D(t4) ** -1.5, dP1(t4), 1j * (-t3 * D(t3) ** -1.5)]
    for k in range(7):
        assert abs(h[k] - ref[k]) < 1e-6 * max(abs(ref[k]), 1.0), (k + 1, h[k], ref[k])


def _force_case(rng, L=60):
    gr = rng.normal(size=(5, L)).astype(np.float32)
    gi = rng.normal(size=(5, L)).astype(np.float32)
    return np.array([L], np.int32), gr, gi


def test_force_displacement_rotation_invariance():
    """Turning the horizontal force and the receiver azimuth by the same angle leaves the
    (r, theta, phi) displacement unchanged; a vertical force moves nothing transversely and does
    not depend on azimuth."""
    rng = np.random.default_rng(5)
    nwn, gr, gi = _force_case(rng)
    th = [0.7] * 4
    phi = np.array([0.2, 1.1, 2.9, 5.0])
    f = np.array([0.3, 1.2, -0.7])
    _, z0 = oracle.synthetics(1, 1, 1.0 / 64, 0.0, nwn, gr, gi, th, phi, f, 4, 16.0)
    for a in (0.4, 2.2):
        fr = np.array([f[0], f[1] * np.cos(a) - f[2] * np.sin(a), f[1] * np.sin(a) + f[2] * np.cos(a)])
        _, z1 = oracle.synthetics(1, 1, 1.0 / 64, 0.0, nwn, gr, gi, th, phi + a, fr, 4, 16.0)
        np.testing.assert_allclose(z1[:, :, 0], z0[:, :, 0], rtol=1e-12, atol=1e-12 * np.abs(z0).max())
    _, zv = oracle.synthetics(1, 1, 1.0 / 64, 0.0, nwn, gr, gi, th, phi, [1.0, 0.0, 0.0], 4, 16.0)
    zv = zv[:, :, 0]
    assert np.all(np.abs(zv[:, 2]) == 0.0)
    np.testing.assert_allclose(zv, np.broadcast_to(zv[0], zv.shape), rtol=1e-13)
    # the horizontal force's radial motion is linear in (Ft cos phi + Ff sin phi): it vanishes
    # for the azimuth perpendicular to the force
    ph0 = np.arctan2(f[1], -f[2])
    _, zp = oracle.synthetics(1, 1, 1.0 / 64, 0.0, nwn, gr, gi, [0.7], [ph0], [0.0, f[1], f[2]], 4, 16.0)
    assert abs(zp[0, 0, 0]) < 1e-13 * np.abs(z0).max() and abs(zp[0, 1, 0]) < 1e-13 * np.abs(z0).max()


def _force_wn_closed_form(dwn, x, kint, a=1000.0, kmax=0.04):
    """Force wavenumber integral of G(k) = exp(-a k) 2 pi / R^2 with closed forms of Laplace transforms:
    kint 1: int k J0 e^{-ak} = a / (a^2+x^2)^{3/2};  kint 2: -(2/R) int J1 e^{-ak} = -(2/R)(1 - a/r)/x;
    kint 3: -R int k^2 J1 e^{-ak} = -R 3 a x / r^5;  kint 5, 7: -+(2/x) int J1 e^{-ak} = -+(2/x)(1 - a/r)/x;
    kint 4, 6: -+2 int k (J0 - J1/(kx)) e^{-ak} = -+2 (a / r^3 - (1 - a/r) / x^2);  r = sqrt(a^2 + x^2)."""
    pi32, re = float(np.float32(np.pi)), 6371000.0
    nwn = int(round(kmax / dwn))
    k = np.arange(nwn) * dwn
    gr = np.zeros((5, nwn), np.float32)
    gi = np.zeros((5, nwn), np.float32)
    spec = {1: 0, 2: 1, 3: 2, 4: 3, 5: 4, 6: 4, 7: 3}[kint]
    scale = {2: 1e26, 3: 1e14}.get(kint, 1e20)               # keep the FP32 spectra and results in range
    gr[spec] = (np.exp(-a * k) * 2 * pi32 / re ** 2 * scale).astype(np.float32)
    h = oracle.bessel_sums(1, 1, np.array([nwn], np.int32), gr, gi, dwn, re, [x], istyp=0)[0, 0, kint - 1].real / scale
    r = np.sqrt(a * a + x * x)
    r4 = 2 * (a / r ** 3 - (1 - a / r) / x ** 2)
    ref = {1: a / r ** 3, 2: -2 / re * (1 - a / r) / x, 3: -re * 3 * a * x / r ** 5, 4: -r4,
           5: -2 / x * (1 - a / r) / x, 6: r4, 7: 2 / x * (1 - a / r) / x}[kint]
    return h, ref


@pytest.mark.parametrize("kint", [1, 2, 3, 4, 5, 6, 7])
def test_force_wavenumber_integrals_converge_to_closed_form(kint):
    x = 5000.0
    errs = []
    for dwn in (2e-5, 1e-5, 5e-6):
        h, ref = _force_wn_closed_form(dwn, x, kint)
        errs.append(abs(h - ref) / abs(ref))
    assert errs[2] < 1e-3, errs
    for a, b in zip(errs, errs[1:]):
        assert 3.5 < a / b < 4.5, errs          # O(dwn^2)


# ---- event and station filters ----

@pytest.mark.parametrize("n1,f1,n2,f2", [(2, 0.01, 4, 0.08), (1, 0.02, 3, 0.05), (0, 0.01, 5, 0.1), (3, 0.005, 0, 1.0)])
def test_butterworth_bandpass_vs_scipy_zpk(n1, f1, n2, f2):
    """The event filter equals scipy's analog Butterworth prototypes H(s) evaluated at
    s = i (f - i sigma / 2 pi) / fc, the complex frequencies of the damped spectrum (order 0: unity)."""
    from scipy import signal
    df, sigma, nf1, nf2 = 1.0 / 512, 5.0 / 512, 3, 200
    pi32 = float(np.float32(np.pi))
    h = oracle.butterworth_bandpass([n1, f1, n2, f2], nf1, nf2, df, sigma)
    assert np.all(h[:nf1 - 1] == 0)
    f = (np.arange(nf1, nf2 + 1) - 1) * df - 1j * sigma / (2 * pi32)

    def analog(n, fc, btype):
        if n == 0:
            return np.ones_like(f)
        z, p, k = signal.butter(n, 1.0, btype=btype, analog=True, output="zpk")
        s = 1j * f / fc
        num = np.prod([s - zz for zz in z], axis=0) if len(z) else np.ones_like(s)
        return k * num / np.prod([s - pp for pp in p], axis=0)

    ref = analog(n1, f1, "high") * analog(n2, f2, "low")
    # the reference's unit roots use its FP32 pi, scipy's the exact one: ~1e-7 apart near the corners
    np.testing.assert_allclose(h[nf1 - 1:], ref, rtol=1e-6, atol=1e-9)


def test_station_filter_equals_time_shift():
    """A station filter e^{i omega tau} e^{sigma tau} gives the seismogram of the time shift tau, and a
    constant event filter scales it (the filters multiply the spectrum as the time shift does)."""
    rng = np.random.default_rng(6)
    L, nf1, nf2, nsamp = 40, 2, 30, 64
    df, dt = 1.0 / 64, 1.0
    sigma = 5 * df
    nwn = np.zeros(nf2, np.int32)
    nwn[nf1 - 1:] = L
    gr = rng.normal(size=(10, L * (nf2 - nf1 + 1))).astype(np.float32)
    gi = rng.normal(size=(10, L * (nf2 - nf1 + 1))).astype(np.float32)
    th, ph = [0.5, 1.9], [0.3, 4.0]
    mt = rng.normal(size=6)
    tau = np.array([3.0, 7.5])
    pi32 = float(np.float32(np.pi))
    om = 2 * pi32 * (np.arange(nf2)) * df
    stf = np.exp(1j * om[None, None, :] * tau[:, None, None]) * np.exp(sigma * tau)[:, None, None] * np.ones((2, 3, 1))
    ts_a, z_a = oracle.synthetics(nf1, nf2, df, sigma, nwn, gr, gi, th, ph, mt, nsamp, dt, tred=tau)
    ts_b, z_b = oracle.synthetics(nf1, nf2, df, sigma, nwn, gr, gi, th, ph, mt, nsamp, dt, stafilter=stf,
                                  evfilter=np.full(nf2, 2.0 - 0.5j))
    np.testing.assert_allclose(z_b[:, :, nf1 - 1:], (2.0 - 0.5j) * z_a[:, :, nf1 - 1:], rtol=1e-12,
                               atol=1e-12 * np.abs(z_a).max())


def _rotate_mt(mt, a):
    """Moment tensor (Mrr, Mtt, Mff, Mrt, Mrf, Mtf) turned by a about the vertical: (Mrt, Mrf) turns as
    a horizontal vector, the (theta, phi) block as a 2-D tensor."""
    c, s = np.cos(a), np.sin(a)
    R = np.array([[c, -s], [s, c]])
    H = R @ np.array([[mt[1], mt[5]], [mt[5], mt[2]]]) @ R.T
    v = R @ np.array([mt[3], mt[4]])
    return np.array([mt[0], H[0, 0], H[1, 1], v[0], v[1], H[0, 1]])


def test_moment_displacement_rotation_invariance():
    """Turning the moment tensor and the receiver azimuth by the same angle leaves the (r, theta, phi)
    displacement unchanged; this pins every azimuthal term of the moment combination."""
    rng = np.random.default_rng(7)
    L = 60
    gr = rng.normal(size=(10, L)).astype(np.float32)
    gi = rng.normal(size=(10, L)).astype(np.float32)
    nwn = np.array([L], np.int32)
    th = [0.9] * 4
    phi = np.array([0.1, 1.4, 3.3, 5.5])
    mt = rng.normal(size=6)
    _, z0 = oracle.synthetics(1, 1, 1.0 / 64, 0.0, nwn, gr, gi, th, phi, mt, 4, 16.0)
    for a in (0.3, 2.0, 4.4):
        _, z1 = oracle.synthetics(1, 1, 1.0 / 64, 0.0, nwn, gr, gi, th, phi + a, _rotate_mt(mt, a), 4, 16.0)
        np.testing.assert_allclose(z1[:, :, 0], z0[:, :, 0], rtol=1e-12, atol=1e-12 * np.abs(z0).max())


def _small_source(rng, L=60, nf1=2, nf2=24):
    """Random FP32 spectra of a moment source at frequencies nf1..nf2, L degrees each."""
    nwn = np.zeros(nf2, np.int32)
    nwn[nf1 - 1:] = L
    nkf = int(nwn.sum())
    gr = (rng.normal(size=(10, nkf)) * np.exp(-np.tile(np.arange(L), nf2 - nf1 + 1) / 20.0)).astype(np.float32)
    gi = (rng.normal(size=(10, nkf)) * np.exp(-np.tile(np.arange(L), nf2 - nf1 + 1) / 20.0)).astype(np.float32)
    return nwn, gr, gi


def test_zne_rotation_follows_the_propagation_direction():
    """vectorDcmplxZNEfromRLTAxesRotation (axesRotation.f90:560-568) with propdir measured at the
    receiver from south over east (computeSyntheticSeismograms.f90:315-317): built here from the
    geometry instead, the longitudinal unit vector is L = (E, N) = (sin a, -cos a) (a = 0 points
    south) and the transverse one T = up x L.  Z = R, (N, E) = ul L + ut T for random directions,
    the horizontal norm is kept, a = 0 gives (N, E) = (-L, T) and a = pi/2 gives (T, L)."""
    rng = np.random.default_rng(21)
    nwn, gr, gi = _small_source(rng)
    R = 5
    theta, phi = rng.uniform(0.3, 2.8, R), rng.uniform(0, 2 * np.pi, R)
    mt = rng.normal(size=6)
    T = 64.0
    args = (2, 24, 1.0 / T, 5.0 / T, nwn, gr, gi, theta, phi, mt, 64, T / 64)
    _, rlt = oracle.synthetics(*args)
    for a in (rng.uniform(0, 2 * np.pi, R), np.zeros(R), np.full(R, np.pi / 2)):
        _, zne = oracle.synthetics(*args, propdir=a)
        for r in range(R):
            Lv = np.array([np.sin(a[r]), -np.cos(a[r]), 0.0])             # (E, N, U)
            Tv = np.cross([0.0, 0.0, 1.0], Lv)
            ul, ut = rlt[r, 1], rlt[r, 2]
            np.testing.assert_allclose(zne[r, 0], rlt[r, 0], rtol=1e-14, atol=0)
            np.testing.assert_allclose(zne[r, 1], ul * Lv[1] + ut * Tv[1], rtol=1e-12, atol=1e-14 * np.abs(ul).max())
            np.testing.assert_allclose(zne[r, 2], ul * Lv[0] + ut * Tv[0], rtol=1e-12, atol=1e-14 * np.abs(ul).max())
            h0 = np.abs(ul) ** 2 + np.abs(ut) ** 2
            h1 = np.abs(zne[r, 1]) ** 2 + np.abs(zne[r, 2]) ** 2
            np.testing.assert_allclose(h1, h0, rtol=1e-12, atol=1e-24 * h0.max())
    _, z0 = oracle.synthetics(*args, propdir=np.zeros(R))
    np.testing.assert_allclose(z0[:, 1], -rlt[:, 1], rtol=1e-14)
    np.testing.assert_allclose(z0[:, 2], rlt[:, 2], rtol=1e-14)


def test_velocity_is_the_time_derivative_of_displacement():
    """The velocity branch multiplies by i (omega - i sigma) (computeSyntheticSeismograms.f90:376-378),
    which after the transform and the exp(sigma t) undamping is exactly d/dt of the displacement
    trace.  Checked in the time domain by central differences of densely sampled displacement traces:
    the residual falls at second order as dt halves (a real omega, or the damping term with the wrong
    sign, leaves an O(sigma/omega) error that does not shrink)."""
    rng = np.random.default_rng(22)
    nwn, gr, gi = _small_source(rng)
    theta, phi, mt = np.array([0.7, 2.1]), np.array([0.4, 3.9]), rng.normal(size=6)
    T = 64.0
    errs = []
    for nsamp in (512, 1024, 2048):
        dt = T / nsamp
        args = (2, 24, 1.0 / T, 5.0 / T, nwn, gr, gi, theta, phi, mt, nsamp, dt)
        u, _ = oracle.synthetics(*args)
        v, _ = oracle.synthetics(*args, velocity=True)
        u, v = u.astype(np.float64), v.astype(np.float64)
        du = (u[:, :, 2:] - u[:, :, :-2]) / (2 * dt)
        vv = v[:, :, 1:-1]
        # the first and last 10 % of the window hold the wrap-around of the periodic transform
        k0, k1 = nsamp // 10, nsamp - nsamp // 10
        errs.append(np.linalg.norm(du[:, :, k0:k1] - vv[:, :, k0:k1]) / np.linalg.norm(vv[:, :, k0:k1]))
    print(f"velocity vs central difference of displacement: {errs}")
    assert errs[-1] < 3e-3
    assert errs[0] / errs[1] > 3.0 and errs[1] / errs[2] > 3.0


# ---- travel times from a ray table (f90/travelTimes.f90) ----
# In a homogeneous sphere rays are straight: the first arrival between radii
# rs and re at angular distance d takes chord / v with
# chord^2 = rs^2 + re^2 - 2 rs re cos d, and its slowness is dT/dd =
# rs re sin d / (v chord).  The table method interpolates linearly in the
# turning-point index, so its error must fall at second order as the table is
# refined; an index or sign slip in the interpolation leaves an O(1) error.
_R, _V = 6371e3, 8000.0


def _chord(rs, re, d):
    return np.sqrt(rs * rs + re * re - 2 * rs * re * np.cos(d)) / _V


@pytest.mark.parametrize("rs,re", [(6000e3, 6371e3), (6371e3, 6000e3), (5800e3, 6300e3)])
def test_travel_times_turning_rays_converge_to_chord(rs, re):
    dmin = np.arccos(min(rs, re) / max(rs, re))      # the ray grazing the shallower radius
    d = np.linspace(dmin + 0.15, 2.0, 400)
    errs = []
    for n in (400, 800):
        tt, p, st = oracle.travel_times(oracle.homogeneous_ray_table(_R, _V, n), rs, np.full(d.size, re), d)
        assert (st == 0).all()
        ex = _chord(rs, re, d)
        pe = rs * re * np.sin(d) / (_V * _V * ex)
        errs.append(np.max(np.abs(tt - ex) / ex))
        assert np.max(np.abs(p - pe) / pe) < 5e-4
    assert errs[0] < 3e-5 and errs[1] < 6e-6
    assert errs[0] / errs[1] > 2.5                    # second order


def test_travel_times_steep_rays():
    """Below the smallest turning distance the reference takes the steep ray,
    distance and time of the source leg minus those of the receiver leg
    (f90/travelTimes.f90:300-354).  That is the geometric ray for a source above
    the receiver; for a source below it the sign makes both negative, and the
    restatement keeps the reference's result (an O(1) departure from the chord)."""
    rs, re = 6371e3, 6000e3
    d = np.linspace(0.03, np.arccos(re / rs) - 0.05, 200)
    errs = []
    for n in (400, 800):
        tt, _, st = oracle.travel_times(oracle.homogeneous_ray_table(_R, _V, n), rs, np.full(d.size, re), d)
        assert (st == 0).all()
        errs.append(np.max(np.abs(tt - _chord(rs, re, d)) / _chord(rs, re, d)))
    assert errs[0] < 5e-5 and errs[0] / errs[1] > 2.5
    tt, _, st = oracle.travel_times(oracle.homogeneous_ray_table(_R, _V, 400), re, np.full(d.size, rs), d)
    assert (st == 0).all() and np.max(np.abs(tt - _chord(rs, re, d)) / _chord(rs, re, d)) > 0.5


def test_travel_times_beyond_deepest_turning_point():
    """Past the largest tabulated distance the arrival travels along the deepest
    turning point: continuous there, with that ray's slowness as slope."""
    T = oracle.homogeneous_ray_table(_R, _V, 400, rmin_frac=0.55)
    rs, re = 6100e3, 6371e3
    d = np.linspace(1.5, 3.1, 500)
    tt, p, st = oracle.travel_times(T, rs, np.full(d.size, re), d)
    assert (st == 0).all()
    far = p == np.float32(T["slowness"][0])
    assert far.any() and (~far).any()
    k = np.argmax(far)
    assert abs(tt[k] - tt[k - 1]) < 2 * p[k] * (d[k] - d[k - 1])         # no jump
    np.testing.assert_allclose(np.diff(tt[far]), p[far][1:] * np.diff(d[far]), rtol=1e-9)


def test_travel_times_errors():
    T = oracle.homogeneous_ray_table(_R, _V, 200)
    d = np.array([0.5])
    for rs, re, code in ((0.2 * _R, _R, 1), (1.01 * _R, _R, 2), (_R, 0.2 * _R, 4), (_R, 1.01 * _R, 6)):
        tt, _, st = oracle.travel_times(T, rs, np.array([re]), d)
        assert st[0] == code and tt[0] == -1.0
    # within 1e-5 r of the top radius the table is extrapolated
    assert oracle.travel_times(T, _R * (1 + 5e-6), np.array([_R]), d)[2][0] == 0
    # teleseismic tables have no steep rays
    tt, _, st = oracle.travel_times(dict(T, istele=True), 6000e3, np.array([6371e3]), np.array([0.05]))
    assert st[0] == 8
