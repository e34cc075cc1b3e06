"""Closed-form solutions of the radial elastic problem in a homogeneous medium.

Textbook results (spherical Bessel functions of complex argument from scipy),
independent of the oracle's formulas, used to pin its start values, ODE systems
and adjoint propagation.  Variables follow the reference's systems:

* solid spheroidal, y = (U, R, V, S): displacement U r^ + V grad_1, radial and
  tangential tractions R, S;
* fluid, y = (U, R): R = kappa div u (fluidOdeSystem.f90:78-95);
* toroidal, y = (W, T), T = mu (W' - W/r) (toroidalOdeSystem.f90:81-102);
* minors m1..m6 = M12, M13, M14, M23, M24, M34 of two solid solutions
  (M_ij = x_i y_j - x_j y_i), the ordering of spheroidalMinors.f90:306-323.

In a homogeneous isotropic medium with Lame moduli lam, mu (complex), density
rho and complex frequency om (k_p^2 = om^2 rho/(lam+2mu), k_s^2 = om^2 rho/mu):
  P  (potential f(k_p r)):  U = k_p f',  V = f/r
  SV (potential f(k_s r)):  U = l(l+1) f/r,  V = f/r + k_s f'
  R = (lam+2mu) U' + lam (2U - l(l+1)V)/r,  S = mu (V' - V/r + U/r)
  fluid:  U = k f',  R = -om^2 rho f;   toroidal:  W = f(k_s r)
with f = j_l (regular), h_l^(1) = j_l + i y_l (incoming for exp(+i om t)) or h_l^(2) = j_l - i y_l
(outgoing).
"""
from __future__ import annotations

import numpy as np
import scipy.special as sp


def sbessel(l, z, kind):
    """f, f', f'' of the spherical Bessel function of the given kind at complex z."""
    z = np.asarray(z, dtype=np.complex128)
    if kind == "j":
        f, fp = sp.spherical_jn(l, z), sp.spherical_jn(l, z, derivative=True)
    elif kind == "h1":
        f = sp.spherical_jn(l, z) + 1j * sp.spherical_yn(l, z)
        fp = sp.spherical_jn(l, z, derivative=True) + 1j * sp.spherical_yn(l, z, derivative=True)
    elif kind == "h2":
        f = sp.spherical_jn(l, z) - 1j * sp.spherical_yn(l, z)
        fp = sp.spherical_jn(l, z, derivative=True) - 1j * sp.spherical_yn(l, z, derivative=True)
    else:
        raise ValueError(kind)
    fpp = -2.0 / z * fp - (1.0 - l * (l + 1) / z ** 2) * f
    return f, fp, fpp


def solid_matrix(r, l, om, rho, lam, mu):
    """Textbook 4x4 system of y = (U, R, V, S) in an isotropic solid (Takeuchi and Saito)."""
    L = l * (l + 1.0)
    c = lam + 2 * mu
    A = np.zeros((4, 4), np.complex128)
    A[0] = [-2 * lam / (c * r), 1 / c, L * lam / (c * r), 0]
    z1 = c - lam ** 2 / c - mu
    A[1] = [-om ** 2 * rho + 4 * z1 / r ** 2, 2 * lam / (c * r) - 2 / r, -2 * L * z1 / r ** 2, L / r]
    A[2] = [-1 / r, 0, 1 / r, 1 / mu]
    A[3] = [-2 * z1 / r ** 2, -lam / (c * r), -om ** 2 * rho + (L * z1 + (L - 2) * mu) / r ** 2, -3 / r]
    return A


def ratio_solution(l, z, zeta):
    """f, f', f'' at z of the solution of the spherical Bessel equation whose ratio z f_{l+1}/f_l is
    zeta (f_{l+1} = (l/z) f_l - f_l', so f'/f = (l - zeta)/z), normalised to f = 1."""
    z = complex(z)
    fp = (l - zeta) / z
    fpp = -2.0 / z * fp - (1.0 - l * (l + 1) / z ** 2)
    return 1.0 + 0j, fp, fpp


def solid_solutions(r, l, om, rho, lam, mu, kind="j", zeta=None):
    """P and SV solutions y = (U, R, V, S) at radius r: array [2, 4].  kind names the spherical
    Bessel function; zeta = (zeta_P, zeta_S) instead gives the solutions of prescribed ratio
    z f_{l+1}/f_l at r (ratio_solution)."""
    L = l * (l + 1.0)
    kp = np.sqrt(om ** 2 * rho / (lam + 2 * mu))
    ks = np.sqrt(om ** 2 * rho / mu)
    out = np.zeros((2, 4), np.complex128)
    fs = (lambda z, i: sbessel(l, z, kind)) if zeta is None else (lambda z, i: ratio_solution(l, z, zeta[i]))
    f, fp, fpp = fs(kp * r, 0)
    U, Up = kp * fp, kp ** 2 * fpp
    V, Vp = f / r, kp * fp / r - f / r ** 2
    out[0] = [U, (lam + 2 * mu) * Up + lam * (2 * U - L * V) / r, V, mu * (Vp - V / r + U / r)]
    f, fp, fpp = fs(ks * r, 1)
    U, Up = L * f / r, L * (ks * fp / r - f / r ** 2)
    V, Vp = f / r + ks * fp, ks * fp / r - f / r ** 2 + ks ** 2 * fpp
    out[1] = [U, (lam + 2 * mu) * Up + lam * (2 * U - L * V) / r, V, mu * (Vp - V / r + U / r)]
    return out


def minors(pair):
    """m1..m6 = M12, M13, M14, M23, M24, M34 of two solid solutions pair[2, 4]."""
    x, y = pair
    M = lambda i, j: x[i] * y[j] - x[j] * y[i]
    return np.array([M(0, 1), M(0, 2), M(0, 3), M(1, 2), M(1, 3), M(2, 3)])


def fluid_solution(r, l, om, rho, kappa):
    k = np.sqrt(om ** 2 * rho / kappa)
    f, fp, _ = sbessel(l, k * r, "j")
    return np.array([k * fp, -om ** 2 * rho * f])


def fluid_matrix(r, l, om, rho, kappa):
    return np.array([[-2 / r, 1 / kappa - l * (l + 1.0) / (om ** 2 * rho * r * r)], [-om ** 2 * rho, 0]])


def toroidal_solution(r, l, om, rho, mu, kind="j", zeta=None):
    ks = np.sqrt(om ** 2 * rho / mu)
    f, fp, _ = sbessel(l, ks * r, kind) if zeta is None else ratio_solution(l, ks * r, zeta)
    return np.array([f, mu * (ks * fp - f / r)])


def hankel_ratio(l, z, kind):
    """z h_{l+1}(z)/h_l(z) of the spherical Hankel function of the first or second kind."""
    s = 1j if kind == 1 else -1j
    return z * (sp.spherical_jn(l + 1, z) + s * sp.spherical_yn(l + 1, z)) / (sp.spherical_jn(l, z) + s * sp.spherical_yn(l, z))


def reference_kind2_ratio(l, z):
    """What sphancfInitialValues returns for kind 2 (initialValues.f90:487-526): it evaluates the
    continued fraction of the first kind (b_j = 2(z + j i)) and only flips the signs of the final
    assembly, n+1 + i z + i CF instead of n+1 - i z - i CF, so its result is 2(l+1) minus the
    first-kind ratio.  That equals z h2_{l+1}/h2_l only asymptotically (z >> l); near the turning
    point z ~ l it differs by O(1)."""
    return 2.0 * (l + 1) - hankel_ratio(l, z, 1)


def toroidal_matrix(r, l, om, rho, mu):
    return np.array([[1 / r, 1 / mu], [-om ** 2 * rho + (l * (l + 1.0) - 2) * mu / r ** 2, -3 / r]])


def ode_residual(sol, mat, r, dr=None):
    """max |y' - A y| / |A y| with y' by a fourth-order central difference: a self-check of the
    closed forms above against the textbook matrices."""
    dr = dr or r * 1e-4
    d = (-sol(r + 2 * dr) + 8 * sol(r + dr) - 8 * sol(r - dr) + sol(r - 2 * dr)) / (12 * dr)
    Ay = mat(r) @ sol(r)
    return float(np.max(np.abs(d - Ay)) / np.max(np.abs(Ay)))
