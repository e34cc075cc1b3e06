"""Pins of the oracle's start values, fluid, radial (l = 0) and adjoint (solidt)
systems and of the unit-jump Green functions away from the source, against the
textbook closed-form solutions of a homogeneous medium (tests/closed_forms.py).

Each check compares per component or through a bilinear invariant, so a wrong
sign, a dropped term or a transposed entry in any single matrix element or
start-value component fails it, whatever its magnitude relative to the others.
"""
import numpy as np
import pytest

import oracle
from closed_forms import (fluid_solution, hankel_ratio, minors, ode_residual, reference_kind2_ratio, solid_matrix,
                          solid_solutions, toroidal_solution, toroidal_matrix)
from gemini_unified_amd import MC_PI, ExternalRadialNodes, NodeEarthmodel, dll1_of
from gemini_unified_amd.earthmodel import _layered_json

R = 6371000.0
OMIM = -0.0024


def homogeneous(vp=8000.0, vs=4500.0, rho=3300.0, n=60, fluid_band=None):
    """Homogeneous sphere (central sphere r <= 100 km plus one layer to R); with fluid_band = (r1, r2),
    the band r1..r2 is a homogeneous fluid (vs = 0) of the same vp and density."""
    layer = lambda r, fl: dict(r=r, fluid=fl, vp=np.full(r.size, vp), vs=np.full(r.size, 0.0 if fl else vs),
                               density=np.full(r.size, rho), qk=np.full(r.size, 57823.0),
                               qm=np.full(r.size, -1.0 if fl else 300.0))
    lays = [layer(np.array([100000.0]), False)]
    if fluid_band is None:
        lays.append(layer(np.linspace(100000.0, R, n), False))
        return NodeEarthmodel.from_json(_layered_json("homogeneous", R, lays))
    r1, r2 = fluid_band                       # solid "inner core" below, the band as outer core (layer 3)
    lays += [layer(np.linspace(100000.0, r1, 8), False), layer(np.linspace(r1, r2, n), True),
             layer(np.linspace(r2, R, n), False)]
    return NodeEarthmodel.from_json(_layered_json("homogeneous", R, lays, nic=2, noc=3, ric=r1, roc=r2))


def moduli(orc, r, nl):
    """(rho, lambda, mu, kappa) of the isotropic medium at r from the oracle's spline tables (the VTI
    moduli A = C, F, L = N are pinned separately against their closed forms)."""
    rho, z, _ = orc.eval_elcon(r, nl)
    C, F, L = z[1], z[2], z[3]
    return rho, F, L, z[5]


def ratio_spread(a, b):
    """max relative spread of the componentwise ratio a_k / b_k (a parallel to b componentwise)."""
    q = a / b
    return float(np.abs(q / q[np.argmax(np.abs(b))] - 1.0).max())


def test_closed_forms_solve_the_textbook_systems():
    """Self-check of tests/closed_forms.py: each closed form solves its textbook system."""
    om = 2 * np.pi * 0.02 + 1j * OMIM
    rho, mu = 3300.0, 3300.0 * 4500.0 ** 2 * (1 + 3e-3j)
    lam = 3300.0 * 8000.0 ** 2 - 2 * mu
    for l in (0, 7, 60):
        for kind in ("j", "h1", "h2"):
            for i in range(2):
                assert ode_residual(lambda x: solid_solutions(x, l, om, rho, lam, mu, kind)[i],
                                    lambda x: solid_matrix(x, l, om, rho, lam, mu), 4.0e6) < 1e-7
            assert ode_residual(lambda x: toroidal_solution(x, l, om, rho, mu, kind),
                                lambda x: toroidal_matrix(x, l, om, rho, mu), 4.0e6) < 1e-7


# ---------------------------------------------------------------------------
# start values (initialValues.f90:202-469) against the closed-form minors
# ---------------------------------------------------------------------------
def _wavenumbers(om, rho, lam, mu):
    omc = om + 1j * OMIM
    return np.sqrt(omc ** 2 * rho / (lam + 2 * mu)), np.sqrt(omc ** 2 * rho / mu)


@pytest.mark.parametrize("f,l,r,kind,bessel", [
    (0.02, 60, 2.0e6, 0, "j"),       # regular solution below the turning point (zetl)
    (0.02, 120, 3.0e6, 0, "j"),
    (0.03, 3, 5.5e6, 0, "h1"),       # above the turning point: Hankel ratio (sphancf kind 1)
    (0.02, 5, 6.0e6, 0, "h1"),
    (0.02, 30, 6.0e6, 4, "ref2"),    # outgoing start of a full space: the reference's kind-2 ratio
    (0.035, 120, 5.0e6, 4, "ref2"),
    (0.05, 300, 6.3e6, 4, "ref2"),
])
def test_minors_start_values_are_closed_form_minors(f, l, r, kind, bessel):
    """Every component m1..m5 of the start vector is the closed-form minor M12, M13, M14, M23, M24 of
    the P and SV solutions, to one common factor (the continued fractions stop at 1e-6).  For the
    full-space start values the solutions are those of the ratio the reference's kind-2 continued
    fraction returns (closed_forms.reference_kind2_ratio, pinned against scipy below), which pins the
    start formula itself."""
    nem = homogeneous()
    nodes = ExternalRadialNodes.split_at_discos(nem, np.array([R]), R)
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    om = 2 * MC_PI * f
    rho, lam, mu, _ = moduli(orc, r, 2)
    rc, ys, _ = orc.initial_values(om, OMIM, dll1_of(l, R), r, 2, kind)
    assert rc == 0
    if bessel == "ref2":
        kp, ks = _wavenumbers(om, rho, lam, mu)
        zeta = (reference_kind2_ratio(l, kp * r), reference_kind2_ratio(l, ks * r))
        m = minors(solid_solutions(r, l, om + 1j * OMIM, rho, lam, mu, zeta=zeta))
    else:
        m = minors(solid_solutions(r, l, om + 1j * OMIM, rho, lam, mu, bessel))
    err = ratio_spread(ys, m[:5])
    print(f"minors start values, kind {kind}, {bessel}: componentwise ratio spread {err:.2e}")
    assert err < 2e-5
    # the dependent minor m6 = M34 = -m1/l(l+1) (spheroidalMinors.f90:137-142,299)
    assert abs(m[5] + m[0] / dll1_of(l, R)) < 1e-9 * np.abs(m).max()


def test_full_space_start_values_are_outgoing_only_asymptotically():
    """A finding recorded here: with the reference's kind-2 ratio the full-space start vector is the
    minors of the outgoing (h2) solutions only where z >> l.  At l = 2 (z ~ 170) it agrees with the
    h2 minors to 1e-4; at l = 30 (z ~ 170 for S, ~ 95 for P) m1 is off by 1 %, and near the P
    turning point (l = 300, z ~ 250) by O(1).  This is the residual the full-space known-answer test
    leaves in m1.  The oracle and the GPU keep the reference's value."""
    nem = homogeneous()
    nodes = ExternalRadialNodes.split_at_discos(nem, np.array([R]), R)
    orc = oracle.Oracle(nem, nodes)
    dev = []
    for f, l, r in ((0.02, 2, 6.0e6), (0.02, 30, 6.0e6), (0.05, 300, 6.3e6)):
        orc.set_frequency(f, 3)
        om = 2 * MC_PI * f
        rho, lam, mu, _ = moduli(orc, r, 2)
        rc, ys, _ = orc.initial_values(om, OMIM, dll1_of(l, R), r, 2, 4)
        assert rc == 0
        dev.append(ratio_spread(ys, minors(solid_solutions(r, l, om + 1j * OMIM, rho, lam, mu, "h2"))[:5]))
    print(f"deviation from the h2 minors: {dev}")
    assert dev[0] < 1e-4 and 3e-3 < dev[1] < 0.1 and dev[2] > 0.3


@pytest.mark.parametrize("f,l,r,kind,bessel", [
    (0.02, 60, 2.0e6, 2, "j"), (0.03, 3, 5.5e6, 2, "h1"), (0.02, 30, 6.0e6, 5, "ref2"),
    (0.05, 300, 6.3e6, 5, "ref2"),
])
def test_toroidal_start_values_are_closed_form(f, l, r, kind, bessel):
    nem = homogeneous()
    nodes = ExternalRadialNodes.split_at_discos(nem, np.array([R]), R)
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    om = 2 * MC_PI * f
    rho, lam, mu, _ = moduli(orc, r, 2)
    rc, ys, _ = orc.initial_values(om, OMIM, dll1_of(l, R), r, 2, kind)
    assert rc == 0
    if bessel == "ref2":
        _, ks = _wavenumbers(om, rho, lam, mu)
        w = toroidal_solution(r, l, om + 1j * OMIM, rho, mu, zeta=reference_kind2_ratio(l, ks * r))
    else:
        w = toroidal_solution(r, l, om + 1j * OMIM, rho, mu, bessel)
    assert ratio_spread(ys[:2], w) < 2e-5


@pytest.mark.parametrize("x,el", [(3.0 + 0.1j, 2.0), (25.0 - 0.5j, 7.0), (12.0 + 0.3j, 20.0), (60.0 - 1j, 3.0),
                                  (45.0 - 0.02j, 30.0), (100.0 - 0.1j, 120.0)])
def test_sphancf_kind2_is_the_reference_formula(x, el):
    """sphancf of kind 2 returns 2(el+1) - x h1_{n+1}/h1_n (the first-kind fraction with the signs of
    the assembly flipped, as written in initialValues.f90:519-523); the true second-kind ratio is
    printed beside it."""
    rc, z = oracle.sphancf(complex(x), el, 2)
    assert rc == 0
    n = int(el)
    want = reference_kind2_ratio(n, x)
    true2 = hankel_ratio(n, x, 2)
    print(f"x={x}, l={n}: |ref - h2 ratio| / |h2 ratio| = {abs(z - true2) / abs(true2):.2e}")
    assert abs(z - want) / abs(want) < 1e-5


# ---------------------------------------------------------------------------
# fluid and radial (l = 0) systems: known answers of homogeneous media
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("f,l", [(0.02, 20), (0.03, 45)])
def test_fluid_system_known_answer(f, l):
    """A homogeneous fluid band whose turning point lies inside it: the upward minors start there
    from the regular fluid solution, and R/U at every fluid node is the closed form
    -om^2 rho j_l(kr) / (k j_l'(kr)), k^2 = om^2 rho / kappa (fluidOdeSystem.f90:78-95)."""
    nem = homogeneous(fluid_band=(300000.0, 3480000.0))
    rn = np.array([1.6e6, 2.1e6, 2.6e6, 3.1e6, 4.5e6, R])
    nodes = ExternalRadialNodes.split_at_discos(nem, rn, R)
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    om, dll1 = 2 * MC_PI * f, dll1_of(l, R)
    rst = orc.start_radius(om, dll1, False)
    assert 300000.0 < rst < 1.6e6
    rc, u, d, ulg, dlg, det = orc.spheroidal_minors(om, OMIM, dll1, 1, nodes.nnod, 1e-11)
    assert rc == 0
    checked = 0
    for k in range(nodes.nnod):
        if nodes.state[k] != 1:
            continue
        r = nodes.rnod[k]
        rho, _, _, kappa = moduli(orc, r, nodes.layer[k])
        y = fluid_solution(r, l, om + 1j * OMIM, rho, kappa)
        q, want = u[k, 1] / u[k, 0], y[1] / y[0]
        print(f"fluid node r={r:.3e}: R/U {q:.6e} closed form {want:.6e}")
        assert abs(q / want - 1) < 1e-6
        checked += 1
    assert checked >= 4


@pytest.mark.parametrize("f", [0.012, 0.031])
def test_radial_system_known_answer(f):
    """l = 0 in a homogeneous solid sphere: the radial system's solution from the centre is the
    regular P solution, R/U = (lam+2mu) k j_1'(kr)/j_1(kr) + 2 lam/r (radialOdeSystem.f90:78-99)."""
    nem = homogeneous()
    rn = np.array([1.5e6, 2.5e6, 3.5e6, 4.5e6, 5.5e6, R])
    nodes = ExternalRadialNodes.split_at_discos(nem, rn, R)
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    om = 2 * MC_PI * f
    rc, u, d, ulg, dlg, det = orc.spheroidal_minors(om, OMIM, 0.0, 1, nodes.nnod, 1e-11)
    assert rc == 0
    for k in range(nodes.nnod):
        r = nodes.rnod[k]
        rho, lam, mu, _ = moduli(orc, r, nodes.layer[k])
        x = solid_solutions(r, 0, om + 1j * OMIM, rho, lam, mu, "j")[0]
        q, want = u[k, 1] / u[k, 0], x[1] / x[0]
        print(f"l=0 node r={r:.3e}: R/U {q:.6e} closed form {want:.6e}")
        assert abs(q / want - 1) < 1e-6


# ---------------------------------------------------------------------------
# adjoint system (solidtOdeSystem.f90:77-114) and the unit-jump Green functions
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("f,l", [(0.02, 30), (0.035, 120)])
def test_adjoint_bilinear_invariant(f, l):
    """The adjoint system is a' = -A^T a, so a(r)^T x(r) is the same at every node for any solution x
    of the solid system: checked with the closed-form P and SV solutions (regular and outgoing) on an
    adjoint solution integrated 3871 km downwards from a random start."""
    nem = homogeneous()
    rn = np.array([2.5e6, 3.5e6, 4.5e6, 5.5e6, 6.2e6, R])
    nodes = ExternalRadialNodes.split_at_discos(nem, rn, R)
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    om, dll1 = 2 * MC_PI * f, dll1_of(l, R)
    rng = np.random.default_rng(7)
    ys = rng.normal(size=4) + 1j * rng.normal(size=4)
    rc, adj, yend = orc.spheroidal_adjoint(om, OMIM, dll1, ys, R, 2.5e6, 1e-11)
    assert rc == 0
    assert np.all(np.abs(adj).max(axis=1) > 0)
    rho, lam, mu, _ = moduli(orc, 4e6, 2)
    for bessel in ("j", "h2"):
        for i in range(2):
            x = np.array([solid_solutions(r, l, om + 1j * OMIM, rho, lam, mu, bessel)[i] for r in nodes.rnod])
            inv = np.sum(adj * x, axis=1)
            scale = np.sum(np.abs(adj * x), axis=1)
            err = np.abs(inv - inv[-1]).max() / scale.max()
            print(f"adjoint invariant ({bessel}, solution {i}): spread {err:.2e}")
            assert err < 1e-7


def symplectic(x, y, r, dll1):
    """r^2 [U_x R_y - R_x U_y + l(l+1) (V_x S_y - S_x V_y)]: constant in r for two solutions of the
    solid system (checked on closed forms in the test below)."""
    return r ** 2 * (x[0] * y[1] - x[1] * y[0] + dll1 * (x[2] * y[3] - x[3] * y[2]))


def test_ujgf_are_solid_solutions_away_from_the_source():
    """With the receiver node 300 km deep, both adjoint legs (up to the surface, down below the
    source) have length.  Between nodes the spheroidal Green functions of every jump solve the solid
    system, so their symplectic product with a closed-form solution is the same at every node on one
    side of the source (toroidal: r^2 (W T' - T W')); the jump at the source changes it."""
    nem = homogeneous()
    depths = np.array([0.0, 120e3, 300e3, 450e3, 600e3, 900e3])
    nodes = ExternalRadialNodes.split_at_discos(nem, np.sort(R - depths), R)
    jre = int(np.argmin(np.abs(nodes.rnod - (R - 300e3)))) + 1
    orc = oracle.Oracle(nem, nodes)
    f, l = 0.03, 40
    orc.set_frequency(f, 3)
    om, dll1 = 2 * MC_PI * f, dll1_of(l, R)
    assert orc.start_radius(om, dll1, False) < nodes.rnod[0]
    rc, gf, ujs, ujt, flag = orc.gfk_column(om, OMIM, dll1, jre, 1, 1e-11)
    assert rc == 0 and flag == 0
    rho, lam, mu, _ = moduli(orc, 6e6, 2)
    omc = om + 1j * OMIM
    # self-check of the invariant on closed-form solutions: the regular and the outgoing P solution
    # (two regular solutions, or a P and an SV solution, have a zero product)
    pv = [symplectic(solid_solutions(r, l, omc, rho, lam, mu, "j")[0], solid_solutions(r, l, omc, rho, lam, mu, "h2")[0],
                     r, dll1) for r in (5.5e6, 6.0e6, 6.3e6)]
    assert np.abs(np.array(pv) / pv[0] - 1).max() < 1e-9
    above = [k for k in range(nodes.nnod) if k > jre - 1]
    below = [k for k in range(nodes.nnod) if k < jre - 1]
    assert len(above) >= 2 and len(below) >= 2
    for bessel in ("j", "h2"):
        xs = {k: solid_solutions(nodes.rnod[k], l, omc, rho, lam, mu, bessel) for k in range(nodes.nnod)}
        wt = {k: toroidal_solution(nodes.rnod[k], l, omc, rho, mu, bessel) for k in range(nodes.nnod)}
        for side in (above, below):
            for j in range(4):
                for i in range(2):
                    v = np.array([symplectic(ujs[k, j], xs[k][i], nodes.rnod[k], dll1) for k in side])
                    sc = max(np.abs(symplectic(np.abs(ujs[k, j]), np.abs(xs[k][i]), nodes.rnod[k], dll1))
                             for k in side)
                    assert np.abs(v - v[0]).max() < 1e-7 * sc, (bessel, j, i)
            for j in range(2):
                v = np.array([nodes.rnod[k] ** 2 * (ujt[k, j, 0] * wt[k][1] - ujt[k, j, 1] * wt[k][0]) for k in side])
                sc = max(nodes.rnod[k] ** 2 * (abs(ujt[k, j, 0] * wt[k][1]) + abs(ujt[k, j, 1] * wt[k][0]))
                         for k in side)
                assert np.abs(v - v[0]).max() < 1e-7 * sc, (bessel, "tor", j)
        # the source jump changes the product: above and below differ by O(1)
        v_up = symplectic(ujs[above[0], 0], xs[above[0]][0], nodes.rnod[above[0]], dll1)
        v_dn = symplectic(ujs[below[-1], 0], xs[below[-1]][0], nodes.rnod[below[-1]], dll1)
        assert abs(v_up - v_dn) > 1e-3 * max(abs(v_up), abs(v_dn))
