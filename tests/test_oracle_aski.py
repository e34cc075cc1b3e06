"""Pins of the oracle's multi-source (ASKI) driver restatement
(computeGreenFKSpectraForASKI.f90:458-590, dsvDerivatives.f90, solidOdeSystem.f90,
getHorcompComplexFluid in fluidOdeSystem.f90:171-188).

- The direct Green functions of a source at node js, seen at the receiver node, equal the
  synthetics driver's Green functions of a source at js obtained by reciprocity from the
  receiver (two different adjoint integrations and assemblies); the same for the force at the
  receiver seen at other mantle nodes.
- Radial derivatives equal central differences of the Green functions at nodes 100 m apart,
  in the mantle and in the fluid outer core (where V and dV/dr come from (U, R)).
- The solid system of the derivatives is the textbook (U, R, V, S) matrix of an isotropic
  medium, and the negative transpose of the adjoint system the oracle integrates.
"""
import numpy as np
import pytest

import oracle
from closed_forms import fluid_matrix, solid_matrix, toroidal_matrix
from gemini_unified_amd import MC_PI, ExternalRadialNodes, dll1_of, global_model
from test_oracle_closed_forms import OMIM, R, homogeneous, moduli


def _nodes(nem, radii):
    return ExternalRadialNodes.split_at_discos(nem, np.sort(np.asarray(radii, float)), nem.rearth)


@pytest.mark.parametrize("f,l", [(0.01, 3), (0.03, 40)])
def test_solid_system_is_the_textbook_matrix(f, l):
    nem = homogeneous()
    nodes = _nodes(nem, [R - 1000.0])
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    om, dll1 = 2 * MC_PI * f, dll1_of(l, R)
    for r in (4.0e6, 6.0e6):
        rho, lam, mu, kappa = moduli(orc, r, 2)
        omc = om + 1j * OMIM
        for sys, ref in ((0, solid_matrix(r, l, omc, rho, lam, mu)), (1, toroidal_matrix(r, l, omc, rho, mu))):
            a = orc.sysmat_at(sys, om, OMIM, dll1, r)
            # entry by entry against its own magnitude (the entries span many decades)
            err = np.abs(a - ref) / np.maximum(np.abs(ref), 1e-300)
            assert np.all((ref == 0) == (np.abs(a) == 0)), (sys, a, ref)
            assert err[ref != 0].max() < 1e-12, (sys, err)
        # the solid system is the negative transpose of the adjoint (solidt) system
        np.testing.assert_array_equal(orc.sysmat_at(0, om, OMIM, dll1, r), -orc.sysmat_at(3, om, OMIM, dll1, r).T)
    # fluid system in a fluid band
    nem = homogeneous(fluid_band=(1.5e6, 3.0e6))
    orc = oracle.Oracle(nem, _nodes(nem, [R - 1000.0]))
    orc.set_frequency(f, 3)
    r = 2.2e6
    rho, lam, mu, kappa = moduli(orc, r, 3)
    ref = fluid_matrix(r, l, om + 1j * OMIM, rho, kappa)
    a = orc.sysmat_at(2, om, OMIM, dll1, r)
    assert np.abs(a - ref)[ref != 0].max() / np.abs(ref).max() < 1e-12


def _aski_case():
    nem = global_model(1, 200)
    # receiver at the surface, sources at 15, 40 and 300 km; a node triple 100 m apart in the mantle
    # (at 900 km depth) and one in the fluid outer core
    d = np.array([0.0, 15e3, 40e3, 300e3, 900e3 - 100.0, 900e3, 900e3 + 100.0])
    rc = 2.6e6
    radii = list(nem.rearth - d) + [rc - 100.0, rc, rc + 100.0]
    nodes = _nodes(nem, radii)
    jre = nodes.node_index(nem.rearth)
    jsnod = [nodes.node_index(nem.rearth - x) for x in (15e3, 40e3, 300e3)] + [jre]
    return nem, nodes, jre, jsnod


@pytest.mark.parametrize("f,l", [(20 / 2048.0, 0), (20 / 2048.0, 7), (50 / 2048.0, 33)])
def test_direct_green_functions_equal_the_reciprocal_ones(f, l):
    nem, nodes, jre, jsnod = _aski_case()
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    om, sig, eps = 2 * MC_PI * f, -5 / 2048.0, 1e-11
    rc, gfa, st = orc.aski_degrees(om, sig, l, 1, jsnod, jre, 1, eps)
    assert rc == 0
    # every source (the moment sources at jsnod and the force at jre) seen at every other source
    # node jr equals the synthetics driver's column of a receiver at jr for a source at that node:
    # U, V (4 or 2 jumps) and W (2 or 1).  At the source node itself the two sides of the jump
    # differ, so jr = jsh is left out.  At l = 0 the synthetics driver's V spectra hold 0/0 from
    # the reciprocity factor 1/l(l+1) and are never read; only U is compared there.
    checked = 0
    for jr in jsnod:
        for istyp in (1, 0):
            rc, gfs, _ = orc.gfk_degrees(om, sig, l, 1, jr, istyp, eps)
            assert rc == 0
            njsph, njtor = (4, 2) if istyp == 1 else (2, 1)
            sel = list(range(2 * njsph)) + [4 * njsph + j for j in range(njtor)]
            keep = np.arange(len(sel)) if l > 0 else np.arange(njsph)
            srcs = [(js, jn) for js, jn in enumerate(jsnod)] if istyp == 1 else [(len(jsnod), jre)]
            for js, jsh in srcs:
                if jsh == jr:
                    continue
                a = gfa[js, jr - 1, sel, 0][keep]
                b = gfs[jsh - 1, :, 0][keep]
                scale = np.abs(b).max()
                assert scale > 0
                assert np.abs(a - b).max() < 1e-7 * scale, (jr, js, a, b)
                checked += 1
    assert checked == len(jsnod) * (len(jsnod) - 1) + len(jsnod) - 1


@pytest.mark.parametrize("f,l", [(20 / 2048.0, 0), (20 / 2048.0, 7), (50 / 2048.0, 33)])
def test_radial_derivatives_are_central_differences(f, l):
    nem, nodes, jre, jsnod = _aski_case()
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    om, sig = 2 * MC_PI * f, -5 / 2048.0
    rc, gf, st = orc.aski_degrees(om, sig, l, 1, jsnod, jre, 1, 1e-12)
    assert rc == 0
    mid_m = nodes.node_index(nem.rearth - 900e3)
    mid_c = nodes.node_index(2.6e6)
    assert nodes.state[mid_c - 1] == 1 and nodes.state[mid_m - 1] == 0
    for js in range(len(jsnod) + 1):
        njsph, njtor = (4, 2) if js < len(jsnod) else (2, 1)
        pairs = [(i, i + 2 * njsph) for i in range(2 * njsph)] + [(4 * njsph + j, 4 * njsph + njtor + j)
                                                                  for j in range(njtor)]
        for mid, fluid in ((mid_m, False), (mid_c, True)):
            k = mid - 1
            h = nodes.rnod[k + 1] - nodes.rnod[k]
            assert abs(h - 100.0) < 1e-6 and abs(nodes.rnod[k] - nodes.rnod[k - 1] - 100.0) < 1e-6
            for iv, idv in pairs:
                if fluid and iv >= 4 * njsph:
                    assert gf[js, k, iv, 0] == 0 and gf[js, k, idv, 0] == 0   # no toroidal motion
                    continue
                if l == 0 and iv >= njsph:
                    # l = 0: V = W = 0 identically; the solid system's V row still gives
                    # dV/dr = -U/r there (A(3,1) U), which the reference stores as computed
                    if iv < 2 * njsph and not fluid:
                        assert abs(gf[js, k, idv, 0] + gf[js, k, iv - njsph, 0] / nodes.rnod[k]) <= \
                            1e-15 * abs(gf[js, k, iv - njsph, 0] / nodes.rnod[k])
                    continue
                y = gf[js, k - 1:k + 2, iv, 0]
                fd = (y[2] - y[0]) / (2 * h)
                dy = gf[js, k, idv, 0]
                if fluid and njsph <= iv < 2 * njsph:
                    # A finding: at a fluid node the reference stores dV/dr = -(U - V)/r
                    # (getHorcompComplexFluid, fluidOdeSystem.f90:186-187).  Differentiating its own
                    # V = -R/(om^2 rho r) with R' = -om^2 rho U gives (U - V)/r - V rho'/rho, which the
                    # central difference confirms; the oracle keeps the reference's value.
                    U, V, r = gf[js, k, iv - njsph, 0], y[1], nodes.rnod[k]
                    assert dy == -((U - V) / r) or abs(dy + (U - V) / r) <= 1e-15 * abs((U - V) / r)
                    rho = [orc.eval_elcon(x, nodes.layer[k])[0] for x in (r - h, r + h)]
                    dy = (U - V) / r - V * (rho[1] - rho[0]) / (2 * h) / orc.eval_elcon(r, nodes.layer[k])[0]
                # second-order difference: relative error (k h)^2 / 6, about 1e-6 with k = om / 6 km/s
                # at 0.025 Hz; a wrong sign, entry or term is an O(1) error
                tol = 1e-5 * max(abs(dy), abs(y[1]) * 2 * np.pi * f / 3000.0) + 1e-300
                assert abs(fd - dy) < tol, (js, fluid, iv, fd, dy)


@pytest.mark.parametrize("f,l", [(20 / 2048.0, 0), (20 / 2048.0, 7), (50 / 2048.0, 33)])
def test_ocean_sources_equal_the_reciprocal_ones(f, l):
    """Sources in the ocean (computeGreenFKSpectraForASKI.f90:359-376, 533-542: two (U, R) jumps,
    no toroidal motion): the direct Green functions of a moment or force source at an ocean node,
    seen at every other source node, equal the synthetics driver's reciprocal ones for a receiver
    at that node and a source at the ocean node (its fluid-node source terms).  At a receiver in
    the ocean the synthetics driver takes V = -R / (r om^2 rho) at the source node's radius and
    density; the direct V uses the receiver's, so the comparison scales by r rho."""
    from ocean_model import ocean_model
    nem = ocean_model()
    ruc = nem.rb[-2]
    rn = np.array([2000e3, 4000e3, ruc - 500.0, ruc + 1000.0, nem.rearth - 500.0])
    la = np.array([nem.layer_index(r)[0] for r in rn], np.int32)
    nodes = ExternalRadialNodes(rearth=nem.rearth, rnod=rn, state=nem.iflso[la - 1].astype(np.int32), layer=la)
    jre = 3
    jsnod = [2, 4, 5, 3]
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    om, sig, eps = 2 * MC_PI * f, -5 / 2048.0, 1e-11
    rc, gfa, st = orc.aski_degrees(om, sig, l, 1, jsnod, jre, 1, eps)
    assert rc == 0
    rho = [orc.eval_elcon(r, k)[0] for r, k in zip(nodes.rnod, nodes.layer)]
    checked = 0
    for jr in (2, 3, 4, 5):
        ocean_r = nodes.state[jr - 1] == 1
        for istyp in (1, 0):
            rc, gfs, _ = orc.gfk_degrees(om, sig, l, 1, jr, istyp, eps)
            assert rc == 0
            njs = 4 if istyp == 1 else 2
            srcs = [(js, jn) for js, jn in enumerate(jsnod)] if istyp == 1 else [(len(jsnod), jre)]
            for js, jsh in srcs:
                if jsh == jr:
                    continue
                nj = 2 if nodes.state[jsh - 1] == 1 else njs
                a = gfa[js, jr - 1, :2 * nj, 0]
                bu, bv = gfs[jsh - 1, :nj, 0], gfs[jsh - 1, njs:njs + nj, 0]
                if ocean_r:
                    bv = bv * (nodes.rnod[jsh - 1] * rho[jsh - 1]) / (nodes.rnod[jr - 1] * rho[jr - 1])
                b = np.concatenate([bu, bv])
                keep = np.arange(2 * nj) if l > 0 else np.arange(nj)
                scale = np.abs(b[keep]).max()
                assert scale > 0
                assert np.abs(a[keep] - b[keep]).max() < 1e-7 * scale, (jr, js, a, b)
                if nodes.state[jsh - 1] == 1:
                    assert not gfa[js, :, 2 * 2 * nj:, 0].any()     # no toroidal motion, 8 spectra
                checked += 1
    assert checked == 4 * 3 + 3


@pytest.mark.parametrize("f,l", [(20 / 2048.0, 3), (50 / 2048.0, 33)])
def test_stress_columns_follow_from_the_direct_derivatives(f, l):
    """Columns with stresses (STRESSFLAG, computeGreenFKSpectraForSynthetics.f90:209-212, 468-489):
    R, S and T of a receiver at node jr for a source at another source node equal those the solid
    and toroidal systems give from the multi-source driver's direct U, V, W and their radial
    derivatives there, U' = A11 U + A12 R + A13 V, V' = A31 U + A33 V + A34 S, W' = A W + B T;
    U, V and W are the columns without stresses."""
    nem, nodes, jre, jsnod = _aski_case()
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    om, sig, eps = 2 * MC_PI * f, -5 / 2048.0, 1e-11
    dll1 = dll1_of(l, nem.rearth)
    rc, gfa, st = orc.aski_degrees(om, sig, l, 1, jsnod, jre, 1, eps)
    assert rc == 0
    checked = 0
    for jr in jsnod:
        r = nodes.rnod[jr - 1]
        A = orc.sysmat_at(0, om, sig, dll1, r)
        At = orc.sysmat_at(1, om, sig, dll1, r)
        for istyp in (1, 0):
            rc, gs, _ = orc.gfk_degrees_stress(om, sig, l, 1, jr, istyp, eps)
            assert rc == 0
            rc, g2, _ = orc.gfk_degrees(om, sig, l, 1, jr, istyp, eps)
            njs, njt = (4, 2) if istyp == 1 else (2, 1)
            assert gs.shape[1] == 4 * njs + 2 * njt
            # U, V, W are those of the columns without stresses, bit for bit
            sel = np.concatenate([np.arange(njs), 2 * njs + np.arange(njs), 4 * njs + np.arange(njt)])
            np.testing.assert_array_equal(gs[:, sel], g2)
            srcs = [(js, jn) for js, jn in enumerate(jsnod)] if istyp == 1 else [(len(jsnod), jre)]
            for js, jsh in srcs:
                if jsh == jr:
                    continue
                d = gfa[js, jr - 1, :, 0]
                U, V, dU, dV = d[:njs], d[njs:2 * njs], d[2 * njs:3 * njs], d[3 * njs:4 * njs]
                W, dW = d[4 * njs:4 * njs + njt], d[4 * njs + njt:4 * njs + 2 * njt]
                R = (dU - A[0, 0] * U - A[0, 2] * V) / A[0, 1]
                S = (dV - A[2, 0] * U - A[2, 2] * V) / A[2, 3]
                T = (dW - At[0, 0] * W) / At[0, 1]
                c = gs[jsh - 1, :, 0]
                for got, want in ((c[njs:2 * njs], R), (c[3 * njs:4 * njs], S), (c[4 * njs + njt:], T)):
                    assert np.abs(got - want).max() < 1e-7 * np.abs(want).max(), (jr, js, got, want)
                checked += 1
    assert checked == len(jsnod) * (len(jsnod) - 1) + len(jsnod) - 1
