"""Pins of the oracle's receiver in the ocean (unitJumpGreenFunctions.f90:242-298, the
synthetics driver's ocean columns computeGreenFKSpectraForSynthetics.f90:293-300,445-481).

- U is continuous across the ocean bottom: the Green functions of a receiver 1 mm above it (the
  ocean branch: jumps in (U, R), the mantle propagator from the ocean bottom, the outer and inner
  core scaled by its value at the CMB) equal those of a receiver 1 mm below it (the solid branch,
  pinned elsewhere), for sources in the inner core, the outer core, the mantle and the ocean; the
  difference falls linearly with the distance (2.3e-4 per metre at most here).
- R at a receiver in the ocean follows from U through the fluid system, U' = A11 U + A12 R; the
  central difference of U over three receivers 10 m apart gives R, and with it the stored V,
  -R / (r om^2 rho) taken at the node radius as the reference does.
"""
import numpy as np
import pytest

import oracle
from gemini_unified_amd import MC_PI, ExternalRadialNodes, dll1_of
from ocean_model import ocean_model

OMIM = -5 / 2048.0


def ocean_receivers(nem, d=0.001, h=10.0):
    """inner core, outer core, mantle, receivers d either side of the ocean bottom, three ocean
    receivers h apart 1 km above it, and a node near the surface"""
    ruc = nem.rb[-2]
    rn = np.array([600e3, 2000e3, 4000e3, ruc - 500.0, ruc - d, ruc + d, ruc + 1000.0 - h, ruc + 1000.0,
                   ruc + 1000.0 + h, nem.rearth - 500.0])
    la = np.array([nem.layer_index(r)[0] for r in rn], np.int32)
    st = nem.iflso[la - 1].astype(np.int32)
    return ExternalRadialNodes(rearth=nem.rearth, rnod=rn, state=st, layer=la)


CASES = [(10 / 2048.0, 0), (10 / 2048.0, 2), (40 / 2048.0, 5), (40 / 2048.0, 30)]


@pytest.mark.parametrize("f,l", CASES)
@pytest.mark.parametrize("istyp", [1, 0])
def test_u_continuous_across_the_ocean_bottom(f, l, istyp):
    nem = ocean_model()
    nodes = ocean_receivers(nem)
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    om, dll1 = 2 * MC_PI * f, dll1_of(l, nem.rearth)
    js, jf = 5, 6                                    # 1 mm below / above the ocean bottom
    assert nodes.state[js - 1] == 0 and nodes.state[jf - 1] == 1
    rc, gs, *_ = orc.gfk_column(om, OMIM, dll1, js, istyp, 1e-10)
    assert rc == 0
    rc, go, ujs, ujt, flag = orc.gfk_column(om, OMIM, dll1, jf, istyp, 1e-10)
    assert rc == 0 and flag == 0
    njsph, njtor = (4, 2) if istyp == 1 else (2, 1)
    assert go.shape[1] == 2 * njsph and gs.shape[1] == 2 * njsph + njtor
    assert not ujt.any() and not ujs[:, 2:].any()    # no toroidal jumps, two spheroidal ones
    checked = 0
    for n in range(nodes.nnod):
        if n + 1 in (js, jf):
            continue
        a, b = go[n, :njsph], gs[n, :njsph]
        scale = np.abs(b).max()
        if scale == 0:                               # below the start radius: zero on both sides
            assert not a.any()
            continue
        assert np.abs(a - b).max() < 1e-6 * scale, (n, a, b)
        checked += 1
    assert checked >= (8 if l < 30 else 6)
    # the inner-core node is reached at low degree
    if l <= 2:
        assert np.abs(go[0, :njsph]).max() > 0


@pytest.mark.parametrize("f,l", CASES)
def test_v_follows_from_u_by_the_fluid_system(f, l):
    nem = ocean_model()
    nodes = ocean_receivers(nem)
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    om, dll1 = 2 * MC_PI * f, dll1_of(l, nem.rearth)
    omc = om + 1j * OMIM
    jm, j0, jp = 7, 8, 9
    assert (nodes.state[[jm - 1, j0 - 1, jp - 1]] == 1).all()
    g = {j: orc.gfk_column(om, OMIM, dll1, j, 1, 1e-11) for j in (jm, j0, jp)}
    assert all(v[0] == 0 for v in g.values())
    re = nodes.rnod[j0 - 1]
    h = nodes.rnod[jp - 1] - re
    A = orc.sysmat_at(2, om, OMIM, dll1, re)
    rho_w = orc.eval_elcon(re, nodes.layer[j0 - 1])[0]
    for n in range(nodes.nnod):
        if n + 1 in (jm, j0, jp):
            continue
        U = g[j0][1][n, :4]
        if not np.abs(U).any():
            continue
        dU = (g[jp][1][n, :4] - g[jm][1][n, :4]) / (2 * h)
        R = (dU - A[0, 0] * U) / A[0, 1]
        v_true = -R / (re * omc ** 2 * rho_w)
        rn = nodes.rnod[n]
        rho_n = orc.eval_elcon(rn, nem.layer_index(rn)[0])[0]
        want = v_true * (re * rho_w) / (rn * rho_n)
        got = g[j0][1][n, 4:8]
        assert np.abs(got - want).max() < 1e-5 * np.abs(want).max(), (n, got, want)


def test_receiver_in_the_outer_core_is_unsupported():
    nem = ocean_model()
    nodes = ocean_receivers(nem)
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(0.01, 3)
    rc, *_ = orc.gfk_column(2 * MC_PI * 0.01, OMIM, dll1_of(3, nem.rearth), 2, 1, 1e-8)
    assert rc == 8
