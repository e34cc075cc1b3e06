"""The library's own kernel source run on the CPU (tests/hostsim: gm_device.inc and
gm_kernels.inc compiled as host C++, one thread per block), against the oracle
and between step engines.  No GPU needed.

* The exact variant is the oracle bit for bit, work counters included: the
  kernels restate the reference's arithmetic in its order.
* Learned control on the fused sweep against the sequential engine (the
  pairing that did not finish in one GPU run and stays off in the library):
  on the CPU both take the same steps, so the algorithm itself terminates.
* The per-instance policy of the trained network, and the fixed target of try
  7, against the reference's target 6 at equal spectrum error.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "hostsim"))
import mlp_ref  # noqa: E402
import oracle  # noqa: E402
from hostsim import HostSim  # noqa: E402

from gemini_unified_amd import MC_PI, ExternalRadialNodes, global_model, proposer, regional_model, dll1_of  # noqa: E402


def _rel_l2(a, b):
    k = a.shape[-1]
    both = ~np.isfinite(a) & ~np.isfinite(b)
    a, b = np.where(both, 0, a), np.where(both, 0, b)
    num = np.linalg.norm((a - b).reshape(-1, k), axis=0)
    den = np.linalg.norm(b.reshape(-1, k), axis=0)
    return np.where(den > 0, num / np.where(den > 0, den, 1), num)


@pytest.mark.parametrize("f,lmin,nl", [(40 / 2048.0, 0, 24), (90 / 2048.0, 300, 16), (7 / 2048.0, 590, 11)])
def test_kernel_source_matches_oracle_bitwise(c1_setup, f, lmin, nl):
    nem, nodes, jre = c1_setup
    sim = HostSim(nem, nodes, "exact")
    r = sim.gfk([f], lmin, nl, 5 / 2048.0, 1e-6, jre, 1)
    assert np.all(r["status"] == 0)
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    orc.stats[:] = 0
    rc, ref, st = orc.gfk_degrees(2 * MC_PI * f, -5 / 2048.0, lmin, nl, jre, 1, 1e-6)
    assert rc == 0
    g = r["gf"]
    both = np.isnan(g) & np.isnan(ref)
    assert np.array_equal(np.where(both, 0, g), np.where(both, 0, ref))
    np.testing.assert_array_equal(r["stats"], orc.stats)


def test_kernel_source_regional_and_minors_match_oracle():
    """Regional wavenumbers (dll1 = (kR)^2) and the minors outputs of a 16-node global batch."""
    nem = regional_model(6)
    nodes = ExternalRadialNodes.from_blocks(nem, [4], [5000.0])
    jre = nodes.node_index(nem.rearth)
    dwn = 2 * np.pi / 8.0e6
    sim = HostSim(nem, nodes, "exact")
    f = 1.3
    r = sim.gfk([f], 2, 40, 5 / 102.4, 1e-6, jre, 1, dwn=dwn)
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    rc, ref, status = orc.gfk_wavenumbers(2 * MC_PI * f, -5 / 102.4, 2, 40, dwn, jre, 1, 1e-6)
    assert rc == 0 and np.all(status == 0)
    assert _rel_l2(r["gf"], ref).max() == 0.0
    gnem = global_model(1, 200)
    gnodes = ExternalRadialNodes.from_blocks(gnem, [15], [700000.0 / 15])
    gjre = gnodes.node_index(gnem.rearth)
    gsim = HostSim(gnem, gnodes, "exact")
    m = gsim.minors([0.11], 1200, 6, 5 / 4096.0, 1e-7, gjre, gjre)
    gorc = oracle.Oracle(gnem, gnodes)
    gorc.set_frequency(0.11, 3)
    for k in range(6):
        rc, u, d, ulg, dlg, det = gorc.spheroidal_minors(2 * MC_PI * 0.11, -5 / 4096.0, dll1_of(1200 + k, gnem.rearth),
                                                         gjre, gjre, 1e-7)
        assert rc == 0
        np.testing.assert_array_equal(m["u_sph"][:, :, k], u)
        np.testing.assert_array_equal(m["d_sph"][:, :, k], d)
        np.testing.assert_array_equal(m["det_sph"][:, k], det)


def _train_cpu(X, T, epochs, batch, lr=2e-3, seed=1, w=(1.0, 0.25)):
    """Adam on the NumPy reference network (the GPU trainer's algorithm, on the CPU)."""
    norm = proposer.normalisation(X, T)
    mu, sd, omu, osd = norm[:9], norm[9:18], norm[18:20], norm[20:22]
    Xn, Tn = ((X - mu) / sd).astype(np.float64), ((T - omu) / osd).astype(np.float64)
    p = mlp_ref.init(seed).astype(np.float64)
    m1, m2, step = np.zeros_like(p), np.zeros_like(p), 0
    rng = np.random.default_rng(seed)
    for _ in range(epochs):
        perm = rng.permutation(len(Xn))
        for b in range(max(1, len(Xn) // batch)):
            idx = perm[b * batch:(b + 1) * batch]
            _, g = mlp_ref.loss_grad(p, Xn[idx], Tn[idx], *w)
            step += 1
            m1 = 0.9 * m1 + 0.1 * g
            m2 = 0.999 * m2 + 0.001 * g * g
            p -= lr * (m1 / (1 - 0.9 ** step)) / (np.sqrt(m2 / (1 - 0.999 ** step)) + 1e-8)
    return p.astype(np.float32), norm


@pytest.fixture(scope="module")
def small_case(c1_setup):
    nem, nodes, jre = c1_setup
    f, lmin, nl = np.array([40 / 2048.0, 80 / 2048.0]), np.array([0, 100], np.int32), np.array([48, 48], np.int32)
    return nem, nodes, jre, f, lmin, nl, 5.0 / 2048.0


def test_learned_control_on_fused_sweep_takes_the_sequential_steps(small_case):
    """The 96 instances of the GPU learned-control test at eps 1e-7, network consulted every 4 steps,
    safety 0.7: on the CPU the fused sweep and the sequential engine make the same decisions (same
    accepted and rejected steps, same statuses) and both finish.  So the GPU run of this pairing that
    did not finish within minutes has no cause in the step-control algorithm; the library keeps the
    pairing off (kLearnedFused = false)."""
    nem, nodes, jre, f, lmin, nl, sigma = small_case
    sim = HostSim(nem, nodes, "fast")
    o = sim.steplog(f, lmin, nl, sigma, 1e-7, jre, jre, maxrec=400_000)
    assert 0 < o["nlog"] <= 400_000
    X, T, ok = proposer.targets_from_log(proposer.canonical(o["log"]))
    params, norm = _train_cpu(X, T, epochs=4, batch=1024)
    out = {}
    for v in ("fast", "fast_fused"):
        s = HostSim(nem, nodes, v)
        s.set_proposer(params, norm, safety=0.7, every=4)
        out[v] = s.gfk(f, lmin, nl, sigma, 1e-7, jre, 1, learned=True)
    a, b = out["fast"], out["fast_fused"]
    assert np.all(a["status"] == 0) and np.all(b["status"] == 0)
    np.testing.assert_array_equal(a["hist"], b["hist"])
    assert _rel_l2(b["gf"], a["gf"]).max() < 1e-10


def _sample(seed, nf, per, T=4096.0):
    rng = np.random.default_rng(seed)
    fr = np.sort(rng.choice(np.arange(2, 820), nf, replace=False)) / T
    l0 = np.array([int(rng.integers(0, 2500 - per)) for _ in fr], np.int32)
    return fr, l0, np.full(nf, per, np.int32)


def test_policy_and_order_target_beat_the_reference_rule_at_equal_error():
    """Held-out sample of the 16-node global workload (8 seeded frequencies x 8 degrees): every
    controller at the loosest eps of a short ladder that keeps the FK-column error <= 1e-6 against the
    oracle at eps 1e-10.  The fused sweep computes all seven sequences at every step, so targeting try 7
    takes about half the steps of the reference's target 6; the trained network (one evaluation per
    instance, trained on step logs of a disjoint sample) proposes try 7 and a first step."""
    nem = global_model(1, 200)
    nodes = ExternalRadialNodes.from_blocks(nem, [15], [700000.0 / 15])
    jre = nodes.node_index(nem.rearth)
    sigma = 5 / 4096.0
    test, train = _sample(5, 8, 8), _sample(9, 12, 16)
    orc = oracle.Oracle(nem, nodes)
    truth = []
    for f, a, n in zip(*test):
        orc.set_frequency(f, 3)
        rc, ref, st = orc.gfk_degrees(2 * MC_PI * f, -sigma, int(a), int(n), jre, 1, 1e-10, oracle.num_threads())
        assert rc == 0
        truth.append(ref)
    truth = np.concatenate(truth, axis=2)
    log = HostSim(nem, nodes, "fast").steplog(*train, sigma, 4e-9, jre, jre, maxrec=1_500_000)
    X, T, _ = proposer.policy_targets_from_log(log["log"])
    assert np.all(T[:, 1] == 7)
    params, norm = _train_cpu(X, T, epochs=40, batch=32)
    ladder = [2e-8, 1e-8, 6e-9, 4e-9, 2.5e-9, 1.5e-9]
    steps = {}
    for name in ("reference", "order7", "policy"):
        s = HostSim(nem, nodes, "fast")
        if name == "order7":
            s.order_target = 7
        if name == "policy":
            s.set_proposer(params, norm, safety=0.8, mode="policy")
        for eps in ladder:
            r = s.gfk(*test, sigma, eps, jre, 1)
            if _rel_l2(r["gf"], truth).max() <= 1e-6:
                h = r["hist"]
                steps[name] = int(h[0:6].sum() + h[8:14].sum())      # spheroidal + toroidal steps
                print(f"{name}: eps {eps:g}, {steps[name]} steps")
                if name == "policy":
                    assert np.all(r["policy"][:, 0] == 7)
                break
    assert set(steps) == {"reference", "order7", "policy"}
    assert steps["order7"] < 0.65 * steps["reference"]
    assert steps["policy"] < 0.65 * steps["reference"]


@pytest.mark.parametrize("model,blocks,f,attmode", [
    ("global", ([40], [150000.0]), None, 0),
    ("global", ([40], [150000.0]), 0.037, 3),
    ("regional", ([12], [4000.0]), 1.3, 1),
])
def test_node_properties_match_oracle(model, blocks, f, attmode):
    """k_node_props (the GFK file's externalNodes group) against the oracle's spline evaluation:
    bitwise in the exact variant, nodes in every layer including the central sphere."""
    nem = global_model(1, 200) if model == "global" else regional_model(6)
    nodes = ExternalRadialNodes.from_blocks(nem, *blocks)
    if model == "global":        # one node in the central sphere (layer 1, below 100 km)
        prn = np.concatenate(([50000.0], nem.rearth - 150000.0 * np.arange(40, -1, -1)))
        nodes = ExternalRadialNodes.split_at_discos(nem, prn, nem.rearth)
    f = nem.fref if f is None else f
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, attmode)
    ref = [orc.eval_elcon(r, nl) for r, nl in zip(nodes.rnod, nodes.layer)]
    ro_ref = np.array([x[0] for x in ref])
    el_ref = np.array([x[1] for x in ref])
    q_ref = np.array([x[2] for x in ref])
    if model == "global":
        assert set(nodes.layer.tolist()) == set(range(1, nem.nlay + 1))
    for variant in ("exact", "fast"):
        ro, el, q = HostSim(nem, nodes, variant).node_properties(f, attmode)
        if variant == "exact":
            np.testing.assert_array_equal(ro, ro_ref)
            np.testing.assert_array_equal(el, el_ref)
            np.testing.assert_array_equal(q, q_ref)
        else:
            np.testing.assert_allclose(ro, ro_ref, rtol=1e-13)
            np.testing.assert_allclose(el, el_ref, rtol=1e-12, atol=1e-12 * np.abs(el_ref).max())
            np.testing.assert_allclose(q, q_ref, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("istyp", [1, 0])
def test_kernel_source_core_nodes_match_oracle(istyp):
    """Nodes in the inner core, the fluid outer core (both sides of the ICB and the CMB) and the
    mantle: the inner-core propagator job, the fluid-node reciprocity and source terms, bit for bit
    with the oracle, work counters included."""
    nem = global_model(1, 200)
    prn = np.array([600e3, 1000e3, 2000e3, 3000e3, 4000e3, 5500e3, nem.rearth])
    nodes = ExternalRadialNodes.split_at_discos(nem, prn, nem.rearth)
    assert set(nodes.state.tolist()) == {0, 1} and nodes.rnod[0] < nem.ric
    jre = nodes.nnod
    for f, lmin, nl in [(40 / 2048.0, 0, 12), (10 / 2048.0, 1, 6), (30 / 2048.0, 100, 5)]:
        sim = HostSim(nem, nodes, "exact")
        r = sim.gfk([f], lmin, nl, 5 / 2048.0, 1e-6, jre, istyp)
        orc = oracle.Oracle(nem, nodes)
        orc.set_frequency(f, 3)
        orc.stats[:] = 0
        rc, ref, st = orc.gfk_degrees(2 * MC_PI * f, -5 / 2048.0, lmin, nl, jre, istyp, 1e-6)
        assert rc == 0 and np.all(r["status"] == 0)
        g = r["gf"]
        both = np.isnan(g) & np.isnan(ref)
        assert np.array_equal(np.where(both, 0, g), np.where(both, 0, ref))
        np.testing.assert_array_equal(r["stats"], orc.stats)


def test_kernel_source_ocean_nodes_match_oracle():
    """Ocean-bottom receiver with nodes in the ocean above it (fluid branch of the up leg) and in the
    outer core below it: bit for bit with the oracle, moment and force sources."""
    from ocean_model import ocean_model, ocean_nodes
    nem = ocean_model()
    nodes, jre = ocean_nodes(nem)
    for istyp in (1, 0):
        for f, lmin, nl in [(40 / 2048.0, 0, 8), (30 / 2048.0, 50, 5)]:
            r = HostSim(nem, nodes, "exact").gfk([f], lmin, nl, 5 / 2048.0, 1e-6, jre, istyp)
            orc = oracle.Oracle(nem, nodes)
            orc.set_frequency(f, 3)
            orc.stats[:] = 0
            rc, ref, st = orc.gfk_degrees(2 * MC_PI * f, -5 / 2048.0, lmin, nl, jre, istyp, 1e-6)
            assert rc == 0 and np.all(r["status"] == 0)
            g = r["gf"]
            both = np.isnan(g) & np.isnan(ref)
            assert np.array_equal(np.where(both, 0, g), np.where(both, 0, ref))
            np.testing.assert_array_equal(r["stats"], orc.stats)


def test_kernel_source_receiver_in_the_ocean_matches_oracle():
    """Receivers in the ocean (unitJumpGreenFunctions.f90:242-298): the mantle propagator from the
    ocean bottom, nodes in the inner core, the outer core, the mantle and the ocean; bit for bit with
    the oracle, moment and force sources, with the same work counters."""
    from ocean_model import ocean_model
    from test_oracle_ocean import ocean_receivers
    nem = ocean_model()
    nodes = ocean_receivers(nem)
    for jre in (6, 8):
        for istyp in (1, 0):
            for f, lmin, nl in [(10 / 2048.0, 0, 6), (40 / 2048.0, 25, 5)]:
                r = HostSim(nem, nodes, "exact").gfk([f], lmin, nl, 5 / 2048.0, 1e-6, jre, istyp)
                orc = oracle.Oracle(nem, nodes)
                orc.set_frequency(f, 3)
                orc.stats[:] = 0
                rc, ref, st = orc.gfk_degrees(2 * MC_PI * f, -5 / 2048.0, lmin, nl, jre, istyp, 1e-6)
                assert rc == 0 and np.all(r["status"] == 0)
                assert r["gf"].shape == ref.shape == (nodes.nnod, 8 if istyp == 1 else 4, nl)
                np.testing.assert_array_equal(r["gf"], ref)
                np.testing.assert_array_equal(r["stats"], orc.stats)
                assert np.abs(ref[0]).max() > 0 if lmin == 0 else True     # the inner-core node


@pytest.mark.parametrize("case", ["mantle", "core"])
def test_adjoint_pairs_keep_the_error(case):
    """Two adjoint integrations per leg (GM_OPT_ADJOINT_PAIR) in the kernel source: the four unit-jump
    start values span a plane, so combining two integrated adjoints gives the Green functions with the
    error of the four separate integrations (against the oracle at eps 1e-10), using fewer propagations."""
    import hostsim
    nem = global_model(1, 200)
    if case == "mantle":
        nodes = ExternalRadialNodes.from_blocks(nem, [15], [700000.0 / 15])
    else:
        prn = np.array([600e3, 1000e3, 2000e3, 3000e3, 4000e3, 5500e3, nem.rearth])
        nodes = ExternalRadialNodes.split_at_discos(nem, prn, nem.rearth)
    jre = nodes.node_index(nem.rearth)
    try:
        for f, lmin, nl in [(40 / 2048.0, 1, 24), (90 / 2048.0, 300, 16)]:
            orc = oracle.Oracle(nem, nodes)
            orc.set_frequency(f, 3)
            rc, truth, _ = orc.gfk_degrees(2 * MC_PI * f, -5 / 2048.0, lmin, nl, jre, 1, 1e-10)
            assert rc == 0
            res = {}
            for pair in (0, 1):
                hostsim.lib().sim_set_adjoint_pair(pair)
                res[pair] = HostSim(nem, nodes, "exact").gfk([f], lmin, nl, 5 / 2048.0, 1e-6, jre, 1)
                assert np.all(res[pair]["status"] == 0)
            e0, e1 = _rel_l2(res[0]["gf"], truth), _rel_l2(res[1]["gf"], truth)
            assert e1.max() <= 1.1 * e0.max() + 1e-9
            assert _rel_l2(res[1]["gf"], res[0]["gf"]).max() < 1e-4
            assert res[1]["stats"][0] < res[0]["stats"][0]
    finally:
        hostsim.lib().sim_set_adjoint_pair(0)


@pytest.mark.parametrize("istyp", [1, 0])
def test_kernel_source_multi_source_matches_oracle(istyp):
    """Multi-source batch (gm_aski_batch) from the kernel source: the adjoints of a jump at each
    source node in turn and k_aski_assemble's direct Green functions with their radial derivatives,
    bit for bit with the oracle (exact variant) and work counters included; nodes in the inner core,
    the fluid outer core and the mantle, sources at three mantle depths and the receiver."""
    nem = global_model(1, 200)
    prn = np.array([600e3, 2000e3, 3000e3, 5000e3, nem.rearth - 300e3, nem.rearth - 40e3, nem.rearth - 15e3,
                    nem.rearth])
    nodes = ExternalRadialNodes.split_at_discos(nem, prn, nem.rearth)
    jre = nodes.node_index(nem.rearth)
    jsnod = [nodes.node_index(nem.rearth - d) for d in (15e3, 300e3, 40e3)] + [jre]
    for variant in ("exact", "fast"):
        for f, lmin, nl in [(40 / 2048.0, 0, 10), (10 / 2048.0, 1, 5), (60 / 2048.0, 120, 5)]:
            r = HostSim(nem, nodes, variant).aski([f], lmin, nl, 5 / 2048.0, 1e-6, jsnod, jre, istyp)
            orc = oracle.Oracle(nem, nodes)
            orc.set_frequency(f, 3)
            orc.stats[:] = 0
            rc, ref, st = orc.aski_degrees(2 * MC_PI * f, -5 / 2048.0, lmin, nl, jsnod, jre, istyp, 1e-6)
            assert rc == 0 and np.all(r["status"] == 0)
            g = r["gf"]
            assert g.shape == ref.shape
            if variant == "exact":
                assert np.array_equal(g, ref)
                np.testing.assert_array_equal(r["stats"], orc.stats)
            else:
                assert _rel_l2(g, ref).max() < 1e-12


@pytest.mark.parametrize("istyp", [1, 0])
def test_kernel_source_ocean_sources_match_oracle(istyp):
    """Multi-source batch with sources in the ocean (computeGreenFKSpectraForASKI.f90:359-376) and
    a receiver 500 m below the ocean bottom, nodes in the inner core, the outer core, the mantle and
    the ocean: bit for bit with the oracle, work counters included."""
    from ocean_model import ocean_model
    nem = ocean_model()
    ruc = nem.rb[-2]
    rn = np.array([600e3, 2000e3, 4000e3, ruc - 500.0, ruc + 1000.0, nem.rearth - 500.0])
    la = np.array([nem.layer_index(r)[0] for r in rn], np.int32)
    nodes = ExternalRadialNodes(rearth=nem.rearth, rnod=rn, state=nem.iflso[la - 1].astype(np.int32), layer=la)
    jre, jsnod = 4, [3, 5, 6, 4]
    for f, lmin, nl in [(10 / 2048.0, 0, 8), (40 / 2048.0, 30, 5)]:
        r = HostSim(nem, nodes, "exact").aski([f], lmin, nl, 5 / 2048.0, 1e-6, jsnod, jre, istyp)
        orc = oracle.Oracle(nem, nodes)
        orc.set_frequency(f, 3)
        orc.stats[:] = 0
        rc, ref, st = orc.aski_degrees(2 * MC_PI * f, -5 / 2048.0, lmin, nl, jsnod, jre, istyp, 1e-6)
        assert rc == 0 and np.all(r["status"] == 0)
        assert np.array_equal(r["gf"], ref)
        np.testing.assert_array_equal(r["stats"], orc.stats)
        assert np.abs(ref[1, :, :8]).max() > 0 and not ref[1, :, 8:].any()    # an ocean source: 8 spectra


def test_kernel_source_stress_columns_match_oracle():
    """Columns with stresses (STRESSFLAG: U, R, V, S and W, T; in the ocean U, R, V and zero S slots,
    computeGreenFKSpectraForSynthetics.f90:209-212, 293-300, 468-489): bit for bit with the oracle,
    solid and ocean receivers, moment and force sources."""
    from ocean_model import ocean_model
    from test_oracle_ocean import ocean_receivers
    nem = ocean_model()
    nodes = ocean_receivers(nem)
    for jre in (4, 8):                     # mantle, ocean
        for istyp in (1, 0):
            f, lmin, nl = 30 / 2048.0, 3, 5
            r = HostSim(nem, nodes, "exact").gfk([f], lmin, nl, 5 / 2048.0, 1e-6, jre, istyp, stress=True)
            orc = oracle.Oracle(nem, nodes)
            orc.set_frequency(f, 3)
            rc, ref, st = orc.gfk_degrees_stress(2 * MC_PI * f, -5 / 2048.0, lmin, nl, jre, istyp, 1e-6)
            assert rc == 0 and np.all(r["status"] == 0)
            njs, njt = (4, 2) if istyp == 1 else (2, 1)
            assert r["gf"].shape == ref.shape == (nodes.nnod, 4 * njs + (0 if jre == 8 else 2 * njt), nl)
            np.testing.assert_array_equal(r["gf"], ref)
            if jre == 8:
                assert not ref[:, 3 * njs:].any() and np.abs(ref[:, njs:2 * njs]).max() > 0


def test_kernel_source_regional_stress_columns_match_oracle():
    """Regional wavenumbers with stresses (STRESSFLAG) for a buried receiver: bit for bit with the
    oracle; the U, V, W rows are the oracle's columns without stresses."""
    nem = regional_model(6)
    nodes = ExternalRadialNodes.from_blocks(nem, [4], [5000.0])
    jre = nodes.nnod - 2
    dwn = 2 * np.pi / 8.0e6
    f = 1.3
    for istyp in (1, 0):
        r = HostSim(nem, nodes, "exact").gfk([f], 2, 30, 5 / 102.4, 1e-6, jre, istyp, dwn=dwn, stress=True)
        orc = oracle.Oracle(nem, nodes)
        orc.set_frequency(f, 3)
        rc, ref, status = orc.gfk_wavenumbers_stress(2 * MC_PI * f, -5 / 102.4, 2, 30, dwn, jre, istyp, 1e-6)
        assert rc == 0 and np.all(status == 0)
        np.testing.assert_array_equal(r["gf"], ref)
        rc, g2, _ = orc.gfk_wavenumbers(2 * MC_PI * f, -5 / 102.4, 2, 30, dwn, jre, istyp, 1e-6)
        njs, njt = (4, 2) if istyp == 1 else (2, 1)
        uvw = np.concatenate([np.arange(njs), 2 * njs + np.arange(njs), 4 * njs + np.arange(njt)])
        np.testing.assert_array_equal(ref[:, uvw], g2)
        assert np.abs(ref[:, njs:2 * njs]).max() > 0 and np.abs(ref[:, 4 * njs + njt:]).max() > 0
