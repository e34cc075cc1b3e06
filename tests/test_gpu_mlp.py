"""GPU kernels of the learned step proposer against the NumPy reference
(oracle/mlp_ref.py), and the step-log and learned-control modes of the
spheroidal kernel end to end on a small C1 batch."""
import numpy as np
import pytest

import mlp_ref
import oracle
from gemini_unified_amd import GreenFKEngine, MC_PI, proposer

pytestmark = pytest.mark.gpu


def _params(seed):
    p = mlp_ref.init(seed)
    rng = np.random.default_rng(seed + 100)
    p = p + rng.normal(0.0, 0.05, p.size).astype(np.float32)     # non-zero biases everywhere
    return p.astype(np.float32)


def test_forward_matches_reference():
    rng = np.random.default_rng(1)
    p = _params(3)
    X = rng.normal(size=(1537, 9)).astype(np.float32)            # not a multiple of the block size
    Y = proposer.forward(p, X)
    R, _ = mlp_ref.forward(p, X.astype(np.float64))
    err = np.abs(Y - R).max() / np.abs(R).max()
    assert err < 2e-5, err


def test_gradient_matches_reference():
    rng = np.random.default_rng(2)
    p = _params(5)
    X = rng.normal(size=(2000, 9)).astype(np.float32)
    T = rng.normal(size=(2000, 2)).astype(np.float32)
    loss, g = proposer.grad(p, X, T, w=(1.0, 0.3))
    rl, rg = mlp_ref.loss_grad(p, X, T, 1.0, 0.3)
    assert abs(loss - rl) / rl < 1e-5
    assert np.linalg.norm(g - rg) / np.linalg.norm(rg) < 1e-4


def test_training_fits_a_smooth_map():
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, size=(20000, 9)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) + 0.5 * X[:, 1] * X[:, 2], np.round(2 * X[:, 3]) + 5], 1).astype(np.float32)
    params, norm, losses = proposer.train(X, T, epochs=15, batch=512, lr=2e-3)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.1 * losses[0], losses


@pytest.fixture(scope="module")
def small_case(c1_setup):
    nem, nodes, jre = c1_setup
    f, lmin, nl = np.array([40 / 2048.0, 80 / 2048.0]), np.array([0, 100], np.int32), np.array([48, 48], np.int32)
    return nem, nodes, jre, f, lmin, nl, 5.0 / 2048.0


def test_step_log_records(small_case):
    nem, nodes, jre, f, lmin, nl, sigma = small_case
    eng = GreenFKEngine(nem, nodes, variant="fast")
    eng.configure(persistent=0)
    rec, nrec, st = eng.collect_steplog(f, lmin, nl, sigma, 1e-6, jre, jre, maxrec=200_000)
    eng.close()
    assert 0 < nrec == rec.shape[0]
    jobs = rec[:, 16].astype(np.int64)
    assert jobs.min() >= 0 and jobs.max() < nl.sum()
    done = rec[:, 15]
    assert set(np.unique(done)) <= {0.0, 3.0, 4.0, 5.0, 6.0, 7.0}
    # each lane numbers its steps 0, 1, 2, ... without gaps
    seq = rec[:, 17].astype(np.int64)
    for j in np.unique(jobs):
        s = np.sort(seq[jobs == j])
        assert np.array_equal(s, np.arange(s.size)), j
    # every record's error estimates are finite where the step had a try to test
    assert np.isfinite(rec[:, 10:15]).mean() > 0.99
    X, T, ok = proposer.targets_from_log(rec)
    assert ok.sum() > 0.5 * nrec and np.all(np.isfinite(T))
    assert np.all((T[:, 1] >= 5) & (T[:, 1] <= 7))


def test_learned_control_stays_accurate(small_case):
    """Learned control changes the step sizes, not the error test: the spectra
    stay within a small multiple of eps of the oracle's at a tight tolerance."""
    nem, nodes, jre, f, lmin, nl, sigma = small_case
    eng = GreenFKEngine(nem, nodes, variant="fast")
    eng.configure(persistent=0, propose_every=4)
    rec, nrec, _ = eng.collect_steplog(f, lmin, nl, sigma, 1e-7, jre, jre, maxrec=400_000)
    X, T, ok = proposer.targets_from_log(rec)
    params, norm, losses = proposer.train(X, T, epochs=8, batch=1024)
    eng.set_proposer(params, norm, safety=0.7)
    res = eng.gfk(f, lmin, nl, sigma, 1e-7, jre, 1, want_fp32=False)
    eng.set_proposer(enable=False)
    eng.close()
    orc = oracle.Oracle(nem, nodes)
    for j, (fj, l0, n) in enumerate(zip(f, lmin, nl)):
        orc.set_frequency(fj, 3)
        rc, ref, _ = orc.gfk_degrees(2.0 * MC_PI * fj, -sigma, int(l0), int(n), jre, 1, 1e-10)
        assert rc == 0
        k0 = int(nl[:j].sum())
        g = res["gf"][..., k0:k0 + n]
        both = np.isnan(g) & np.isnan(ref)
        g, o = np.where(both, 0, g), np.where(both, 0, ref)
        num = np.linalg.norm((g - o).reshape(-1, n), axis=0)
        den = np.linalg.norm(o.reshape(-1, n), axis=0)
        err = num[den > 0] / den[den > 0]
        assert err.max() < 1e-4, (fj, err.max())


def _rel_l2_cols(a, b):
    k = a.shape[-1]
    both = ~np.isfinite(a) & ~np.isfinite(b)
    a, b = np.where(both, 0, a), np.where(both, 0, b)
    num = np.linalg.norm((a - b).reshape(-1, k), axis=0)
    den = np.linalg.norm(b.reshape(-1, k), axis=0)
    return np.where(den > 0, num / np.where(den > 0, den, 1), num)


def test_fixed_target_matches_host_build(small_case):
    """The step rule around a fixed target try count (GM_OPT_ORDER_TARGET 7) on the GPU is the CPU
    build of the same kernel source bit for bit (exact variant), work counters included."""
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "hostsim"))
    from hostsim import HostSim
    nem, nodes, jre, f, lmin, nl, sigma = small_case
    eng = GreenFKEngine(nem, nodes, variant="exact")
    eng.configure(persistent=0, order_target=7)
    g = eng.gfk(f, lmin, nl, sigma, 1e-7, jre, 1, want_fp32=False)
    eng.close()
    sim = HostSim(nem, nodes, "exact")
    sim.order_target = 7
    r = sim.gfk(f, lmin, nl, sigma, 1e-7, jre, 1)
    both = np.isnan(g["gf"]) & np.isnan(r["gf"])
    assert np.array_equal(np.where(both, 0, g["gf"]), np.where(both, 0, r["gf"]))
    np.testing.assert_array_equal(g["stats"], r["stats"])


def test_learned_policy_needs_fewer_steps_at_equal_error():
    """What the trained network gains over the method without it: on a held-out sample of the 16-node
    global workload, the per-instance policy (network trained on the plain integrator's step logs of a
    disjoint sample) and the reference rule each run at the loosest eps of a ladder that keeps the
    FK-column error <= 1e-6 against the oracle at 1e-10.  The policy needs far fewer steps (it targets
    try 7, where the fused sweep takes its largest steps); its gain over the fixed target 7 without a
    network is printed beside it."""
    from gemini_unified_amd import ExternalRadialNodes, global_model
    nem = global_model(1, 200)
    nodes = ExternalRadialNodes.from_blocks(nem, [15], [700000.0 / 15])
    jre = nodes.node_index(nem.rearth)
    sigma = 5 / 4096.0
    rng = np.random.default_rng(5)
    fr = np.sort(rng.choice(np.arange(2, 820), 16, replace=False)) / 4096.0
    l0 = np.array([int(rng.integers(0, 2484)) for _ in fr], np.int32)
    nl = np.full(fr.size, 16, np.int32)
    orc = oracle.Oracle(nem, nodes)
    truth = []
    for fj, a in zip(fr, l0):
        orc.set_frequency(fj, 3)
        rc, ref, _ = orc.gfk_degrees(2.0 * MC_PI * fj, -sigma, int(a), 16, jre, 1, 1e-10, oracle.num_threads())
        assert rc == 0
        truth.append(ref)
    truth = np.concatenate(truth, axis=2)
    eng = GreenFKEngine(nem, nodes, variant="fast")
    eng.configure(persistent=0)
    trng = np.random.default_rng(9)
    tf = np.sort(trng.choice(np.arange(2, 820), 32, replace=False)) / 4096.0
    tl = np.array([int(trng.integers(0, 2476)) for _ in tf], np.int32)
    rec, nrec, _ = eng.collect_steplog(tf, tl, np.full(tf.size, 24, np.int32), sigma, 4e-9, jre, jre,
                                       maxrec=3_000_000)
    assert nrec <= 3_000_000
    X, T, _ = proposer.policy_targets_from_log(rec)
    params, norm, losses = proposer.train(X, T, epochs=40, batch=64, lr=2e-3, seed=1)
    ladder = [2e-8, 1e-8, 6e-9, 4e-9, 2.5e-9, 1.5e-9, 1e-9]
    result = {}
    for name in ("plain", "order7", "policy"):
        eng.set_proposer(enable=False)
        eng.configure(order_target=7 if name == "order7" else 6)
        if name == "policy":
            eng.set_proposer(params, norm, safety=0.8, mode="policy")
        for eps in ladder:
            r = eng.gfk(fr, l0, nl, sigma, eps, jre, 1, want_fp32=False)
            if _rel_l2_cols(r["gf"], truth).max() <= 1e-6:
                # the fused sweep computes all seven sequences at every step, so its work is the number of
                # steps (accepted + rejected); the reference's matvec counter charges a try-7 step 79
                # matvecs and a try-6 step 54
                result[name] = (eps, int(r["stats"][1] + r["stats"][2]), int(r["stats"][4]))
                break
    eng.close()
    print(f"equal-error operating points (eps, steps, reference matvec count): {result}")
    assert set(result) == {"plain", "order7", "policy"}
    assert result["policy"][1] < 0.7 * result["plain"][1]
