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
