"""Learned Bulirsch-Stoer step proposer: data, targets, GPU training.

The plain integrator (reference step control) is run over a training subset of
the (omega, l) grid with step logging: for every spheroidal-minors step it
records the features the network sees and the Bulirsch error estimate after
each of the 3..7 modified-midpoint sequences (all seven are computed).  The
error estimate after k sequences scales as h^q_k; the exponents Q_ORDER were
measured on rejected steps, which are retried from the same point with h/2
(log2 of the error ratio of the two attempts, median over C1 step logs at
eps = 1e-6 .. 1e-8).  Each record then gives the largest step h_k that the
error test would accept after k sequences, h_k = h (eps / err_k)^(1/q_k), and
the work per unit length MV(k) / h_k (MV = matvecs of k sequences).  The
regression targets are the cheapest k* among 5..7 and log10(h_k*/R).

The ninth feature is the effective order of the previous step: the number of
sequences at which its error estimate crossed eps, interpolated in log error
between the last two tests.  With the previous step size it tells the network
how far below or above the tolerance the previous step was.

The network is trained on the GPU by the library's own forward / backward /
Adam kernels.  In the "step" mode it proposes during the integration; the
reference error test still decides, and a rejected step is halved as in the
reference.  In the "policy" mode (policy_targets_from_log) it is trained on one
sample per (omega, l) and chooses, before the integration, each instance's
target try count and first step; the kernels then run the reference's step
rule around that target on the fused seven-sequence sweep.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import F, check

Q_ORDER = np.array([2.2, 3.6, 4.6, 5.2, 6.3])            # error exponent after k = 3..7 sequences
MATVECS = np.array([15.0, 24.0, 37.0, 54.0, 79.0])       # sum_{j<=k} (n_j + 1), k = 3..7
NIN, NOUT = 9, 2
FEATURES = ["log10 omega", "l/1000", "r/R", "vp km/s", "vs km/s", "rho g/cc", "log10 Q_mu", "log10 h_prev/R",
            "effective order of the previous step"]


def nparams() -> int:
    return int(_lib.lib().gm_mlp_nparams())


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    shapes = [(NIN, 64), (64,), (64, 64), (64,), (64, 64), (64,), (64, NOUT), (NOUT,)]
    parts = [rng.normal(0.0, np.sqrt(2.0 / s[0]), size=s).ravel() if len(s) == 2 else np.zeros(s) for s in shapes]
    return np.concatenate(parts).astype(np.float32)


def targets_from_log(rec, kmin=5):
    """(features X[n,9], targets T[n,2] = (log10(h*/R), k*), mask of usable records)."""
    X = rec[:, :NIN].astype(np.float64)
    lh = rec[:, 9].astype(np.float64)
    e = rec[:, 10:15].astype(np.float64)                       # log10(err_k / eps), k = 3..7
    ok = np.all(np.isfinite(e), axis=1) & np.isfinite(lh) & np.all(np.isfinite(X), axis=1)
    ok &= rec[:, 8] > 0                                        # first step of a propagation has no history
    lhk = lh[:, None] - e / Q_ORDER[None, :]                   # log10(h_k / R)
    cost = np.log10(MATVECS)[None, :] - lhk                    # log10(work per unit length)
    cost[:, :kmin - 3] = np.inf
    kbest = np.argmin(np.where(np.isfinite(cost), cost, np.inf), axis=1)
    T = np.stack([lhk[np.arange(len(lh)), kbest], kbest + 3.0], axis=1)
    return X[ok].astype(np.float32), T[ok].astype(np.float32), ok


def canonical(rec):
    """Step records in (instance, step) order.  Lanes claim log slots with an atomic counter, so the
    order in which records arrive depends on which lane wins each slot; sorted, a log is the same from
    run to run."""
    rec = np.asarray(rec)
    return rec[np.lexsort((rec[:, 17], rec[:, 16]))]


def subsample(rec, nmax, seed=4):
    """A seeded subset of at most nmax records, kept in their order."""
    if len(rec) <= nmax:
        return rec
    idx = np.sort(np.random.default_rng(seed).choice(len(rec), size=nmax, replace=False))
    return rec[idx]


def policy_targets_from_log(rec, kmin=5):
    """Per-instance targets of the policy mode (one network evaluation per (omega, l)).

    Features: those of each instance's first logged step, at its spheroidal start radius (no step
    history, as the policy kernel sees them).  Targets: the try count k* in kmin..7 that would cover the
    instance's logged steps in the fewest steps of the fused sweep, where every step computes all seven
    sequences and so costs the same: a logged step of size h covers h / h_k = (err_k / eps)^(1/q_k)
    steps of the largest step try k accepts there; and log10(h*/R) of the first step that k* accepts.
    Returns (X[n, 9], T[n, 2], instance ids)."""
    rec = canonical(rec)
    job = rec[:, 16].astype(np.int64)
    e = rec[:, 10:15].astype(np.float64)
    starts = np.flatnonzero(np.r_[True, job[1:] != job[:-1]])
    ok = np.all(np.isfinite(e), axis=1)
    terms = np.where(ok[:, None], 10.0 ** (np.where(ok[:, None], e, 0.0) / Q_ORDER[None, :]), 0.0)
    cost = np.add.reduceat(terms, starts, axis=0)
    cost[:, :kmin - 3] = np.inf
    kbest = np.argmin(cost, axis=1)
    first = rec[starts]
    e0 = first[:, 10:15].astype(np.float64)
    lh = first[:, 9].astype(np.float64) - e0[np.arange(len(first)), kbest] / Q_ORDER[kbest]
    T = np.stack([lh, kbest + 3.0], axis=1)
    good = (first[:, 17] == 0) & np.all(np.isfinite(first[:, :NIN]), axis=1) & np.all(np.isfinite(T), axis=1)
    return first[good, :NIN].astype(np.float32), T[good].astype(np.float32), job[starts][good]


def normalisation(X, T):
    mu, sd = X.mean(0), X.std(0)
    sd = np.where(sd > 1e-6, sd, 1.0)
    omu, osd = T.mean(0), T.std(0)
    osd = np.where(osd > 1e-6, osd, 1.0)
    return np.concatenate([mu, sd, omu, osd]).astype(np.float32)


def train(X, T, epochs=30, batch=1024, lr=2e-3, seed=1, device=0, w=(1.0, 0.25)):
    """Normalise, train on the GPU, return (params, norm, per-epoch losses)."""
    norm = normalisation(X, T)
    mu, sd, omu, osd = norm[:NIN], norm[NIN:2 * NIN], norm[2 * NIN:2 * NIN + 2], norm[2 * NIN + 2:]
    Xn = np.ascontiguousarray((X - mu) / sd, dtype=np.float32)
    Tn = np.ascontiguousarray((T - omu) / osd, dtype=np.float32)
    p = init_params(seed)
    losses = np.zeros(epochs, np.float32)
    check(None, _lib.lib().gm_mlp_train(int(device), p.ctypes.data_as(F), Xn.shape[0], Xn.ctypes.data_as(F),
                                        Tn.ctypes.data_as(F), float(w[0]), float(w[1]), int(epochs), int(batch),
                                        float(lr), int(seed), losses.ctypes.data_as(F)))
    return p, norm, losses


def forward(params, X, device=0):
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.zeros((X.shape[0], NOUT), np.float32)
    check(None, _lib.lib().gm_mlp_forward(int(device), np.ascontiguousarray(params, np.float32).ctypes.data_as(F),
                                          X.shape[0], X.ctypes.data_as(F), Y.ctypes.data_as(F)))
    return Y


def grad(params, X, T, w=(1.0, 1.0), device=0):
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    g = np.zeros(nparams(), np.float32)
    loss = C.c_float()
    check(None, _lib.lib().gm_mlp_grad(int(device), np.ascontiguousarray(params, np.float32).ctypes.data_as(F),
                                       X.shape[0], X.ctypes.data_as(F), T.ctypes.data_as(F), float(w[0]),
                                       float(w[1]), C.byref(loss), g.ctypes.data_as(F)))
    return loss.value, g


def split_grid(grid, frac=0.1, seed=4):
    """Seeded split of the (omega, l) grid: returns (train, test) as lists of (freq, l-array)."""
    rng = np.random.default_rng(seed)
    train, test = [], []
    for f, n in zip(grid.freqs, grid.nwn):
        m = rng.random(int(n)) < frac
        ls = np.arange(int(n))
        train.append((f, ls[m]))
        test.append((f, ls[~m]))
    return train, test


def as_batch(pairs):
    """(freq, l-array) pairs -> per-slot arrays (freqs, lmin, nl) of contiguous l runs."""
    fr, lm, nl = [], [], []
    for f, ls in pairs:
        if len(ls) == 0:
            continue
        # contiguous runs of degrees
        br = np.flatnonzero(np.diff(ls) != 1) + 1
        for run in np.split(ls, br):
            fr.append(f); lm.append(int(run[0])); nl.append(len(run))
    return np.array(fr), np.array(lm, np.int32), np.array(nl, np.int32)
