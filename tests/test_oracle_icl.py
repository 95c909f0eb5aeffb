"""Pins of the oracle's ICL criterion and GMM entropy (CPU only).

ICL is the reference's default criterion (src/core/similarity.h:66):
ICL = BIC - 2E (src/core/pairwise_clustering.cpp:167-172), with E from
GMM::calcEntropy (src/core/pairwise_gmm.cpp:348-360), the sum over the
points of log(loggamma[label][i]), evaluated on whatever `loggamma` holds
when the EM loop ends (:413-431):

* a fit that ran all 100 iterations ends after an M step, which
  exponentiated `loggamma` in place, so E = sum of log(gamma[label]) =
  sum of the log responsibilities at the labels;
* a converged fit breaks out after an E step, so `loggamma` still holds the
  log responsibilities and E = sum of log(log responsibility at the label).

The reference's log responsibilities are log N_k(x) - log p(x), without
log pi_k (:158-187), and p(x) = sum_j pi_j N_j(x) <= max_j N_j(x), so the
largest one, the label's, is >= 0: a converged fit's E is finite unless a
point's maximum rounds to 0 (-inf) or below (NaN).  With K = 1 every log
responsibility is exactly 0 (log p(x) = log N(x) + log(exp(0))), so a
one-component fit's E is -inf and its ICL +inf, and under ICL it is never
chosen.

The entropy is pinned against an independent double-precision EM (the one of
test_oracle_em_fp64.py) run for the oracle's number of E steps, on converged
and 100-iteration fits; two planted slips in the oracle's entropy (taken on
gamma after a converged fit, log pi kept in the responsibilities) must fail
that comparison.
"""
import math

import numpy as np
import pytest

import oracle
from test_oracle_em_fp64 import blobs, lcg_draws


def f32(v):
    return float(np.float32(v))


def ulp(v):
    return float(np.spacing(np.float32(abs(v))))


# -------------------------------------------------------------------- ICL

def test_icl_with_zero_entropy_is_bic():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        K = int(rng.integers(1, 9))
        N = int(rng.integers(1, 33000))
        logL = f32(rng.uniform(-1e5, 1e3))
        assert np.float32(oracle.icl(K, logL, N, 0.0)).view(np.uint32) == \
            np.float32(oracle.bic(K, logL, N)).view(np.uint32)


def test_icl_is_bic_minus_twice_entropy():
    # ICL - BIC = -2E up to the two float roundings of the results; a sign
    # slip (+2E) or a dropped factor (-E) misses by |E| or more
    rng = np.random.default_rng(4)
    for _ in range(2000):
        K = int(rng.integers(1, 9))
        N = int(rng.integers(30, 33000))
        logL = f32(rng.uniform(-1e5, 1e3))
        E = f32(rng.uniform(-500, 500))
        b = oracle.bic(K, logL, N)
        i = oracle.icl(K, logL, N, E)
        tol = 2 * ulp(max(abs(b), abs(i), abs(2 * E)))
        assert abs((i - b) - (-2.0 * E)) <= tol, (K, N, logL, E, b, i)
    # larger entropy, smaller ICL
    assert oracle.icl(3, -900.0, 400, 10.0) < oracle.icl(3, -900.0, 400, -10.0)


def test_icl_special_entropies():
    assert oracle.icl(2, -500.0, 300, -math.inf) == math.inf
    assert oracle.icl(2, -500.0, 300, math.inf) == -math.inf
    assert math.isnan(oracle.icl(2, -500.0, 300, math.nan))


# ---------------------------------------------------------------- entropy

def em64_log_resp(X, K, T):
    """Log responsibilities log N_k(x) - log p(x) (the reference's loggamma
    after an E step) of the T-th E step of a double-precision EM from the
    same initial draws."""
    X = X.astype(np.float64)
    N = len(X)
    mu = X[lcg_draws(K, N)].copy()
    sig = np.tile(np.eye(2), (K, 1, 1))
    logpi = np.full(K, math.log(1.0 / K))
    lg = None
    for t in range(T):
        det = sig[:, 0, 0] * sig[:, 1, 1] - sig[:, 0, 1] * sig[:, 1, 0]
        inv = np.linalg.inv(sig)
        nrm = -0.5 * (2.0 * math.log(2.0 * math.pi) + np.log(det))
        d = X[None, :, :] - mu[:, None, :]
        lpdf = nrm[:, None] - 0.5 * np.einsum("kni,kij,knj->kn", d, inv, d)
        a = logpi[:, None] + lpdf
        m = a.max(0)
        lpx = m + np.log(np.exp(a - m).sum(0))
        lg = lpdf - lpx
        if t == T - 1:
            break
        mg = lg.max(1)
        logGamma = mg + np.log(np.exp(lg - mg[:, None]).sum(1))
        b = logpi + logGamma
        logpi = logpi + logGamma - (b.max() + math.log(np.exp(b - b.max()).sum()))
        g = np.exp(lg)
        G = np.exp(logGamma)
        mu = (g @ X) / G[:, None]
        d = X[None, :, :] - mu[:, None, :]
        sig = np.einsum("kn,kni,knj->kij", g, d, d) / G[:, None, None]
    return lg


def case_data(seed):
    r = np.random.default_rng(100 + seed)
    K = int(r.integers(2, 5))
    nc = int(r.integers(1, 4))
    centers = [tuple(c) for c in r.uniform(-4, 4, (nc, 2))]
    X = blobs(200 + seed, centers, int(r.integers(60, 200)), float(r.uniform(0.8, 2.0)))
    return X, K


# converged (c) and 100-iteration (h) fits of overlapping blobs, K = 2-4
CASES = [0, 1, 2, 5, 6, 8, 10, 11, 14, 15, 16, 17, 19, 24, 33, 35]


def entropy_agrees(seed):
    """The oracle's entropy against the FP64 evaluation of the same
    formula at the oracle's labels; returns (agrees, converged)."""
    X, K = case_data(seed)
    ok, labels, logL, E, st = oracle.gmm_fit(X, K)
    assert ok
    T, converged = int(st[0]), bool(st[2])
    lg = em64_log_resp(X, K, T)
    at = lg[labels.astype(np.int64), np.arange(len(X))]
    assert np.all(at >= -1e-12)          # the label's log responsibility is >= 0
    with np.errstate(divide="ignore"):
        terms = np.log(at) if converged else at
    E64 = terms.sum()
    tol = 1e-3 * np.abs(terms).sum() + 1e-3
    return bool(np.isfinite(E64) and abs(E - E64) <= tol), converged


@pytest.mark.parametrize("seed", CASES)
def test_entropy_matches_fp64(seed):
    good, _ = entropy_agrees(seed)
    assert good


def test_entropy_cases_cover_both_endings():
    kinds = {entropy_agrees(s)[1] for s in CASES}
    assert kinds == {True, False}


@pytest.mark.parametrize("plant", [1, 2])
def test_entropy_pin_rejects_planted_slips(plant):
    oracle.set_plant(plant)
    try:
        results = [entropy_agrees(s)[0] for s in CASES]
    finally:
        oracle.set_plant(0)
    assert not all(results)


def test_one_component_entropy_is_minus_inf():
    # every log responsibility is log N(x) - (log N(x) + log 1) = 0
    for seed in range(6):
        X, _ = case_data(seed)
        ok, labels, logL, E, st = oracle.gmm_fit(X, 1)
        assert ok and E == -math.inf
        assert int(st[0]) == 3 and st[2] == 1     # converged on the third E step
        assert oracle.icl(1, logL, len(X), E) == math.inf


def test_icl_never_selects_one_component_and_can_select_converged_fits():
    from kinc_amd import gemgen
    from kinc_amd import _native as native
    x = gemgen.modules(40, 200, seed=21, nan_rate=0.1)
    p = native.make_params(clusmethod="gmm", corrmethod="pearson", crit="ICL")
    K, labels, corr, iters = oracle.execute_block(p, x, 0, 780)
    assert np.count_nonzero(K == 1) == 0
    assert np.count_nonzero(K > 1) > 0
    chosen = iters[np.arange(len(K))[K > 1], K[K > 1].astype(np.int64) - 1]
    assert np.any(chosen & (1 << 12)), "no converged fit chosen on this sample"
    # with BIC the same pairs do choose one component
    pb = native.make_params(clusmethod="gmm", corrmethod="pearson", crit="BIC")
    Kb = oracle.execute_block(pb, x, 0, 780)[0]
    assert np.count_nonzero(Kb == 1) > 0
