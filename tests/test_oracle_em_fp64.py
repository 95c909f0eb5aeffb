"""The oracle's GMM fit against an independent double-precision EM.

The FP64 EM below is written with numpy from the reference's parametrisation
of the algorithm (src/core/pairwise_gmm.cpp:143-316 for the E and M steps,
:367-445 for GMM::fit):

* initial means: the points drawn by the per-fit LCG, `rand() % N` with
  seed 1, one draw per component (src/opencl/outlier.cl:14-18); unit
  covariances, pi = 1/K;
* E step: log N_k(x) = normalizer_k - 0.5 (x - mu_k)^T Sigma_k^-1 (x - mu_k),
  log p(x) = logsumexp_k(log pi_k + log N_k(x)), logL = sum of log p(x), and
  the kept log responsibilities are log N_k(x) - log p(x) (without log pi_k,
  as the reference keeps them);
* M step: logGamma_k = logsumexp over the points of those responsibilities,
  log pi_k += logGamma_k - logsumexp_k(log pi_k + logGamma_k), then mu_k and
  Sigma_k from the exponentiated responsibilities divided by Gamma_k.

That is textbook EM in disguise (the missing pi_k in the responsibilities
cancels in every update), so in exact arithmetic logL never decreases; the
FP64 trajectory must show it.  The oracle's FP32 fit must match the FP64 fit
run for the same number of E steps: logL to a relative 5e-6, the labels
exactly (the cases are separated well enough that no point sits near a
decision boundary at FP32 resolution), and the same success flag.  A dropped
term, a wrong sign or a transposed operand in the oracle's EM moves logL by
far more than that (the six cases agree to 7e-8 - 6e-7).
"""
import math

import numpy as np
import pytest

import oracle


def lcg_draws(K, N):
    s = 1
    out = []
    for _ in range(K):
        s = (s * 1103515245 + 12345) % (1 << 64)
        out.append(((s // 65536) % 32768) % N)
    return out


def em_fp64(X, K, T, mistake=None):
    """T E steps with an M step between consecutive ones; returns
    (ok, labels, logL trajectory).  `mistake` plants one plausible error
    (for the test that the comparison notices it)."""
    X = X.astype(np.float64)
    N = len(X)
    mu = X[lcg_draws(K, N)].copy()
    sig = np.tile(np.eye(2), (K, 1, 1))
    logpi = np.full(K, math.log(1.0 / K))
    traj = []
    lg = None
    for t in range(T):
        det = sig[:, 0, 0] * sig[:, 1, 1] - sig[:, 0, 1] * sig[:, 1, 0]
        if np.any(det <= 0):
            return False, None, traj
        inv = np.linalg.inv(sig)
        nrm = -0.5 * (2.0 * math.log(2.0 * math.pi) + np.log(det))
        d = X[None, :, :] - mu[:, None, :]
        quad = np.einsum("kni,kij,knj->kn", d, inv, d)
        lpdf = nrm[:, None] - 0.5 * quad
        a = logpi[:, None] + lpdf
        m = a.max(0)
        lpx = m + np.log(np.exp(a - m).sum(0))
        traj.append(lpx.sum())
        lg = (a if mistake == "logpi_in_gamma" else lpdf) - lpx
        if t == T - 1:
            break
        mg = lg.max(1)
        logGamma = mg + np.log(np.exp(lg - mg[:, None]).sum(1))
        b = logpi + logGamma
        logGammaSum = b.max() + math.log(np.exp(b - b.max()).sum())
        logpi = logpi + logGamma - (0.0 if mistake == "no_pi_norm" else logGammaSum)
        g = np.exp(lg)
        G = np.exp(logGamma)
        mu = (g @ X) / G[:, None]
        d = X[None, :, :] - mu[:, None, :]
        sig = np.einsum("kn,kni,knj->kij", g, d, d) / G[:, None, None]
        if mistake == "half_quad":
            sig = 2.0 * sig
    return True, lg.argmax(0).astype(np.int8), traj


def blobs(seed, centers, n, sd):
    rng = np.random.default_rng(seed)
    parts = [rng.normal(c, sd, (n, 2)) for c in centers]
    X = np.concatenate(parts)
    return X[rng.permutation(len(X))].astype(np.float32)


CASES = [
    # seed, K, centers, points per center, sd
    (11, 2, [(0, 0), (6, 4)], 150, 1.0),
    (12, 3, [(0, 0), (6, 0), (3, 6)], 120, 0.8),
    (13, 2, [(8, 8), (8, 14)], 200, 1.2),
    (14, 3, [(-5, 0), (0, 5), (5, 0)], 100, 1.0),
    (15, 4, [(0, 0), (8, 0), (0, 8), (8, 8)], 80, 1.0),
    (16, 5, [(0, 0), (9, 0), (0, 9), (9, 9), (4.5, 4.5)], 70, 0.9),
]


def agrees(X, K, mistake=None):
    """The comparison of the oracle's fit with the FP64 EM."""
    ok, labels, logL, _, st = oracle.gmm_fit(X, K)
    T = int(st[0])
    with np.errstate(all="ignore"):
        ok64, labels64, traj = em_fp64(X, K, T, mistake)
    return (ok and ok64 and T >= 3 and math.isfinite(traj[-1])
            and abs(logL - traj[-1]) <= 5e-6 * abs(traj[-1])
            and np.array_equal(labels, labels64)), traj


@pytest.mark.parametrize("seed,K,centers,n,sd", CASES)
def test_oracle_em_matches_fp64_em(seed, K, centers, n, sd):
    X = blobs(seed, centers, n, sd)
    good, traj = agrees(X, K)
    assert good
    # exact EM never lowers the likelihood (FP64 rounding aside)
    steps = np.diff(np.array(traj))
    assert np.all(steps >= -1e-9 * np.abs(np.array(traj[1:]))), steps


@pytest.mark.parametrize("mistake", ["logpi_in_gamma", "no_pi_norm", "half_quad"])
def test_comparison_rejects_planted_mistakes(mistake):
    """Each plausible slip in an EM restatement (log pi kept in the
    responsibilities, pi left unnormalised, a covariance off by a factor of
    two) fails the comparison on at least one of the cases."""
    assert not all(agrees(blobs(seed, centers, n, sd), K, mistake)[0]
                   for seed, K, centers, n, sd in CASES)


