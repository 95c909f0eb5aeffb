"""Pins of the RMT oracle (oracle/kinc_oracle_rmt.c, RMT::process of
src/core/rmt.cpp:33-439) against what fixes each step independently:
SciPy's Akima interpolator (the same published 1970 method GSL implements),
LAPACK's symmetric eigensolver through numpy, brute-force maxima, the block
structure of the pruned matrix, and the chi-square statistic's behaviour on
spectra whose nearest-neighbour spacings are known to be Poisson (small) or
Wigner-Dyson (large)."""
import numpy as np
import pytest
from scipy.interpolate import Akima1DInterpolator

import oracle


@pytest.mark.parametrize("seed", range(4))
def test_akima_matches_scipy(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 60))
    x = np.sort(rng.uniform(-3, 7, n))
    x = np.unique(x)
    y = np.cumsum(rng.exponential(1.0, x.size)) * rng.choice([-1, 1])
    xe = np.concatenate([x, np.linspace(x[0], x[-1], 777)])
    got = oracle.akima(x, y, xe)
    want = Akima1DInterpolator(x, y)(xe)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(y).max())
    assert np.array_equal(got[:x.size], y)        # passes through the knots


def test_akima_reproduces_a_line_and_rejects_bad_knots():
    x = np.array([0.0, 0.5, 1.7, 2.0, 4.0, 9.0])
    y = 3.0 * x - 2.0
    xe = np.linspace(0, 9, 101)
    assert np.allclose(oracle.akima(x, y, xe), 3.0 * xe - 2.0, rtol=0, atol=1e-12)
    with pytest.raises(ValueError):
        oracle.akima(x[::-1], y, xe)               # knots must increase
    with pytest.raises(ValueError):
        oracle.akima(x[:4], y[:4], xe)             # Akima needs 5 knots


@pytest.mark.parametrize("n", [1, 2, 7, 60, 150])
def test_eigenvalues_match_lapack(n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n)).astype(np.float32)
    low = np.tril(a.astype(np.float64))
    full = low + np.tril(low, -1).T                # the solver reads the lower triangle
    w = np.linalg.eigvalsh(full)
    w = w[np.argsort(np.abs(w), kind="stable")].astype(np.float32)
    got = oracle.rmt_eigenvalues(a)
    assert np.allclose(got, w, rtol=2e-6, atol=2e-6 * np.abs(w).max())
    assert np.all(np.diff(np.abs(got)) >= -1e-6)   # |.|-ascending


def random_corr(N, K, seed, density=0.05):
    rng = np.random.default_rng(seed)
    m = np.zeros((N, N, K), np.float32)
    for k in range(K):
        r = rng.uniform(-1, 1, (N, N)).astype(np.float32)
        r[rng.random((N, N)) > density] = 0
        r = np.tril(r, -1)
        m[:, :, k] = r + r.T
    return m


def test_maximums_brute_force():
    m = random_corr(40, 3, 1)
    got = oracle.rmt_maximums(m).reshape(40, 3)
    a = np.abs(m)
    for i in range(40):
        a[i, i, :] = 0
    assert np.array_equal(got, a.max(axis=1))


def test_prune_block_structure():
    N, K, t = 50, 3, np.float32(0.6)
    m = random_corr(N, K, 2, density=0.2)
    mx = oracle.rmt_maximums(m)
    idx, pr = oracle.rmt_prune(m, mx, t)
    assert np.array_equal(idx, np.nonzero(mx >= t)[0])
    assert np.all(np.triu(pr, 1) == 0) and np.all(np.diag(pr) == 1)
    # entries only between indices of the same cluster, each one the
    # thresholded correlation of its gene pair
    gi, ki = idx // K, idx % K
    for a in range(idx.size):
        for b in range(a):
            v = m[gi[a], gi[b], ki[a]] if ki[a] == ki[b] else 0
            assert pr[a, b] == (v if abs(v) >= t else 0)
    # the spectrum is the union of the per-cluster blocks' spectra
    full = np.tril(pr.astype(np.float64))
    full = full + np.tril(full, -1).T
    blocks = [np.linalg.eigvalsh(full[np.ix_(ki == k, ki == k)]) for k in range(K)]
    assert np.allclose(np.sort(np.concatenate(blocks)), np.linalg.eigvalsh(full), atol=1e-9)


def poisson_spectrum(n, seed):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0, n)).astype(np.float32) + 1.0


def goe_spectrum(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n))
    w = np.linalg.eigvalsh((a + a.T) / np.sqrt(2 * n))
    w = w + 3.0                                   # all positive: |.| order = value order
    return w.astype(np.float32)


def test_chi_square_separates_poisson_from_goe():
    # uncorrelated levels (Poisson spacings) fit the reference's Poisson
    # expectation; GOE levels repel (Wigner-Dyson spacings) and do not
    for seed in range(3):
        assert oracle.rmt_chi_square(poisson_spectrum(1200, seed)) < 99.607
        assert oracle.rmt_chi_square(goe_spectrum(1200, seed)) > 200


def test_chi_square_needs_enough_unique_eigenvalues():
    e = poisson_spectrum(60, 1)
    e[10:] = e[10]                                # 10 unique values after the first
    assert oracle.rmt_chi_square(e) == -1
    assert oracle.rmt_chi_square(poisson_spectrum(50, 1)) == -1   # 49 after skipping the first
    assert oracle.rmt_chi_square(poisson_spectrum(51, 1)) != -1


def test_chi_square_knots_must_increase():
    # negative and positive eigenvalues interleaved by |.| give non-increasing
    # spline knots, where the reference's GSL call fails
    e = poisson_spectrum(400, 3)
    e[1::2] *= -1
    e = e[np.argsort(np.abs(e), kind="stable")]
    assert oracle.rmt_chi_square(e) == -2


def pairs_and_block(n_pairs=100, block=500, seed=1):
    """See kinc_amd.gemgen.rmt_pairs_and_block."""
    from kinc_amd import gemgen
    return gemgen.rmt_pairs_and_block(n_pairs, block, seed)


def modular_corr(N, modules, seed, noise=1.0):
    """Pearson correlations of a module-structured sample matrix, with the
    |r| < 0.5 entries dropped as the Similarity writer's default would.  Its
    pruned matrices have negative eigenvalues."""
    rng = np.random.default_rng(seed)
    S = 120
    f = rng.normal(size=(modules, S))
    mod = rng.integers(0, modules, N)
    x = f[mod] * rng.uniform(0.5, 2.0, (N, 1)) + noise * rng.normal(size=(N, S))
    r = np.corrcoef(x).astype(np.float32)
    np.fill_diagonal(r, 0)
    r[np.abs(r) < 0.5] = 0
    return r[:, :, None]


def test_run_finds_a_threshold_and_logs_each_step():
    m = pairs_and_block()
    rc, t, rows = oracle.rmt_run(m)
    assert rc > 0 and len(rows) == rc
    assert abs(t - 0.541) < 1e-4
    ts = [r[0] for r in rows]
    sizes = [r[1] for r in rows]
    assert all(a > b for a, b in zip(ts, ts[1:]))             # decreasing thresholds
    assert all(a <= b for a, b in zip(sizes, sizes[1:]))      # the pruned matrix only grows
    # thresholds are accumulated in float, as the reference does
    t32 = np.float32(0.99)
    for row in rows:
        assert row[0] == float(t32)
        t32 = np.float32(t32 - np.float32(0.001))
    # the final threshold is the last one whose chi-square was below the
    # critical value (the isolated pairs: Poisson levels), and the sweep
    # stopped at the first chi-square of 200 or more after it (the block
    # enters: repelling levels)
    sub = [r for r in rows if r[2] != -1 and r[2] < 99.607]
    assert sub and sub[-1][0] == t
    assert rows[-1][2] >= 200 and rows[-1][1] == 700
    assert all(r[2] < 200 for r in rows[:-1])


def test_run_reports_the_stop_threshold():
    rc, t, rows = oracle.rmt_run(pairs_and_block(), stop=0.6)
    assert rc == -1


def test_run_reports_failing_spline_knots():
    # pruned Pearson matrices of noisy modules have negative eigenvalues; in
    # |.| order they are not increasing spline knots, where GSL fails
    rc, t, rows = oracle.rmt_run(modular_corr(300, 10, 5))
    assert rc == -2
