"""Seeded synthetic gene expression matrices (GEMs) for tests and benchmarks.

All matrices are FP32 [genes][samples] with NaN for missing values, drawn
from numpy's PCG64 generator with one independent substream per role
(loadings, scores, noise, mask).  N(m, s) below means standard deviation s.

  latent   x = 8 + 2*eps + L.F, L[g,3] ~ N(0,1), F[3,s] ~ N(0,1.5)
  modules  genes split contiguously into 50 modules, module m has c_m ~ U{1,2,3}
           conditions, each (module, sample) draws a condition c, and
           x = mu_g + a[m,c] + lam[g,c] * f[m,s] + 0.5*eps with mu_g ~ N(8,1),
           a ~ N(0,2), lam ~ N(0,1), f ~ N(0,1), eps ~ N(0,1); pairs of one
           module therefore carry 1-3 true clusters.
NaN masks are Bernoulli per entry, or per-gene rates r_g ~ Beta(1,4)
(inverse CDF r = 1 - (1-u)^(1/4)).
"""
import numpy as np

N_MODULES = 50


def _streams(seed):
    root = np.random.Generator(np.random.PCG64(seed))
    return root.spawn(4)


def latent(G, S, seed, nan_rate=0.05):
    r_load, r_score, r_noise, r_mask = _streams(seed)
    L = r_load.normal(0.0, 1.0, size=(G, 3))
    F = r_score.normal(0.0, 1.5, size=(3, S))
    x = 8.0 + 2.0 * r_noise.standard_normal((G, S)) + L @ F
    x = x.astype(np.float32)
    x[r_mask.random((G, S)) < nan_rate] = np.nan
    return x


def modules(G, S, seed, nan_rate=0.10, nan_beta=False, degenerate=0.0):
    r_load, r_score, r_noise, r_mask = _streams(seed)
    module = (np.arange(G) * N_MODULES) // G
    cm = r_load.integers(1, 4, size=N_MODULES)
    cond = np.minimum((r_score.random((N_MODULES, S)) * cm[:, None]).astype(np.int64), cm[:, None] - 1)
    mu = r_load.normal(8.0, 1.0, size=G)
    a = r_load.normal(0.0, 2.0, size=(N_MODULES, 3))
    lam = r_load.normal(0.0, 1.0, size=(G, 3))
    f = r_score.normal(0.0, 1.0, size=(N_MODULES, S))
    x = np.empty((G, S), np.float32)
    step = 4096
    for g0 in range(0, G, step):
        g1 = min(G, g0 + step)
        m = module[g0:g1]
        c = cond[m]
        rows = np.arange(g0, g1)[:, None]
        val = (mu[g0:g1, None] + a[m[:, None], c] + lam[rows, c] * f[m]
               + 0.5 * r_noise.standard_normal((g1 - g0, S)))
        x[g0:g1] = val.astype(np.float32)
    if degenerate > 0:
        _degenerate(x, degenerate, r_load)
    if nan_beta:
        u = r_mask.random(G)
        rate = 1.0 - (1.0 - u) ** 0.25
        mask = r_mask.random((G, S)) < rate[:, None]
    else:
        mask = r_mask.random((G, S)) < nan_rate
    x[mask] = np.nan
    return x


def _degenerate(x, frac, rng):
    """Near-degenerate structure on a fraction of genes: constant sub-blocks,
    duplicated sample columns, values quantised to 8 levels (Spearman ties),
    and 30-35-point near-collinear clusters (y = a*x + 1e-6*eps)."""
    G, S = x.shape
    genes = rng.choice(G, size=max(1, int(frac * G)), replace=False)
    for i, g in enumerate(genes):
        kind = i % 4
        if kind == 0:
            lo = int(rng.integers(0, S - 40))
            x[g, lo:lo + int(rng.integers(30, 40))] = x[g, lo]
        elif kind == 1:
            src = rng.choice(S, size=max(1, S // 20), replace=False)
            dst = rng.choice(S, size=src.size, replace=False)
            x[g, dst] = x[g, src]
        elif kind == 2:
            lo, hi = np.nanmin(x[g]), np.nanmax(x[g])
            x[g] = (lo + np.round((x[g] - lo) / max(hi - lo, 1e-6) * 7) / 7 * (hi - lo)).astype(np.float32)
        else:
            partner = (g + 1) % G
            n = int(rng.integers(30, 36))
            idx = rng.choice(S, size=n, replace=False)
            slope = rng.normal(0.0, 1.0)
            x[g, idx] = (slope * x[partner, idx] + 1e-6 * rng.standard_normal(n)).astype(np.float32)


# The workloads of the benchmark configurations, by name.
CONFIGS = {
    "A": dict(kind="latent", G=1000, S=200, seed=1, nan_rate=0.05,
              params=dict(clusmethod="none", corrmethod="pearson", minsamp=30)),
    "B": dict(kind="modules", G=5000, S=500, seed=2, nan_rate=0.10,
              params=dict(clusmethod="gmm", corrmethod="spearman", maxclus=5, crit="BIC",
                          preout=True, postout=True, minsamp=30)),
    "H": dict(kind="modules", G=20000, S=500, seed=3, nan_rate=0.10,
              params=dict(clusmethod="gmm", corrmethod="spearman", maxclus=5, crit="BIC",
                          preout=True, postout=True, minsamp=30)),
    "C": dict(kind="modules", G=20000, S=1000, seed=4, nan_beta=True,
              params=dict(clusmethod="kmeans", corrmethod="pearson", maxclus=5, crit="BIC",
                          minsamp=30)),
    "D": dict(kind="modules", G=60000, S=475, seed=5, nan_rate=0.08,
              params=dict(clusmethod="gmm", corrmethod="pearson", maxclus=5, crit="BIC",
                          minsamp=30)),
    "E": dict(kind="modules", G=2000, S=5000, seed=6, nan_rate=0.10, degenerate=0.2,
              params=dict(clusmethod="gmm", corrmethod="spearman", maxclus=5, crit="BIC",
                          minsamp=30)),
}


def make(name, G=None, S=None):
    """GEM of a named configuration, optionally at a reduced size."""
    c = dict(CONFIGS[name])
    G = G or c["G"]
    S = S or c["S"]
    if c["kind"] == "latent":
        return latent(G, S, c["seed"], c.get("nan_rate", 0.05))
    return modules(G, S, c["seed"], nan_rate=c.get("nan_rate", 0.10),
                   nan_beta=c.get("nan_beta", False), degenerate=c.get("degenerate", 0.0))


def rmt_pairs_and_block(n_pairs=100, block=500, seed=1, rb=0.5405, eps=2e-4):
    """A correlation matrix [G][G][1] for the RMT analytic whose pruned
    versions stay positive definite (so the reference's |value|-ordered
    eigenvalues are increasing spline knots): n_pairs isolated gene pairs
    with r ~ U(0.55, 0.99) (eigenvalues 1 - r and 1 + r: independent levels,
    Poisson spacings) and one block of `block` genes whose correlations lie
    within eps of rb (it enters between the thresholds 0.541 and 0.540; its
    levels near 1 - rb repel like a GOE matrix's).  Genes are shuffled."""
    rng = np.random.default_rng(seed)
    G = 2 * n_pairs + block
    m = np.zeros((G, G), np.float32)
    r = rng.uniform(0.55, 0.99, n_pairs).astype(np.float32)
    p = np.arange(n_pairs)
    m[2 * p, 2 * p + 1] = r
    m[2 * p + 1, 2 * p] = r
    e = np.triu(rng.uniform(-eps, eps, (block, block)), 1)
    blk = (rb + e + e.T).astype(np.float32)
    np.fill_diagonal(blk, 0)
    m[2 * n_pairs:, 2 * n_pairs:] = blk
    perm = rng.permutation(G)
    return np.ascontiguousarray(m[np.ix_(perm, perm)][:, :, None])
