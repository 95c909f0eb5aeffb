"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Every case runs the same seeded synthetic block through the oracle and the
device, in both execution forms, and requires bit-exact K, labels,
correlations and EM/Lloyd iteration records.
"""
import numpy as np
import pytest

import oracle
from kinc_amd import _native, gemgen
from parity_util import assert_block_equal

pytestmark = pytest.mark.gpu

# (name, generator kwargs, analytic flags, start, count)
CASES = [
    ("pearson_none_latent", dict(kind="latent", G=120, S=200, seed=11, nan_rate=0.05),
     dict(clusmethod="none", corrmethod="pearson"), 0, 7140),
    ("spearman_none_latent", dict(kind="latent", G=80, S=150, seed=12, nan_rate=0.05),
     dict(clusmethod="none", corrmethod="spearman"), 0, 3160),
    ("gmm_spearman_bic_out", dict(kind="modules", G=40, S=300, seed=13, nan_rate=0.10),
     dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", preout=True, postout=True), 0, 300),
    ("gmm_pearson_bic", dict(kind="modules", G=40, S=256, seed=14, nan_rate=0.08),
     dict(clusmethod="gmm", corrmethod="pearson", crit="BIC"), 100, 300),
    ("gmm_pearson_icl_out", dict(kind="modules", G=30, S=200, seed=15, nan_rate=0.10),
     dict(clusmethod="gmm", corrmethod="pearson", crit="ICL", preout=True, postout=True), 0, 200),
    ("kmeans_pearson_bic", dict(kind="modules", G=30, S=200, seed=16, nan_rate=0.10),
     dict(clusmethod="kmeans", corrmethod="pearson", crit="BIC", postout=True), 0, 150),
    ("kmeans_spearman_minexpr", dict(kind="modules", G=30, S=180, seed=17, nan_rate=0.05),
     dict(clusmethod="kmeans", corrmethod="spearman", crit="BIC", minexpr=7.0), 0, 150),
    ("gmm_maxclus8_small_S", dict(kind="modules", G=40, S=64, seed=18, nan_rate=0.05),
     dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", maxclus=8, minclus=2), 0, 200),
    ("gmm_maxclus8_S700_memstore", dict(kind="modules", G=16, S=700, seed=24, nan_rate=0.10),
     dict(clusmethod="gmm", corrmethod="pearson", crit="BIC", maxclus=7), 0, 120),
    # above 1024 samples: the memory-store instance
    ("gmm_spearman_S1500", dict(kind="modules", G=14, S=1500, seed=19, nan_rate=0.10),
     dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", preout=True, postout=True), 0, 91),
    ("kmeans_pearson_S1300", dict(kind="modules", G=12, S=1300, seed=20, nan_rate=0.15),
     dict(clusmethod="kmeans", corrmethod="pearson", crit="BIC"), 0, 66),
    ("gmm_spearman_E_S5000", dict(kind="modules", G=10, S=5000, seed=6, nan_rate=0.10,
                                  degenerate=0.2),
     dict(clusmethod="gmm", corrmethod="spearman", crit="BIC"), 0, 45),
    # the largest sample count the device takes (32768): memory-store fits
    # of about 26,500 points, Spearman sums far past the int32 wrap
    ("gmm_spearman_S32768_max", dict(kind="modules", G=3, S=32768, seed=31, nan_rate=0.10),
     dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", preout=True, postout=True), 0, 3),
    ("pearson_none_S32768_max", dict(kind="latent", G=6, S=32768, seed=32, nan_rate=0.05),
     dict(clusmethod="none", corrmethod="pearson"), 0, 15),
    # parameter corners and slot-count boundaries: a sample count at the edge
    # of a register slot (64, 65), of the register range (1024) and just past
    # it (1025, the memory store), one cluster only, minimum samples above
    # most pairs' N, ICL with K-means, minclus > 1
    ("kmeans_spearman_icl_out", dict(kind="modules", G=24, S=150, seed=30, nan_rate=0.10),
     dict(clusmethod="kmeans", corrmethod="spearman", crit="ICL", preout=True, postout=True),
     0, 120),
    ("gmm_pearson_minclus3_icl", dict(kind="modules", G=24, S=160, seed=31, nan_rate=0.10),
     dict(clusmethod="gmm", corrmethod="pearson", crit="ICL", minclus=3, maxclus=4), 0, 120),
    ("gmm_spearman_maxclus1", dict(kind="modules", G=24, S=120, seed=32, nan_rate=0.10),
     dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", maxclus=1, postout=True), 0, 150),
    ("gmm_pearson_minsamp_high", dict(kind="modules", G=24, S=140, seed=33, nan_rate=0.12),
     dict(clusmethod="gmm", corrmethod="pearson", crit="BIC", minsamp=110), 0, 150),
    ("gmm_spearman_S64", dict(kind="modules", G=24, S=64, seed=34, nan_rate=0.0),
     dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", minsamp=20), 0, 120),
    ("gmm_spearman_S65", dict(kind="modules", G=24, S=65, seed=35, nan_rate=0.0),
     dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", minsamp=20), 0, 120),
    ("gmm_pearson_S1024", dict(kind="modules", G=10, S=1024, seed=36, nan_rate=0.0),
     dict(clusmethod="gmm", corrmethod="pearson", crit="BIC"), 0, 45),
    ("gmm_pearson_S1025", dict(kind="modules", G=10, S=1025, seed=37, nan_rate=0.0),
     dict(clusmethod="gmm", corrmethod="pearson", crit="BIC"), 0, 45),
    # the FP64 Pearson option: double sums and quotient on both sides
    ("pearson_none_fp64", dict(kind="latent", G=100, S=300, seed=38, nan_rate=0.05),
     dict(clusmethod="none", corrmethod="pearson", pearson_fp64=True), 0, 4950),
    ("gmm_pearson_fp64_out", dict(kind="modules", G=30, S=220, seed=39, nan_rate=0.10),
     dict(clusmethod="gmm", corrmethod="pearson", crit="BIC", preout=True, postout=True,
          pearson_fp64=True), 0, 200),
    ("gmm_pearson_fp64_S1100", dict(kind="modules", G=10, S=1100, seed=40, nan_rate=0.05),
     dict(clusmethod="gmm", corrmethod="pearson", crit="ICL", pearson_fp64=True), 0, 45),
    # Pearson without clustering (one pair per lane in the persistent form):
    # expression floor, above the register range, a ragged launch at an offset
    ("pearson_none_minexpr", dict(kind="latent", G=90, S=150, seed=41, nan_rate=0.05),
     dict(clusmethod="none", corrmethod="pearson", minexpr=7.0), 0, 4005),
    ("pearson_none_S1300", dict(kind="latent", G=40, S=1300, seed=42, nan_rate=0.10),
     dict(clusmethod="none", corrmethod="pearson"), 0, 780),
    ("pearson_none_ragged", dict(kind="latent", G=50, S=77, seed=43, nan_rate=0.30),
     dict(clusmethod="none", corrmethod="pearson", minsamp=20), 100, 1000),
    # smallest shapes: one pair; fewer samples than a slot
    ("gmm_spearman_one_pair", dict(kind="modules", G=2, S=90, seed=44, nan_rate=0.05),
     dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", preout=True, postout=True,
          minsamp=10), 0, 1),
    ("gmm_spearman_S24", dict(kind="modules", G=30, S=24, seed=45, nan_rate=0.05),
     dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", minsamp=10), 0, 435),
    ("pearson_none_S24", dict(kind="latent", G=30, S=24, seed=46, nan_rate=0.10),
     dict(clusmethod="none", corrmethod="pearson", minsamp=10), 0, 435),
    ("kmeans_pearson_maxclus8", dict(kind="modules", G=20, S=300, seed=47, nan_rate=0.05),
     dict(clusmethod="kmeans", corrmethod="pearson", crit="BIC", maxclus=8), 0, 190),
]


def _gem(kind, G, S, seed, **kw):
    if kind == "latent":
        return gemgen.latent(G, S, seed, kw.get("nan_rate", 0.05))
    return gemgen.modules(G, S, seed, nan_rate=kw.get("nan_rate", 0.1),
                          degenerate=kw.get("degenerate", 0.0))


_ORACLE = {}


def _oracle_block(name, p, x, start, count):
    """Oracle output of a case, computed once for both execution forms."""
    if name not in _ORACLE:
        _ORACLE[name] = oracle.execute_block(p, x, start, count)
    return _ORACLE[name]


@pytest.mark.parametrize("form", ["plain", "persistent"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_block_parity(case, form):
    name, gk, flags, start, count = case
    x = _gem(**gk)
    G, S = x.shape
    count = min(count, _native.pair_count(G) - start)
    p = _native.make_params(form=form, ksize=1000, **flags)
    want = _oracle_block(name, p, x, start, count)
    eng = _native.Engine(x, p)
    try:
        got = eng.execute(start, count)
    finally:
        eng.close()
    print("case", name, form, "pairs", count, "K hist", np.bincount(want[0].astype(np.int64) + 0))
    assert_block_equal(got, want, S, where="%s/%s" % (name, form))


def _with_infinities(x, seed):
    """Expression data as a log transform of zero counts leaves it: a fifth
    of the genes hold a few -inf entries, one gene a +inf, and every seventh
    gene is constant."""
    rng = np.random.default_rng(seed)
    x = x.copy()
    G, S = x.shape
    for g in range(1, G, 5):
        x[g, rng.choice(S, 3, replace=False)] = -np.inf
    x[2, rng.integers(0, S)] = np.inf
    x[::7] = 4.0
    return x


@pytest.mark.parametrize("form", ["plain", "persistent"])
@pytest.mark.parametrize("flags", [
    dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", preout=True, postout=True),
    dict(clusmethod="gmm", corrmethod="pearson", crit="ICL"),
    dict(clusmethod="kmeans", corrmethod="pearson", crit="BIC", postout=True),
    dict(clusmethod="none", corrmethod="pearson"),
    dict(clusmethod="none", corrmethod="spearman"),
], ids=["gmm_spearman", "gmm_pearson_icl", "kmeans_pearson", "none_pearson", "none_spearman"])
def test_parity_with_infinities_and_constant_genes(flags, form):
    x = _with_infinities(gemgen.modules(28, 180, 48, nan_rate=0.05), 49)
    G, S = x.shape
    count = _native.pair_count(G)
    p = _native.make_params(form=form, ksize=1000, **flags)
    want = oracle.execute_block(p, x, 0, count)
    eng = _native.Engine(x, p)
    try:
        got = eng.execute(0, count)
    finally:
        eng.close()
    print("infinities", flags, form, "pairs", count, flush=True)
    assert_block_equal(got, want, S, where="infinities %s/%s" % (flags, form))
