"""max_clusters 9..64 (the reference's limit, src/core/similarity_input.cpp:
180-189; MAX_CLUSTER_SIZE = 64, src/core/pairwise_index.h:36) on the
run-time-K memory-store instance, bit for bit against the oracle: K, labels,
correlations and per-fit iteration records, both execution forms, GMM and
K-means, BIC and ICL, Pearson (stale-workspace reads, clusters in groups of
twelve) and Spearman.  With 63-64 components on ~150 points most fits fail
(singular covariances) or leave clusters empty, which must also match."""
import numpy as np
import pytest

import oracle
from kinc_amd import _native, gemgen
from parity_util import assert_block_equal

pytestmark = pytest.mark.gpu

CASES = [
    # name, genes, samples, NaN rate, params, start, count
    ("gmm_spearman_k1_12", 40, 120, 0.10,
     dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", minclus=1, maxclus=12,
          preout=True, postout=True), 0, 200),
    ("gmm_pearson_icl_k9_20", 36, 150, 0.15,
     dict(clusmethod="gmm", corrmethod="pearson", crit="ICL", minclus=9, maxclus=20), 30, 150),
    ("kmeans_pearson_k9_16", 40, 200, 0.10,
     dict(clusmethod="kmeans", corrmethod="pearson", crit="BIC", minclus=9, maxclus=16), 100, 300),
    ("kmeans_spearman_k60_64", 24, 180, 0.10,
     dict(clusmethod="kmeans", corrmethod="spearman", crit="BIC", minclus=60, maxclus=64), 0, 60),
    ("gmm_pearson_k62_64", 20, 160, 0.05,
     dict(clusmethod="gmm", corrmethod="pearson", crit="BIC", minclus=62, maxclus=64), 50, 40),
]


@pytest.mark.parametrize("form", ["persistent", "plain"])
@pytest.mark.parametrize("name,G,S,nan,flags,start,count", CASES, ids=[c[0] for c in CASES])
def test_maxclus_up_to_64_matches_oracle(name, G, S, nan, flags, start, count, form):
    print("maxclus case %s %s" % (name, form), flush=True)
    x = gemgen.modules(G, S, seed=51, nan_rate=nan)
    p = _native.make_params(form=form, ksize=64, **flags)
    eng = _native.Engine(x, p)
    try:
        got = eng.execute(start, count)
    finally:
        eng.close()
    want = oracle.execute_block_threads(p, x, start, count, threads=8)
    assert_block_equal(got, want, S, name)
    if flags["minclus"] < 12:
        assert np.count_nonzero(got[0] > 1) > 0
