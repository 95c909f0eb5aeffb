"""Parity of the HIP path with the oracle over randomly drawn parameter sets.

Each case draws (from a fixed seed) the clustering and correlation methods,
criterion, outlier switches, expression floor, minimum samples, cluster
range, sample count (on both sides of the register range and the 5-wave LDS
limit), NaN rate and execution form, and requires bit-exact K, labels,
correlations and iteration records on a block of a small matrix, as the
fixed cases in test_gpu_parity.py do."""
import numpy as np
import pytest

import oracle
from kinc_amd import _native, gemgen
from parity_util import assert_block_equal

pytestmark = pytest.mark.gpu

N_CASES = 64


def _draw(i):
    rng = np.random.default_rng(1000 + i)
    clus = ["none", "gmm", "kmeans"][int(rng.integers(0, 3))]
    corr = ["pearson", "spearman"][int(rng.integers(0, 2))]
    S = int(rng.choice([20, 48, 64, 97, 150, 256, 300, 420, 513, 700, 1030, 1400]))
    maxclus = int(rng.choice([1, 2, 3, 5, 5, 5, 8]))
    if maxclus == 8 and S > 700:
        maxclus = 5
    flags = dict(clusmethod=clus, corrmethod=corr,
                 crit=["BIC", "ICL"][int(rng.integers(0, 2))],
                 preout=bool(rng.integers(0, 2)), postout=bool(rng.integers(0, 2)),
                 minsamp=int(rng.choice([5, 10, 30, max(5, S // 2)])),
                 minclus=int(rng.integers(1, maxclus + 1)), maxclus=maxclus,
                 pearson_fp64=bool(rng.random() < 0.2))
    if rng.random() < 0.3:
        flags["minexpr"] = float(rng.uniform(6.0, 8.5))
    nan_rate = float(rng.choice([0.0, 0.05, 0.15, 0.3]))
    budget = 40000 if clus == "none" else 900
    G = int(max(4, min(60, np.sqrt(2 * budget / max(1.0, S / 100.0)))))
    kind = "latent" if clus == "none" and rng.random() < 0.5 else "modules"
    form = ["plain", "persistent"][int(rng.integers(0, 2))]
    seed = int(rng.integers(0, 1 << 30))
    return flags, S, G, nan_rate, kind, form, seed


@pytest.mark.parametrize("i", range(N_CASES))
def test_random_parity(i):
    flags, S, G, nan_rate, kind, form, seed = _draw(i)
    if kind == "latent":
        x = gemgen.latent(G, S, seed, nan_rate)
    else:
        x = gemgen.modules(G, S, seed, nan_rate=nan_rate)
    n = _native.pair_count(G)
    start = n // 7
    count = min(n - start, 200 if flags["clusmethod"] != "none" else 4000)
    if flags["clusmethod"] != "none" and S > 1024:
        count = min(count, 40)
    p = _native.make_params(form=form, ksize=97, **flags)
    print("case", i, form, "G", G, "S", S, "nan", nan_rate, flags, "pairs", count, flush=True)
    want = oracle.execute_block(p, x, start, count)
    eng = _native.Engine(x, p)
    try:
        got = eng.execute(start, count)
    finally:
        eng.close()
    assert_block_equal(got, want, S, where="random case %d" % i)
