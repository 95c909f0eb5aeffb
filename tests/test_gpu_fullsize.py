"""The 20,000-gene matrices at full size on the GPU: whole 32,768-pair work
blocks, including the ragged last block that ends at the largest pair index
(19999, 19998), checked through properties that do not need the CPU oracle to
finish the block:

  - the persistent and plain forms agree bit for bit;
  - sample masks: label -9 exactly where either gene's value is NaN
    (fetchPair, similarity_serial.cpp:108-141), read off the matrix itself
    through the pair index (Pairwise::Index, pairwise_index.cpp:13-107);
  - K lies in {0} or [minclus, maxclus]; labels name only clusters below K
    (cluster 0 at K = 0 and 1) or the markers -6..-9, with no
    post-clustering outlier (-8) at K = 1;
  - correlation k is NaN for k >= K and, below K, exactly when the cluster
    has fewer than minsamp members; |r| <= 1 otherwise;
  - for Spearman, whose pairs are independent, a sample of the block's pairs
    through the oracle, bit-exact.
Pearson pairs read the block's shared X workspace, so a sample of them is not
independent of the pairs between; Pearson parity over whole blocks is in the
smaller parity cases."""
import numpy as np
import pytest

import oracle
from kinc_amd import _native, gemgen
from parity_util import assert_block_equal

pytestmark = pytest.mark.gpu
BLOCK = 32768
_GEM = {}


def _gem(name):
    if name not in _GEM:
        _GEM.clear()
        _GEM[name] = gemgen.make(name)
    return _GEM[name]


def _pairs_of(start, count):
    k = np.arange(start, start + count, dtype=np.int64)
    x = ((1 + np.sqrt(1 + 8 * k.astype(np.float64))) // 2).astype(np.int64)
    x -= (x * (x - 1) // 2 > k)
    x += ((x + 1) * x // 2 <= k)
    y = k - x * (x - 1) // 2
    assert ((x * (x - 1) // 2 + y) == k).all() and (y >= 0).all() and (y < x).all()
    return x, y


def _run(x, p, start, count):
    eng = _native.Engine(x, p)
    try:
        return eng.execute(start, count)
    finally:
        eng.close()


def _check_invariants(x, start, got, p):
    K, L, C, _ = got
    count = K.shape[0]
    gx, gy = _pairs_of(start, count)
    nan = np.isnan(x)
    want_missing = nan[gx] | nan[gy]
    assert ((L == -9) == want_missing).all(), "sample masks differ from the matrix's NaN"
    MC, mn, minsamp = p.max_clusters, p.min_clusters, p.min_samples
    Ki = K.astype(np.int64)
    assert ((Ki == 0) | ((Ki >= mn) & (Ki <= MC))).all()
    Li = L.astype(np.int64)
    assert ((Li >= -9) & (Li < np.maximum(Ki, 1)[:, None])).all()
    assert not ((Li == -8) & (Ki == 1)[:, None]).any()
    for k in range(MC):
        members = (Li == k).sum(axis=1)
        c = C[:, k]
        above = Ki <= k
        assert np.isnan(c[above]).all()
        below = ~above
        assert (np.isnan(c[below]) == (members[below] < minsamp)).all()
        fin = np.isfinite(c)
        assert (np.abs(c[fin]) <= 1.0).all()
    return gx, gy


@pytest.mark.parametrize("block", [3000, 6103])
def test_headline_gmm_spearman_full_blocks(block):
    """20,000 x 500, GMM + Spearman with outlier removal: a middle block and
    the last, ragged block (6,896 pairs)."""
    x = _gem("H")
    G, S = x.shape
    total = _native.pair_count(G)
    start = block * BLOCK
    count = min(BLOCK, total - start)
    print("case H block", block, "pairs", count, flush=True)
    flags = dict(gemgen.CONFIGS["H"]["params"])
    got = _run(x, _native.make_params(form="persistent", **flags), start, count)
    plain = _run(x, _native.make_params(form="plain", **flags), start, count)
    assert_block_equal(plain, got, S, where="H block %d plain/persistent" % block)
    gx, gy = _check_invariants(x, start, got, _native.make_params(**flags))
    if block == 6103:
        assert count == total - 6103 * BLOCK and (gx[-1], gy[-1]) == (G - 1, G - 2)
    K = got[0]
    assert (K > 1).sum() > count // 20
    # a sample of the block, including its first and last pair
    idx = np.unique(np.concatenate([[0, count - 1], np.linspace(0, count - 1, 46).astype(int)]))
    pairs = np.stack([gx[idx], gy[idx]], axis=1).astype(np.int32)
    want = oracle.execute_pairs(_native.make_params(**flags), x, pairs)
    sub = tuple(a[idx] for a in got)
    assert_block_equal(sub, want, S, where="H block %d sample" % block)


def test_uneven_kmeans_pearson_full_block():
    """20,000 x 1,000 with per-gene NaN rates from Beta(1, 4) (valid samples
    per pair from about 30 to 1,000), K-means + Pearson: one whole block."""
    x = _gem("C")
    G, S = x.shape
    start = 2500 * BLOCK
    print("case C block 2500 pairs", BLOCK, flush=True)
    flags = dict(gemgen.CONFIGS["C"]["params"])
    got = _run(x, _native.make_params(form="persistent", **flags), start, BLOCK)
    plain = _run(x, _native.make_params(form="plain", **flags), start, BLOCK)
    assert_block_equal(plain, got, S, where="C block plain/persistent")
    _check_invariants(x, start, got, _native.make_params(**flags))
    N = S - (got[1] == -9).sum(axis=1)
    assert N.min() < 200 and N.max() > 800
