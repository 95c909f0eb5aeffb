"""Configs C and D at full size on the GPU against the oracle, bit for bit.

* C: 20,000 genes x 1,000 samples with per-gene NaN rates drawn from
  Beta(1,4), so a pair's valid-sample count N runs from about 30 to 1,000;
  one whole 32,768-pair work block with GMM and with K-means (maxclus 5,
  BIC, Pearson), persistent form, every pair of the block compared.
* D: the 60,000 x 475 serving matrix (GMM + Pearson): a query gene against
  every other gene as an explicit pair list cut into two work blocks with
  segment starts (the serving loop's call), the first block whole and the
  head of the second compared.

The oracle runs on host threads in contiguous ranges; each range replays
the fetches of the earlier pairs of its work block, which is all the shared
X workspace of Pearson's stale reads depends on, so the results equal one
serial pass (tests/test_oracle.py checks that on the CPU).
"""
import os

import numpy as np
import pytest

import oracle
from kinc_amd import _native, gemgen
from parity_util import assert_block_equal

pytestmark = pytest.mark.gpu
BLOCK = 32768
_GEM = {}


def _threads():
    # the GPU box gives one GPU 16 host cores, whatever the OS reports
    return max(1, min(16, os.cpu_count() or 1))


def _gem(name):
    if name not in _GEM:
        _GEM.clear()
        _GEM[name] = gemgen.make(name)
    return _GEM[name]


@pytest.mark.parametrize("clus", ["gmm", "kmeans"])
def test_config_c_whole_block(clus):
    x = _gem("C")
    S = x.shape[1]
    p = _native.make_params(clusmethod=clus, corrmethod="pearson", maxclus=5, crit="BIC",
                            minsamp=30)
    start = 97 * BLOCK
    print("config C %s block %d" % (clus, start // BLOCK), flush=True)
    eng = _native.Engine(x, p)
    try:
        got = eng.execute(start, BLOCK)
    finally:
        eng.close()
    K = got[0]
    # the block spans the sample-count range the binned queue is for
    Ns = (~np.isnan(x[oracle.block_pairs(start, BLOCK)[::16]])).all(axis=1).sum(axis=1)
    assert Ns.min() < 200 and Ns.max() > 800
    want = oracle.execute_block_threads(p, x, start, BLOCK, threads=_threads())
    assert_block_equal(got, want, S, "config C %s" % clus)
    assert np.count_nonzero(K > 1) > 0


def test_config_d_query_blocks():
    x = _gem("D")
    G, S = x.shape
    p = _native.make_params(clusmethod="gmm", corrmethod="pearson", maxclus=5, crit="BIC",
                            minsamp=30, batch_capacity=65536)
    q = 12345
    g = np.array([v for v in range(G) if v != q], np.int64)
    pairs = np.stack([np.maximum(g, q), np.minimum(g, q)], 1).astype(np.int32)
    n = pairs.shape[0]
    seg = np.where(np.arange(n) < BLOCK, 0, BLOCK).astype(np.int32)
    print("config D query gene %d: %d pairs in 2 work blocks" % (q, n), flush=True)
    eng = _native.Engine(x, p)
    try:
        got = eng.execute_pairs(pairs, seg)
    finally:
        eng.close()
    head = BLOCK + 4096
    want = oracle.execute_pairs_threads(p, x, pairs[:head], seg[:head], threads=_threads())
    got_head = tuple(a[:head] for a in got)
    assert_block_equal(got_head, want, S, "config D query")
    # the second block reads its own workspace, not the first block's
    assert (seg[BLOCK:] == BLOCK).all()
