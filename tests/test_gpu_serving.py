"""Explicit pair lists, batching of work blocks, and the query server.

Results must be bit-exact with the oracle run on the same pair list with the
same work-block segmentation, and a query's answer must be identical whether
it runs alone or batched with other queries.
"""
import numpy as np
import pytest

import oracle
from kinc_amd import _native, gemgen, serving
from parity_util import assert_block_equal

pytestmark = pytest.mark.gpu


def _segments(n, rng, blocks=4):
    cuts = np.sort(rng.choice(np.arange(1, n), size=blocks - 1, replace=False))
    seg = np.zeros(n, np.int32)
    for c in cuts:
        seg[c:] = c
    return seg


@pytest.mark.parametrize("flags", [
    dict(clusmethod="none", corrmethod="pearson"),
    dict(clusmethod="gmm", corrmethod="pearson", crit="BIC"),
    dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", preout=True, postout=True),
], ids=["pearson", "gmm_pearson", "gmm_spearman"])
def test_pair_list_with_segments(flags):
    rng = np.random.default_rng(5)
    x = gemgen.modules(60, 220, seed=21, nan_rate=0.12)
    G = x.shape[0]
    n = 600 if flags["clusmethod"] == "none" else 160
    a = rng.integers(1, G, n)
    b = (rng.random(n) * a).astype(np.int64)
    pairs = np.stack([a, b], 1).astype(np.int32)
    seg = _segments(n, rng)
    p = _native.make_params(batch_capacity=65536, **flags)
    want = oracle.execute_pairs(p, x, pairs, seg)
    eng = _native.Engine(x, p)
    try:
        got = eng.execute_pairs(pairs, seg)
    finally:
        eng.close()
    print("pair list", flags, "pairs", n, "segments", len(np.unique(seg)), flush=True)
    assert_block_equal(got, want, x.shape[1], where="pair list %s" % flags)


def _oracle_hits(p, x, genes, min_abs):
    pairs = serving.query_pairs(genes, x.shape[0])
    n = pairs.shape[0]
    seg = (np.arange(n) // 32768 * 32768).astype(np.int32)
    K, _, corr, _ = oracle.execute_pairs(p, x, pairs, seg, want_labels=False)
    rows = []
    for q in range(n):
        c = 0
        for k in range(K[q]):
            r = corr[q, k]
            if not np.isnan(r) and min_abs <= abs(r) <= p.max_corr:
                rows.append((pairs[q, 0], pairs[q, 1], c, r))
                c += 1
    return np.array(rows, serving.HIT_DTYPE)


def _same_hits(a, b):
    return (a.shape == b.shape and np.array_equal(a["x"], b["x"]) and np.array_equal(a["y"], b["y"])
            and np.array_equal(a["cluster"], b["cluster"])
            and np.array_equal(a["corr"].view(np.uint32), b["corr"].view(np.uint32)))


def test_server_alone_equals_mix_and_oracle():
    x = gemgen.modules(400, 180, seed=22, nan_rate=0.08)
    p = _native.make_params(clusmethod="gmm", corrmethod="pearson", crit="BIC")
    srv = serving.Server(x, p, hbm_budget=64 << 20, batch_pairs=65536)
    try:
        q_small = [5]
        q_mid = [17, 3, 250]
        q_big = list(range(100, 220, 1))  # 120 genes: about 40k pairs, two work blocks
        # alone
        alone = {}
        for name, genes in (("small", q_small), ("mid", q_mid), ("big", q_big)):
            qid = srv.submit(genes, 0.5)
            assert srv.wait(qid, 300)
            alone[name] = srv.result(qid)
            print("alone", name, "pairs", alone[name][2], "hits", len(alone[name][0]),
                  "latency %.3fs" % alone[name][1], flush=True)
        # mixed: all in flight together, in a different order
        ids = {name: srv.submit(genes, 0.5)
               for name, genes in (("big", q_big), ("small", q_small), ("mid", q_mid))}
        for name, qid in ids.items():
            assert srv.wait(qid, 300)
            hits, lat, npairs = srv.result(qid)
            print("mix", name, "hits", len(hits), "latency %.3fs" % lat, flush=True)
            assert npairs == alone[name][2]
            assert _same_hits(hits, alone[name][0]), name
        pairs_done, batches, busy = srv.stats()
        assert pairs_done == 2 * sum(v[2] for v in alone.values())
    finally:
        srv.close()
    for name, genes in (("small", q_small), ("mid", q_mid)):
        want = _oracle_hits(p, x, genes, np.float32(0.5))
        assert _same_hits(alone[name][0], want), name


def test_server_admission_under_small_budget():
    """A 1 MiB budget holds one 60-gene query's results (about 34k pairs x
    21 bytes) at a time: the others wait for admission and are still served,
    and a query larger than the whole budget is refused."""
    x = gemgen.latent(600, 120, seed=23)
    p = _native.make_params(clusmethod="none", corrmethod="pearson")
    srv = serving.Server(x, p, hbm_budget=1 << 20, batch_pairs=32768)
    try:
        queries = [list(range(i, i + 60)) for i in (0, 200, 400)]
        ids = [srv.submit(q, 0.3) for q in queries]
        res = []
        for qid in ids:
            assert srv.wait(qid, 120)
            res.append(srv.result(qid))
        assert all(r[2] == 60 * 599 - 60 * 59 // 2 for r in res)
        with pytest.raises(_native.KincError):
            srv.submit(list(range(200)), 0.3)
    finally:
        srv.close()
    want = _oracle_hits(p, x, queries[1], np.float32(0.3))
    assert _same_hits(res[1][0], want)


def test_server_small_query_overtakes_large_ones():
    """Fewest remaining blocks first, one-block quanta while several queries
    have work, and admission that a query which does not fit cannot block:
    a 1-gene query submitted behind two large ones (the second of which does
    not fit in the budget until the first is done) finishes first, and every
    answer equals the query's answer alone."""
    import time
    x = gemgen.modules(3000, 160, seed=24, nan_rate=0.08)
    p = _native.make_params(clusmethod="gmm", corrmethod="pearson", crit="BIC")
    big1 = list(range(0, 150))       # about 439k pairs: 14 work blocks
    big2 = list(range(150, 300))
    small = [2500]
    # 21 result bytes per pair: one large query and the small one fit, not two large
    srv = serving.Server(x, p, hbm_budget=12 << 20, batch_pairs=32768)
    try:
        alone = {}
        for name, genes in (("big1", big1), ("big2", big2), ("small", small)):
            qid = srv.submit(genes, 0.5)
            assert srv.wait(qid, 300)
            alone[name] = srv.result(qid)
        t0 = time.perf_counter()
        ids = {"big1": srv.submit(big1, 0.5), "big2": srv.submit(big2, 0.5)}
        time.sleep(0.05)
        ids["small"] = srv.submit(small, 0.5)
        offs = {"big1": 0.0, "big2": 0.0, "small": time.perf_counter() - t0}
        done = {}
        for name, qid in ids.items():
            assert srv.wait(qid, 300)
            wait_s, run_s = srv.timing(qid)
            hits, lat, npairs = srv.result(qid)
            done[name] = offs[name] + lat
            print(name, "latency %.3f s, start wait %.3f s, run %.3f s" % (lat, wait_s, run_s),
                  flush=True)
            assert npairs == alone[name][2]
            assert _same_hits(hits, alone[name][0]), name
        assert done["small"] < done["big1"] < done["big2"]
    finally:
        srv.close()
