"""EMX / CCM / CMX formats and the Similarity writer (CPU only).

Mirrors the reference's three enabled format tests (write random data, read
it back: tests/testexpressionmatrix.cpp:10-62, testclustermatrix.cpp:10-104,
testcorrelationmatrix.cpp:10-89) and pins Similarity::process
(src/core/similarity.cpp:66-127) with a hand-computed byte fixture and an
independent Python model.
"""
import json
import os

import numpy as np
import pytest

from kinc_amd import _native, formats

HERE = os.path.dirname(os.path.abspath(__file__))


def test_emx_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    x = rng.normal(8, 2, (10, 5)).astype(np.float32)
    x[rng.random(x.shape) < 0.2] = np.nan
    p = str(tmp_path / "a.emx")
    formats.write_emx(p, x, genes=["g%d" % i for i in range(10)], transform="log2")
    y = formats.read_emx_native(p)
    assert x.tobytes() == y.tobytes()  # bit-exact, NaN included
    z, meta = formats.read_emx(p)
    assert x.tobytes() == z.tobytes()
    assert meta["genes"][3] == "g3" and meta["transform"] == "log2" and len(meta["samples"]) == 5


def _hex(s):
    return bytes.fromhex(s.replace(" ", ""))


def test_writer_matches_hand_computed_bytes(tmp_path):
    with open(os.path.join(HERE, "golden", "ccm_cmx_bytes.json")) as f:
        fx = json.load(f)
    S, MC = fx["samples"], fx["max_clusters"]
    corr = np.array([[np.nan if v == "nan" else v for v in row] for row in fx["corr"]], np.float32)
    ccm, cmx = str(tmp_path / "o.ccm"), str(tmp_path / "o.cmx")
    w = formats.SimilarityWriter(ccm, cmx, fx["genes"], S, MC, metadata={"genes": ["a", "b", "c"]})
    w.process(0, np.array(fx["K"], np.int8), np.array(fx["labels"], np.int8), corr)
    w.close()
    typ, payload, meta = formats.read_file(ccm)
    want = _hex(fx["ccm_header_hex"]) + b"".join(_hex(r) for r in fx["ccm_records_hex"])
    assert typ == "ccm" and payload == want
    typ, payload, meta = formats.read_file(cmx)
    want = _hex(fx["cmx_header_hex"]) + b"".join(_hex(r) for r in fx["cmx_records_hex"])
    assert typ == "cmx" and payload == want
    assert meta["correlations"] == ["pearson"] and meta["genes"] == ["a", "b", "c"]


def _model_process(start, K, labels, corr, S, mincorr, maxcorr):
    """Independent model of Similarity::process: lists of CCM and CMX records."""
    ccm, cmx = [], []
    for q in range(K.shape[0]):
        x, y = _native.pair_index(start + q)
        keep = [k for k in range(K[q]) if not np.isnan(corr[q, k])
                and mincorr <= abs(corr[q, k]) <= maxcorr]
        if K[q] > 1:
            for c, k in enumerate(keep):
                m = [int(k == l) if l >= 0 else -int(l) for l in labels[q]]
                ccm.append((x, y, c, m))
        if K[q] > 0:
            for c, k in enumerate(keep):
                cmx.append((x, y, c, corr[q, k]))
    return ccm, cmx


@pytest.mark.parametrize("seed", range(5))
def test_writer_random_round_trip(tmp_path, seed):
    rng = np.random.default_rng(seed)
    G, S, MC = 12, 7, 5
    n = G * (G - 1) // 2
    K = rng.integers(0, MC + 1, n).astype(np.int8)
    labels = rng.choice(np.array([-9, -8, -7, -6, -1, 0, 1, 2, 3, 4], np.int8), (n, S))
    corr = rng.uniform(-1, 1, (n, MC)).astype(np.float32)
    corr[rng.random(corr.shape) < 0.1] = np.nan
    ccm, cmx = str(tmp_path / "r.ccm"), str(tmp_path / "r.cmx")
    w = formats.SimilarityWriter(ccm, cmx, G, S, MC, corr_method="spearman", min_corr=0.3)
    split = n // 3
    w.process(0, K[:split], labels[:split], corr[:split])
    w.process(split, K[split:], labels[split:], corr[split:])
    w.close()
    want_ccm, want_cmx = _model_process(0, K, labels, corr, S, np.float32(0.3), 1.0)
    hdr, rec = formats.read_pairwise(ccm)
    assert hdr["clusters"] == len(want_ccm) and hdr["samples"] == S
    assert hdr["pairs"] == len({(a, b) for a, b, _, _ in want_ccm})
    masks = formats.ccm_masks(rec, S)
    for i, (x, y, c, m) in enumerate(want_ccm):
        assert (rec["x"][i], rec["y"][i], rec["cluster"][i]) == (x, y, c)
        assert masks[i].tolist() == m
    hdr, rec = formats.read_pairwise(cmx)
    assert hdr["clusters"] == len(want_cmx) and hdr["meta"]["correlations"] == ["spearman"]
    vals = formats.cmx_values(rec)
    for i, (x, y, c, v) in enumerate(want_cmx):
        assert (rec["x"][i], rec["y"][i], rec["cluster"][i]) == (x, y, c)
        assert vals[i] == v


def test_writer_refuses_out_of_order_blocks(tmp_path):
    # Matrix::write refuses a non-increasing indent (pairwise_matrix.cpp:133-142)
    ccm, cmx = str(tmp_path / "o.ccm"), str(tmp_path / "o.cmx")
    w = formats.SimilarityWriter(ccm, cmx, 5, 4, 3)
    K = np.array([1, 1], np.int8)
    lab = np.zeros((2, 4), np.int8)
    corr = np.full((2, 3), 0.9, np.float32)
    w.process(4, K, lab, corr)
    with pytest.raises(_native.KincError):
        w.process(2, K, lab, corr)
    w.close()


def test_writer_rejects_bad_max_clusters(tmp_path):
    with pytest.raises(_native.KincError):
        formats.SimilarityWriter(str(tmp_path / "a"), str(tmp_path / "b"), 5, 4, 65)


def _check_files(ccm, cmx, want_ccm, want_cmx, S):
    hdr, rec = formats.read_pairwise(ccm)
    assert hdr["clusters"] == len(want_ccm)
    assert hdr["pairs"] == len({(a, b) for a, b, _, _ in want_ccm})
    masks = formats.ccm_masks(rec, S)
    got = list(zip(rec["x"].tolist(), rec["y"].tolist(), rec["cluster"].tolist()))
    assert got == [(x, y, c) for x, y, c, _ in want_ccm]
    assert masks.tolist() == [m for _, _, _, m in want_ccm]
    hdr, rec = formats.read_pairwise(cmx)
    assert hdr["clusters"] == len(want_cmx)
    assert hdr["pairs"] == len({(a, b) for a, b, _, _ in want_cmx})
    got = list(zip(rec["x"].tolist(), rec["y"].tolist(), rec["cluster"].tolist()))
    assert got == [(x, y, c) for x, y, c, _ in want_cmx]
    assert formats.cmx_values(rec).view(np.uint32).tolist() == \
        np.array([v for _, _, _, v in want_cmx], np.float32).view(np.uint32).tolist()


def _random_block(rng, n, S, MC):
    K = rng.integers(0, MC + 1, n).astype(np.int8)
    labels = rng.choice(np.array([-9, -8, -7, -6, -1, 0, 1, 2, 3, 4, 5, 6, 7], np.int8), (n, S))
    corr = rng.uniform(-1, 1, (n, MC)).astype(np.float32)
    corr[rng.random(corr.shape) < 0.1] = np.nan
    return K, labels, corr


@pytest.mark.parametrize("S", [7, 8])
def test_writer_large_block_matches_model(tmp_path, S):
    # large enough that the records are encoded by several threads; an odd
    # sample count leaves a half-filled last mask byte
    rng = np.random.default_rng(S)
    G, MC = 130, 8
    n = G * (G - 1) // 2
    K, labels, corr = _random_block(rng, n, S, MC)
    ccm, cmx = str(tmp_path / "l.ccm"), str(tmp_path / "l.cmx")
    w = formats.SimilarityWriter(ccm, cmx, G, S, MC, min_corr=0.4, max_corr=0.95)
    w.process(0, K, labels, corr)
    w.close()
    want_ccm, want_cmx = _model_process(0, K, labels, corr, S, np.float32(0.4), np.float32(0.95))
    _check_files(ccm, cmx, want_ccm, want_cmx, S)


def test_writer_stops_at_the_first_bad_pair(tmp_path):
    # a pair with K > max_clusters: every record before it is written, none after
    rng = np.random.default_rng(9)
    G, S, MC = 130, 6, 4
    n = G * (G - 1) // 2
    K, labels, corr = _random_block(rng, n, S, MC)
    labels = np.where(labels > 3, 0, labels).astype(np.int8)
    bad = 5000
    K[bad] = MC + 1
    K[bad + 300] = MC + 2
    ccm, cmx = str(tmp_path / "b.ccm"), str(tmp_path / "b.cmx")
    w = formats.SimilarityWriter(ccm, cmx, G, S, MC, min_corr=0.2)
    with pytest.raises(_native.KincError) as e:
        w.process(0, K, labels, corr)
    assert "cluster size exceeding the max" in str(e.value)
    w.close()
    want_ccm, want_cmx = _model_process(0, K[:bad], labels[:bad], corr[:bad], S,
                                        np.float32(0.2), 1.0)
    _check_files(ccm, cmx, want_ccm, want_cmx, S)


def test_writer_single_cluster_runs_over_many_calls(tmp_path):
    # long runs of one-cluster pairs (the writer's branch-free path) between
    # multi-cluster and empty pairs, split over calls of very different sizes
    # (one encoding thread, then the pool, then one again), with
    # correlations on the |r| bounds, NaN and zero
    rng = np.random.default_rng(21)
    G, S, MC = 180, 9, 5
    n = G * (G - 1) // 2
    K, labels, corr = _random_block(rng, n, S, MC)
    labels = np.where(labels >= MC, 0, labels).astype(np.int8)
    runs = rng.random(n) < 0.8
    K[runs] = 1
    lo, hi = np.float32(0.3), np.float32(0.9)
    edge = rng.choice(np.array([0.3, -0.3, 0.9, -0.9, 0.0, np.nan], np.float32), n)
    pick = rng.random(n) < 0.05
    corr[pick, 0] = edge[pick]
    ccm, cmx = str(tmp_path / "r.ccm"), str(tmp_path / "r.cmx")
    w = formats.SimilarityWriter(ccm, cmx, G, S, MC, min_corr=float(lo), max_corr=float(hi))
    sizes = [1, 7, 2048, 9000, 3, 4096, 130, 20000]
    s = 0
    for sz in sizes:
        e = min(n, s + sz)
        w.process(s, K[s:e], labels[s:e], corr[s:e])
        s = e
    w.process(s, K[s:], labels[s:], corr[s:])
    w.close()
    want_ccm, want_cmx = _model_process(0, K, labels, corr, S, lo, hi)
    _check_files(ccm, cmx, want_ccm, want_cmx, S)
    assert len(want_cmx) > 5000
