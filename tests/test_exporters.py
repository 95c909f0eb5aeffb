"""CMX text export (ExportCorrelationMatrix::process,
src/core/exportcorrelationmatrix.cpp:17-117) and network extraction
(Extract::process, src/core/extract.cpp:14-275) against an independent
Python reading of the CCM/CMX/EMX files (kinc_amd.formats' numpy parsers)
and the reference's line layout."""
import os
import subprocess

import numpy as np
import pytest

from kinc_amd import formats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def similarity_files(tmp_path, seed=0, G=14, S=9, MC=4, min_corr=0.3):
    """EMX + CCM + CMX of random Similarity results (through the writer)."""
    rng = np.random.default_rng(seed)
    x = rng.normal(8, 1, (G, S)).astype(np.float32)
    x[rng.random((G, S)) < 0.15] = np.nan
    genes = ["gene_%d" % g for g in range(G)]
    samples = ["s%d" % i for i in range(S)]
    emx = str(tmp_path / "a.emx")
    formats.write_emx(emx, x, genes=genes, samples=samples)
    n = G * (G - 1) // 2
    K = rng.integers(0, MC + 1, n).astype(np.int8)
    labels = rng.choice(np.array([-9, -8, -7, -6, 0, 1, 2, 3], np.int8), (n, S))
    corr = rng.uniform(-1, 1, (n, MC)).astype(np.float32)
    ccm, cmx = str(tmp_path / "a.ccm"), str(tmp_path / "a.cmx")
    w = formats.SimilarityWriter(ccm, cmx, G, S, MC, min_corr=min_corr,
                                 metadata={"genes": genes, "samples": samples})
    w.process(0, K, labels, corr)
    w.close()
    return emx, ccm, cmx, x, genes


def model_clusters(ccm, cmx, S):
    """(x, y, k, n, counts, corr, mask) per CMX cluster, from the files."""
    _, crec = formats.read_pairwise(cmx)
    _, mrec = formats.read_pairwise(ccm)
    cvals = formats.cmx_values(crec)
    masks = formats.ccm_masks(mrec, S)
    mindex = {(int(a), int(b), int(c)): i for i, (a, b, c) in
              enumerate(zip(mrec["x"], mrec["y"], mrec["cluster"]))}
    pairs = {}
    for i in range(crec.size):
        pairs.setdefault((int(crec["x"][i]), int(crec["y"][i])), []).append(i)
    out = []
    for (a, b), idx in pairs.items():
        for k, i in enumerate(idx):
            assert int(crec["cluster"][i]) == k
            if len(idx) > 1:
                m = masks[mindex[(a, b, k)]]
                counts = [int((m == v).sum()) for v in (1, 9, 8, 7, 6)]
                text = "".join(chr(ord("0") + int(v)) for v in m)
            else:
                counts, text = [0] * 5, None
            out.append((a, b, k, len(idx), counts, float(cvals[i]), text))
    return out


def test_export_cmx_text(tmp_path):
    emx, ccm, cmx, x, genes = similarity_files(tmp_path)
    out = str(tmp_path / "a.txt")
    formats.export_cmx_text(ccm, cmx, out)
    S = x.shape[1]
    want = []
    for a, b, k, n, counts, c, text in model_clusters(ccm, cmx, S):
        want.append("\t".join([str(a), str(b), str(k), str(n)] + [str(v) for v in counts]
                              + ["%g" % np.float32(c), text if text else "0" * S]))
    got = open(out).read().split("\n")
    assert got[-1] == ""
    assert got[:-1] == want
    assert len(want) > 20


@pytest.mark.parametrize("mincorr,maxcorr", [(0.85, 1.0), (0.3, 0.9)])
def test_extract_network_and_graphml(tmp_path, mincorr, maxcorr):
    emx, ccm, cmx, x, genes = similarity_files(tmp_path, seed=1)
    out, gml = str(tmp_path / "net.txt"), str(tmp_path / "net.graphml")
    formats.extract_network(emx, ccm, cmx, out, gml, mincorr, maxcorr)
    S = x.shape[1]
    rows, edges = [], []
    for a, b, k, n, counts, c, text in model_clusters(ccm, cmx, S):
        c32 = np.float32(c)
        if abs(c32) < np.float32(mincorr) or np.float32(maxcorr) < abs(c32):
            continue
        if text is None:
            text = "".join("9" if (np.isnan(x[a, i]) or np.isnan(x[b, i])) else "1"
                           for i in range(S))
        rows.append("\t".join([genes[a], genes[b], "%g" % c32, "co", str(k), str(n)]
                              + [str(v) for v in counts] + [text]))
        edges.append('    <edge source="%s" target="%s" samples="%s"/>' % (genes[a], genes[b], text))
    lines = open(out).read().split("\n")
    assert lines[0] == ("Source\tTarget\tsc\tInteraction\tCluster\tNum_Clusters\tCluster_Samples\t"
                        "Missing_Samples\tCluster_Outliers\tPair_Outliers\tToo_Low\tSamples")
    assert lines[1:-1] == rows and lines[-1] == ""
    g = open(gml).read().split("\n")
    assert g[0] == '<?xml version="1.0" encoding="UTF-8"?>'
    nodes = [ln for ln in g if ln.startswith("    <node")]
    assert nodes == ['    <node id="%s"/>' % n for n in genes]
    assert [ln for ln in g if ln.startswith("    <edge")] == edges
    assert g[-3:] == ["  </graph>", "</graphml>", ""]
    assert len(rows) > 0


def test_cli_export_and_extract(tmp_path):
    exe = os.path.join(ROOT, "kinc_amd", "kinc")
    if not os.path.exists(exe):
        pytest.skip("kinc CLI not built")
    emx, ccm, cmx, x, genes = similarity_files(tmp_path, seed=2)
    a = str(tmp_path / "cli.txt")
    r = subprocess.run([exe, "run", "export-cmx", "--ccm", ccm, "--cmx", cmx, "--output", a],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    b = str(tmp_path / "lib.txt")
    formats.export_cmx_text(ccm, cmx, b)
    assert open(a).read() == open(b).read()
    c = str(tmp_path / "net.txt")
    r = subprocess.run([exe, "run", "extract", "--emx", emx, "--ccm", ccm, "--cmx", cmx,
                        "--output", c, "--mincorr", "0.5"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    d = str(tmp_path / "net2.txt")
    formats.extract_network(emx, ccm, cmx, d, None, 0.5, 1.0)
    assert open(c).read() == open(d).read()
