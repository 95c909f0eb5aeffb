"""`kinc run similarity` end to end on the GPU (the ACE driver loop of
Similarity::size/makeWork/execute/process, src/core/similarity.cpp:17-127):
the CCM and CMX it writes are byte-identical to the ones the same writer
produces from the oracle's results, over one block (GMM + Spearman) and over
several blocks (Pearson, where block boundaries matter)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from kinc_amd import _native, formats, gemgen

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "kinc_amd", "kinc")
BLOCK = 32768


def oracle_files(tmp_path, x, flags, meta):
    G, S = x.shape
    p = _native.make_params(**flags)
    n = G * (G - 1) // 2
    ccm, cmx = str(tmp_path / "o.ccm"), str(tmp_path / "o.cmx")
    w = formats.SimilarityWriter(ccm, cmx, G, S, p.max_clusters,
                                 corr_method=flags.get("corrmethod", "pearson"),
                                 min_corr=p.min_corr, max_corr=p.max_corr, metadata=meta)
    for s in range(0, n, BLOCK):
        cnt = min(BLOCK, n - s)
        K, L, C, _ = oracle.execute_block(p, x, s, cnt)
        w.process(s, K, L, C)
    w.close()
    return ccm, cmx


def run_cli(tmp_path, x, flags, form):
    emx = str(tmp_path / "in.emx")
    G, S = x.shape
    meta = {"genes": ["g%d" % g for g in range(G)], "samples": ["s%d" % i for i in range(S)],
            "transform": "none"}
    formats.write_emx(emx, x, genes=meta["genes"], samples=meta["samples"])
    ccm, cmx = str(tmp_path / "c.ccm"), str(tmp_path / "c.cmx")
    args = [EXE, "run", "similarity", "--input", emx, "--ccm", ccm, "--cmx", cmx, "--form", form]
    for k, v in flags.items():
        args += ["--" + k, str(v).lower() if isinstance(v, bool) else str(v)]
    print("cli", " ".join(args[2:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return ccm, cmx, meta


@pytest.mark.parametrize("form", ["persistent", "plain"])
def test_cli_gmm_spearman_one_block(tmp_path, form):
    x = gemgen.modules(40, 180, seed=21, nan_rate=0.1)
    flags = dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", preout=True, postout=True,
                 mincorr=0.3)
    ccm, cmx, meta = run_cli(tmp_path, x, flags, form)
    occm, ocmx = oracle_files(tmp_path, x, flags, meta)
    for a, b in ((ccm, occm), (cmx, ocmx)):
        ta, pa, ma = formats.read_file(a)
        tb, pb, mb = formats.read_file(b)
        assert ta == tb and pa == pb
        assert ma["genes"] == mb["genes"]
    hdr, rec = formats.read_pairwise(cmx)
    assert hdr["clusters"] > 10


def test_cli_pearson_several_blocks(tmp_path):
    # 70,125 pairs: three work blocks, Pearson reads each block's stale X
    x = gemgen.latent(375, 60, seed=5, nan_rate=0.2)
    flags = dict(clusmethod="none", corrmethod="pearson", minsamp=20, mincorr=0.2)
    ccm, cmx, meta = run_cli(tmp_path, x, flags, "persistent")
    occm, ocmx = oracle_files(tmp_path, x, flags, meta)
    for a, b in ((ccm, occm), (cmx, ocmx)):
        assert formats.read_file(a)[1] == formats.read_file(b)[1]
    hdr, rec = formats.read_pairwise(cmx)
    assert hdr["pairs"] > 1000


def test_cli_pearson_fp64_option(tmp_path):
    # --pearson-fp64 true reaches the engine: the files match the writer fed
    # with the oracle's FP64-Pearson results, and differ from the default's
    x = gemgen.latent(150, 120, seed=8, nan_rate=0.1)
    flags = dict(clusmethod="none", corrmethod="pearson", minsamp=20, mincorr=0.1)
    ccm, cmx, meta = run_cli(tmp_path, x, dict(flags, **{"pearson-fp64": True}), "persistent")
    occm, ocmx = oracle_files(tmp_path, x, dict(flags, pearson_fp64=True), meta)
    assert formats.read_file(cmx)[1] == formats.read_file(ocmx)[1]
    d32 = tmp_path / "d32"
    d32.mkdir()
    c32, x32, _ = run_cli(d32, x, flags, "persistent")
    assert formats.read_file(x32)[1] != formats.read_file(cmx)[1]


def test_cli_one_or_two_workers_same_files(tmp_path):
    # blocks taken in turn by two engine handles give the same files as one
    x = gemgen.modules(300, 90, seed=9, nan_rate=0.1)
    flags = dict(clusmethod="gmm", corrmethod="pearson", crit="BIC", mincorr=0.3)
    one = tmp_path / "one"
    one.mkdir()
    c1, x1, _ = run_cli(one, x, dict(flags, workers=1), "persistent")
    two = tmp_path / "two"
    two.mkdir()
    c2, x2, _ = run_cli(two, x, dict(flags, workers=2), "persistent")
    assert formats.read_file(c1)[1] == formats.read_file(c2)[1]
    assert formats.read_file(x1)[1] == formats.read_file(x2)[1]
    assert formats.read_pairwise(x1)[0]["pairs"] > 100
