"""RMT on the GPU (include/kinc_rmt.h) against the oracle
(oracle/kinc_oracle_rmt.c): pruned sizes and thresholds exactly, eigenvalues
to within 2 float ulps (rocSOLVER's dsyevd and the oracle's Jacobi sweeps
are different double-precision solvers; both are rounded to float as the
reference rounds GSL's result), chi-squares within the effect of such a
rounding difference, and the same final threshold or the same failure."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from kinc_amd import formats, rmt
from test_rmt_oracle import modular_corr, pairs_and_block, random_corr

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def close_float(got, want):
    got = np.asarray(got, np.float32)
    want = np.asarray(want, np.float32)
    tol = 2 * np.spacing(np.maximum(np.abs(want), np.float32(1e-3)))
    return got.shape == want.shape and bool(np.all(np.abs(got - want) <= tol))


@pytest.mark.parametrize("case", ["random3", "pairs_block"])
def test_eigenvalues_match_oracle(case):
    m = random_corr(80, 3, 7, density=0.15) if case == "random3" else pairs_and_block()
    print("case", case, m.shape, flush=True)
    h = rmt.Rmt.from_dense(m)
    mx = oracle.rmt_maximums(m)
    for t in (0.99, 0.9, 0.75, 0.6, 0.541, 0.540, 0.5):
        t = np.float32(t)
        idx, pr = oracle.rmt_prune(m, mx, t)
        got = h.eigenvalues(float(t))
        assert got.size == idx.size, (t, got.size, idx.size)
        if idx.size:
            want = oracle.rmt_eigenvalues(pr)
            assert close_float(got, want), (t, np.abs(got - want).max())
    h.close()


def compare_runs(got, want):
    rc_g, t_g, rows_g = got
    rc_w, t_w, rows_w = want
    assert len(rows_g) == len(rows_w)
    for (tg, sg, cg), (tw, sw, cw) in zip(rows_g, rows_w):
        assert tg == tw and sg == sw
        # identical float eigenvalues give identical chi; a one-ulp difference
        # can move an unfolded spacing across a histogram bin edge
        assert (cg == cw) or (cg > 0 and cw > 0 and abs(cg - cw) <= 0.05 * cw), (tg, cg, cw)


def test_run_matches_oracle():
    m = pairs_and_block()
    h = rmt.Rmt.from_dense(m)
    rc, t, rows = h.run()
    assert rc == 0, h.last_error()
    orc, ot, orows = oracle.rmt_run(m)
    assert orc == len(orows)
    compare_runs((rc, t, rows), (orc, ot, orows))
    assert t == ot
    h.close()


def test_run_reports_stop_threshold_like_oracle():
    m = pairs_and_block()
    h = rmt.Rmt.from_dense(m, tstop=0.6)
    rc, t, rows = h.run()
    assert rc != 0 and "stopping threshold" in h.last_error()
    orc, ot, orows = oracle.rmt_run(m, stop=0.6)
    assert orc == -1
    compare_runs((rc, t, rows), (orc, ot, orows))
    h.close()


def test_run_reports_failing_spline_knots_like_oracle():
    m = modular_corr(300, 10, 5)
    h = rmt.Rmt.from_dense(m)
    rc, t, rows = h.run()
    assert rc != 0 and "spline knots" in h.last_error()
    orc, ot, orows = oracle.rmt_run(m)
    assert orc == -2
    compare_runs((rc, t, rows), (orc, ot, orows))
    h.close()


def write_cmx(path, m):
    """The dense matrix as a CMX file, through the Similarity writer: one
    cluster per pair holding r (all |r| >= 0.5 here)."""
    G = m.shape[0]
    n = G * (G - 1) // 2
    xs, ys = np.tril_indices(G, -1)
    order = np.lexsort((ys, xs))
    r = m[xs[order], ys[order], 0]
    K = (r != 0).astype(np.int8)
    corr = r.reshape(n, 1).astype(np.float32)
    w = formats.SimilarityWriter(path + ".ccm", path, G, 2, 1, corr_method="pearson",
                                 min_corr=0.5)
    labels = np.zeros((n, 2), np.int8)
    for s in range(0, n, 32768):
        e = min(n, s + 32768)
        w.process(s, K[s:e], labels[s:e], corr[s:e])
    w.close()


def test_cli_writes_reference_log(tmp_path):
    m = pairs_and_block()
    cmx = str(tmp_path / "p.cmx")
    write_cmx(cmx, m)
    log = str(tmp_path / "p.log")
    exe = os.path.join(ROOT, "kinc_amd", "kinc")
    r = subprocess.run([exe, "run", "rmt", "--input", cmx, "--log", log], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = open(log).read().splitlines()
    orc, ot, orows = oracle.rmt_run(m)
    assert len(lines) == len(orows) + 1
    for line, (t, size, chi) in zip(lines, orows):
        f = line.split("\t")
        assert f[0] == "%g" % t and f[1] == str(size)
        assert abs(float(f[2]) - chi) <= 0.05 * abs(chi) + 1e-6
    assert lines[-1] == "%g" % ot
