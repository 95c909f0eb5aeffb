"""The FP64 Pearson option (kinc_sim_params.pearson_fp64 = 1, which the
reference lacks: its Pearson, src/core/pairwise_pearson.cpp:12-52, sums the
five moments in FP32) in the oracle, pinned against numpy's float64
correlation of the same samples, and the FP64 error reference used by the
bench's error report pinned the same way.

Parity of the device with the oracle in this mode is in the GPU parity cases
whose names end in "fp64"."""
import numpy as np
import pytest

import oracle
from kinc_amd import _native, gemgen


def _full_block(G, S, seed, mean, sd):
    """No NaN and no expression floor: every pair's X is both genes in full
    and its labels all 0, so the reference's full-index read of the X
    workspace sees exactly the pair's samples."""
    rng = np.random.default_rng(seed)
    base = rng.normal(0, 1, (4, S))
    w = rng.normal(0, 1, (G, 4))
    x = mean + sd * (w @ base + 0.5 * rng.normal(0, 1, (G, S)))
    return x.astype(np.float32)


def _numpy_r(x):
    G = x.shape[0]
    xd = x.astype(np.float64)
    out = []
    for g1 in range(1, G):
        for g2 in range(g1):
            out.append(np.corrcoef(xd[g1], xd[g2])[0, 1])
    return np.array(out)


@pytest.mark.parametrize("mean,sd", [(0.0, 1.0), (10.0, 1.0), (1000.0, 1.0)])
def test_fp64_option_matches_numpy(mean, sd):
    x = _full_block(24, 180, 3, mean, sd)
    n = 24 * 23 // 2
    want = _numpy_r(x)
    p64 = _native.make_params(clusmethod="none", corrmethod="pearson", pearson_fp64=True)
    K, L, C, _ = oracle.execute_block(p64, x, 0, n)
    assert (K == 1).all()
    r64 = C[:, 0].astype(np.float64)
    # the one-pass formula in double still cancels by (mean / sd)^2: the
    # double result lies within about n * eps * (1 + (mean / sd)^2) of exact
    eps = np.finfo(np.float64).eps
    dbl = 4 * x.shape[1] * eps * (1.0 + (mean / sd) ** 2) + 1e-13
    # and the option's float is that double rounded once
    assert (np.abs(r64 - want) <= np.abs(want) * 2.0 ** -24 + dbl).all()
    p32 = _native.make_params(clusmethod="none", corrmethod="pearson")
    K2, L2, C2, _ = oracle.execute_block(p32, x, 0, n)
    assert (K2 == K).all() and (L2 == L).all()
    err32 = np.abs(C2[:, 0].astype(np.float64) - want).max()
    if mean >= 1000.0:
        # the one-pass FP32 formula cancels catastrophically at a large mean
        assert err32 > 1e-3
    # the FP64 error reference computes the same double quantity
    ref = oracle.corr_fp64_block(p32, x, 0, n, K2, L2)[:, 0]
    assert np.abs(ref - want).max() <= dbl
    assert (r64 == ref.astype(np.float32)).all()


def test_fp64_option_with_clusters_and_nan():
    """With NaN, outliers and GMM clusters: K and labels are those of the
    FP32 mode, and each cluster's correlation is the FP64 reference's value
    rounded to float (the error reference reads the same samples)."""
    x = gemgen.modules(20, 160, 7, nan_rate=0.1)
    n = 20 * 19 // 2
    flags = dict(clusmethod="gmm", corrmethod="pearson", crit="BIC", preout=True, postout=True)
    K, L, C, _ = oracle.execute_block(_native.make_params(**flags), x, 0, n)
    K6, L6, C6, _ = oracle.execute_block(_native.make_params(pearson_fp64=True, **flags), x, 0, n)
    assert (K6 == K).all() and (L6 == L).all()
    ref = oracle.corr_fp64_block(_native.make_params(**flags), x, 0, n, K, L)
    fin = np.isfinite(ref)
    assert (np.isfinite(C6) == fin).all()
    assert (C6[fin] == ref[fin].astype(np.float32)).all()
    assert (K > 1).sum() > 10


def test_fp64_error_reference_on_ragged_pairs():
    """The FP64 reference for a pair with NaN under clusmethod none, where the
    reference's Pearson reads X[i] for every sample whose label is 0 over the
    compacted workspace: with NaN only at the tail of both genes the
    compacted and full indexes agree, so numpy over the non-NaN prefix pins
    it."""
    S = 120
    rng = np.random.default_rng(5)
    x = (5.0 + rng.normal(0, 1, (2, S))).astype(np.float32)
    x[0, 100:] = np.nan
    x[1, 110:] = np.nan
    p = _native.make_params(clusmethod="none", corrmethod="pearson")
    K, L, C, _ = oracle.execute_block(p, x, 0, 1)
    ref = oracle.corr_fp64_block(p, x, 0, 1, K, L)[0, 0]
    want = np.corrcoef(x[1, :100].astype(np.float64), x[0, :100].astype(np.float64))[0, 1]
    assert abs(ref - want) <= 1e-12
