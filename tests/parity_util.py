"""Helpers for comparing the HIP path with the oracle."""
import numpy as np


def same_float_bits(a, b):
    """Elementwise bit equality of float32 arrays (NaN == NaN of any payload)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    nan = np.isnan(a) & np.isnan(b)
    return nan | (a.view(np.uint32) == b.view(np.uint32))


def assert_block_equal(got, want, S, where=""):
    """got/want: (K, labels, corr, iters) tuples; bit-exact comparison."""
    K1, L1, C1, I1 = got
    K0, L0, C0, I0 = want
    bad = np.nonzero(K1 != K0)[0]
    assert bad.size == 0, "%s K differs at pairs %s: got %s want %s" % (
        where, bad[:10], K1[bad[:10]], K0[bad[:10]])
    if L0 is not None and L1 is not None:
        bad = np.nonzero((L1 != L0).any(axis=1))[0]
        assert bad.size == 0, "%s labels differ at pairs %s" % (where, bad[:10])
    if I0 is not None and I1 is not None:
        bad = np.nonzero((I1 != I0).any(axis=1))[0]
        assert bad.size == 0, "%s iteration records differ at pairs %s: got %s want %s" % (
            where, bad[:5], I1[bad[:5]], I0[bad[:5]])
    eq = same_float_bits(C1, C0)
    bad = np.nonzero(~eq.all(axis=1))[0]
    assert bad.size == 0, "%s correlations differ at pairs %s: got %s want %s" % (
        where, bad[:5], C1[bad[:5]], C0[bad[:5]])


def count_mismatching_pairs(got, want):
    """Number of pairs whose K, labels, iteration record or correlation bits differ."""
    K1, L1, C1, I1 = got
    K0, L0, C0, I0 = want
    bad = K1 != K0
    bad |= (L1 != L0).any(axis=1)
    bad |= (I1 != I0).any(axis=1)
    bad |= ~same_float_bits(C1, C0).all(axis=1)
    return int(np.count_nonzero(bad))
