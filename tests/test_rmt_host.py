"""The host half of the RMT analytic (kinc_rmt_chi_square: unique values,
Akima unfolding, NNSD histogram, chi-square; RMT::computeChiSquare,
src/core/rmt.cpp:282-439) against the oracle, bit for bit, on the spectra
the oracle's own pins use.  Runs without a GPU: no device call is made."""
import numpy as np
import pytest

import oracle
from kinc_amd import rmt
from test_rmt_oracle import goe_spectrum, poisson_spectrum


def spectra():
    yield "poisson", poisson_spectrum(1200, 0)
    yield "goe", goe_spectrum(900, 1)
    yield "small", poisson_spectrum(51, 2)
    yield "few-unique", np.repeat(poisson_spectrum(30, 3), 3)
    e = poisson_spectrum(400, 3)
    e[1::2] *= -1
    yield "interleaved", e[np.argsort(np.abs(e), kind="stable")]
    rng = np.random.default_rng(4)
    yield "mixed", np.sort(np.concatenate([rng.uniform(0.01, 0.5, 300),
                                           rng.uniform(1.5, 6.0, 300)])).astype(np.float32)


@pytest.mark.parametrize("name,e", list(spectra()), ids=[n for n, _ in spectra()])
@pytest.mark.parametrize("kw", [{}, dict(minpace=5, maxpace=12, bins=30)])
def test_chi_square_matches_oracle(native, name, e, kw):
    got = np.float32(rmt.chi_square(e, **kw))
    p = rmt.make_params(**kw)
    want = np.float32(oracle.rmt_chi_square(e, p.min_eigenvalues, p.min_pace, p.max_pace, p.bins))
    assert got.view(np.uint32) == want.view(np.uint32), (got, want)
