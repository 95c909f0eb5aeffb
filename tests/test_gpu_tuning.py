"""The launch-configuration autotuner (kinc_sim_autotune): every
configuration a handle can take (5-wave builds of the small bins on or off,
1, 2 or 4 auxiliary streams for the bins) gives the same bits, equal to the
oracle's, and the tuner keeps one of them."""
import numpy as np
import pytest

import oracle
from kinc_amd import _native, gemgen
from parity_util import assert_block_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("S,flags", [
    (256, dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", preout=True, postout=True)),
    (300, dict(clusmethod="kmeans", corrmethod="pearson", crit="ICL")),
])
def test_every_configuration_gives_the_same_bits(S, flags):
    x = gemgen.modules(140, S, seed=31, nan_rate=0.15)
    p = _native.make_params(**flags)
    eng = _native.Engine(x, p)
    try:
        want = eng.execute(0, 3000)
        t = eng.autotune(500, 2000)
        print("tuning", t, flush=True)
        assert t["candidates"] == 6          # 256 and 300 samples leave room for 5-wave builds
        assert t["aux_streams"] in (1, 2, 4) and t["sample_ms"] > 0
        assert_block_equal(eng.execute(0, 3000), want, S, "tuned")
        for w5 in (False, True):
            for naux in (1, 2, 4):
                eng.set_tuning(w5, naux)
                assert_block_equal(eng.execute(0, 3000), want, S, "w5=%s aux=%d" % (w5, naux))
        with pytest.raises(_native.KincError):
            eng.set_tuning(False, 5)
    finally:
        eng.close()
    assert_block_equal(want, oracle.execute_block(p, x, 0, 3000), S, "oracle")


def test_five_wave_builds_refused_where_lds_leaves_no_room():
    x = gemgen.modules(20, 900, seed=32, nan_rate=0.1)
    eng = _native.Engine(x, _native.make_params(clusmethod="gmm", corrmethod="spearman"))
    try:
        t = eng.autotune(0, 150)
        assert t["w5"] is False and t["candidates"] == 3
        with pytest.raises(_native.KincError):
            eng.set_tuning(True, 4)
    finally:
        eng.close()
