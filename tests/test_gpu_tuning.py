"""The launch-configuration autotuner (kinc_sim_autotune): every
configuration a handle can take (the bins' kernel builds for 4 waves per
SIMD or for 5, 3 or 2 where the LDS makes them useful, 1, 2 or 4 auxiliary
streams for the bins, the GMM fits with 4 and 5 clusters in lockstep or one
after the other) gives the same bits, equal to the oracle's, and the tuner
keeps one of them."""
import numpy as np
import pytest

import oracle
from kinc_amd import _native, gemgen
from parity_util import assert_block_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("S,flags,configs", [
    # 256 samples: the 5-wave builds (small layout) or the lockstep fits (their
    # tile enlarges the layout past room for 20 waves per CU) on 4-wave builds
    (256, dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", preout=True, postout=True),
     [(4, False), (4, True), (5, False)]),
    (300, dict(clusmethod="kmeans", corrmethod="pearson", crit="ICL"), [(4, False), (5, False)]),
])
def test_every_configuration_gives_the_same_bits(S, flags, configs):
    x = gemgen.modules(140, S, seed=31, nan_rate=0.15)
    p = _native.make_params(**flags)
    eng = _native.Engine(x, p)
    try:
        want = eng.execute(0, 3000)
        t = eng.autotune(500, 2000)
        print("tuning", t, flush=True)
        assert t["candidates"] == 3 * len(configs)
        assert (t["pair_waves"], t["joint45"]) in configs
        assert t["aux_streams"] in (1, 2, 4) and t["sample_ms"] > 0
        assert_block_equal(eng.execute(0, 3000), want, S, "tuned")
        for waves, j45 in configs:
            for naux in (1, 2, 4):
                eng.set_tuning(waves, naux, joint45=j45)
                assert_block_equal(eng.execute(0, 3000), want, S,
                                   "waves=%d joint45=%s aux=%d" % (waves, j45, naux))
        with pytest.raises(_native.KincError):
            eng.set_tuning(4, 5, joint45=False)
        with pytest.raises(_native.KincError):
            eng.set_tuning(3, 4, joint45=False)  # no bins of 5 slots or more here
        if (5, True) not in configs and (4, True) in configs:
            with pytest.raises(_native.KincError):
                eng.set_tuning(5, 4, joint45=True)
    finally:
        eng.close()
    assert_block_equal(want, oracle.execute_block(p, x, 0, 3000), S, "oracle")


def test_five_wave_builds_refused_where_lds_leaves_no_room():
    x = gemgen.modules(20, 900, seed=32, nan_rate=0.1)
    eng = _native.Engine(x, _native.make_params(clusmethod="gmm", corrmethod="spearman"))
    try:
        t = eng.autotune(0, 150)
        # 900 samples: no 5-wave builds; 3-wave builds or 4-wave ones, the
        # lockstep fits on or off
        assert t["pair_waves"] in (3, 4) and t["candidates"] == 12
        with pytest.raises(_native.KincError):
            eng.set_tuning(5, 4)
    finally:
        eng.close()


@pytest.mark.parametrize("S,nan,flags", [
    # about 406 valid samples: the 7-slot instance
    (480, 0.08, dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", preout=True,
                     postout=True)),
    # about 380: the 6-slot instance
    (448, 0.08, dict(clusmethod="gmm", corrmethod="pearson", crit="ICL", preout=True,
                     postout=False)),
    # up to 8 clusters (fits 2+3 and 4+5 in lockstep, 6 alone), about 200
    # valid samples: the 4-slot instance of the 8-cluster kernel
    (480, 0.35, dict(clusmethod="gmm", corrmethod="spearman", crit="ICL", preout=False,
                     postout=True, minclus=2, maxclus=6)),
])
def test_lockstep_four_and_five_cluster_fits(S, nan, flags):
    """The fits with 4 and 5 clusters in lockstep (their chain tile over X)
    give the bits of the fits one after the other and of the oracle."""
    x = gemgen.modules(120, S, seed=33, nan_rate=nan)
    p = _native.make_params(**flags)
    eng = _native.Engine(x, p)
    try:
        assert eng.tuning()["joint45"] is True
        got = eng.execute(0, 3000)
        eng.set_tuning(4, 4, joint45=False)
        assert eng.tuning()["joint45"] is False
        assert_block_equal(eng.execute(0, 3000), got, S, "joint45 off")
        eng.set_tuning(4, 2, joint45=True)
        assert_block_equal(eng.execute(0, 3000), got, S, "joint45 on again")
    finally:
        eng.close()
    assert_block_equal(got, oracle.execute_block(p, x, 0, 3000), S, "oracle")


def test_lockstep_fits_refused_without_four_and_five_cluster_fits():
    x = gemgen.modules(20, 480, seed=34, nan_rate=0.1)
    eng = _native.Engine(x, _native.make_params(clusmethod="gmm", corrmethod="spearman",
                                                maxclus=4))
    try:
        assert eng.tuning()["joint45"] is False
        with pytest.raises(_native.KincError):
            eng.set_tuning(4, 4, joint45=True)
    finally:
        eng.close()


@pytest.mark.parametrize("S,flags", [
    # per-wave LDS of about 17.5 KB: 9 waves per CU
    (900, dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", preout=True, postout=True)),
    # about 18.7 KB: 8 waves per CU
    (1024, dict(clusmethod="gmm", corrmethod="spearman", crit="ICL")),
    # K-means + Pearson without outliers (no sort buffers): 10 waves per CU
    (1000, dict(clusmethod="kmeans", corrmethod="pearson", crit="BIC")),
])
def test_wide_register_builds(S, flags):
    """Where the LDS caps residency at 12 waves per CU the handle runs the
    3-wave builds (and the lockstep fits on up to 10 slots); they give the
    bits of the 4-wave builds and of the oracle."""
    x = gemgen.modules(40, S, seed=35, nan_rate=0.2, nan_beta=True)
    p = _native.make_params(**flags)
    eng = _native.Engine(x, p)
    n = 500
    try:
        assert eng.tuning()["pair_waves"] == 3
        got = eng.execute(0, n)
        for j45 in ((False, True) if flags["clusmethod"] == "gmm" else (False,)):
            for w in (3, 4):
                eng.set_tuning(w, 4, joint45=j45)
                assert_block_equal(eng.execute(0, n), got, S, "waves=%d joint45=%s" % (w, j45))
        with pytest.raises(_native.KincError):
            eng.set_tuning(5, 4)
    finally:
        eng.close()
    assert_block_equal(got, oracle.execute_block(p, x, 0, n), S, "oracle")


def test_group_autotune_sets_one_configuration_on_every_worker():
    """kinc_sim_autotune_group: the worker handles of one run are timed
    together and all get the fastest configuration; their results are those
    of an untuned handle and of the oracle.  Handles that differ in
    parameters or appear twice are refused."""
    S = 256
    flags = dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", preout=True, postout=True)
    x = gemgen.modules(140, S, seed=36, nan_rate=0.15)
    p = _native.make_params(**flags)
    engs = [_native.Engine(x, p) for _ in range(2)]
    other = _native.Engine(x, _native.make_params(**dict(flags, maxclus=4)))
    try:
        want = engs[0].execute(0, 3000)
        t = _native.autotune_group(engs, 500, 2000)
        print("group tuning", t, flush=True)
        # 256 samples: (4-wave, lockstep off/on) and (5-wave, off), 3 stream counts
        assert t["candidates"] == 9 and t["sample_ms"] > 0
        for e in engs:
            got = e.tuning()
            assert (got["pair_waves"], got["aux_streams"], got["joint45"]) == \
                (t["pair_waves"], t["aux_streams"], t["joint45"])
            assert_block_equal(e.execute(0, 3000), want, S, "tuned worker")
        for bad in ([engs[0], engs[0]], [engs[0], other]):
            with pytest.raises(_native.KincError) as err:
                _native.autotune_group(bad, 500, 2000)
            assert err.value.code == _native.KINC_E_INVALID
        # a refused group leaves the handles' configuration as it was
        assert engs[0].tuning()["aux_streams"] == t["aux_streams"]
    finally:
        for e in engs + [other]:
            e.close()
    assert_block_equal(want, oracle.execute_block(p, x, 0, 3000), S, "oracle")
