"""The launch plan a handle gets (kinc_sim_plan_get, host only): bins, the
per-wave LDS layouts with and without room for the lockstep 4- and 5-cluster
chain tile, which kernel builds a handle may run and the configuration it
starts with.  These are the rules the GPU tests exercise through handles;
here they are checked on the CPU for the benchmark shapes and the edges."""
import pytest

from kinc_amd import _native, gemgen

LDS_PER_CU = 160 * 1024
ROW = 4 * 68                      # one chain-tile row (68 floats)


def a16(n):
    return (n + 15) // 16 * 16


def head(S):
    """labels, work labels and nine parameter sets before X"""
    return a16(S) + a16(S) + 288


def cfg(name, **over):
    return _native.make_params(**dict(gemgen.CONFIGS[name]["params"], **over))


def test_headline_shape():
    pl = _native.plan(cfg("H"), 20000, 500)
    assert pl["slots"] == [1, 2, 3, 4, 5, 6, 7, 8]
    # X (8 bytes per sample) and the sort buffers (2 x 512 floats) follow the head
    assert pl["lds_small"] == head(500) + 4000 + 2 * 512 * 4
    # the 29-row lockstep tile fits over X and the sort region without growing it
    assert pl["lds_joint45"] == pl["lds_small"]
    assert 29 * ROW <= 4000 + 2 * 512 * 4
    assert pl["lds_small"] * 16 <= LDS_PER_CU < pl["lds_small"] * 20
    assert pl["pair_waves_allowed"] == [4]
    t = pl["tuning"]
    assert (t["pair_waves"], t["joint45"], t["aux_streams"]) == (4, True, 4)


def test_sort_buffers_only_for_outliers_or_spearman():
    sp = _native.plan(cfg("H"), 1000, 500)
    pe = _native.plan(cfg("H", corrmethod="pearson", preout=False, postout=False), 1000, 500)
    po = _native.plan(cfg("H", corrmethod="pearson", preout=True, postout=False), 1000, 500)
    assert pe["lds_small"] < sp["lds_small"] == po["lds_small"]
    # without sorts the shared region holds the lockstep fits' second labels
    # past the entropy row
    assert pe["lds_small"] == head(500) + 4000 + a16(ROW + 500)


def test_lockstep_tile_enlarges_a_small_layout():
    pl = _native.plan(cfg("D"), 60000, 475)
    x = a16(8 * 475)
    assert pl["lds_joint45"] == head(475) + x + (29 * ROW - x)
    assert pl["lds_small"] < pl["lds_joint45"]
    # the small layout would allow the 5-wave builds, the lockstep one not:
    # the handle starts with the lockstep fits on the 4-wave builds
    assert pl["lds_small"] * 20 <= LDS_PER_CU < pl["lds_joint45"] * 20
    assert pl["pair_waves_allowed"] == [4, 5]
    assert pl["tuning"]["pair_waves"] == 4 and pl["tuning"]["joint45"] is True


def test_three_wave_builds_where_lds_caps_residency():
    km = _native.plan(cfg("C"), 20000, 1000)
    assert km["slots"] == [1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16]
    assert LDS_PER_CU // km["lds_small"] < 16
    assert km["pair_waves_allowed"] == [3, 4] and km["tuning"]["pair_waves"] == 3
    assert km["lds_joint45"] == 0 and km["tuning"]["joint45"] is False
    gm = _native.plan(cfg("C", clusmethod="gmm"), 20000, 1000)
    # every GMM tile lies over X (8,000 bytes >= 29 rows), the shared region
    # keeps only the second labels
    assert gm["lds_small"] == gm["lds_joint45"] == head(1000) + 8000 + a16(ROW + 1000)
    assert gm["tuning"]["pair_waves"] == 3 and gm["tuning"]["joint45"] is True
    # without bins of 5 slots or more there are no 3-wave builds to run
    small = _native.plan(cfg("H"), 1000, 256)
    assert 3 not in small["pair_waves_allowed"]


def test_memory_store_above_the_register_range():
    pl = _native.plan(cfg("E"), 2000, 5000)
    assert pl["slots"] == [0]
    assert pl["lds_joint45"] == 0 and pl["pair_waves_allowed"] == [4]
    assert pl["tuning"]["pair_waves"] == 4 and pl["tuning"]["joint45"] is False
    assert _native.plan(cfg("H"), 100, 1024)["slots"][-1] == 16
    assert _native.plan(cfg("H"), 100, 1025)["slots"] == [0]


def test_cluster_bounds_select_instances_and_lockstep_fits():
    eight = _native.plan(cfg("H", maxclus=8), 100, 500)
    assert eight["slots"] == [2, 4, 8]
    assert eight["lds_joint45"] != 0
    four = _native.plan(cfg("H", maxclus=4), 100, 500)
    assert four["lds_joint45"] == 0 and four["tuning"]["joint45"] is False
    # the fits with 4 and 5 clusters run together only if both are fitted
    assert _native.plan(cfg("H", minclus=5), 100, 500)["lds_joint45"] == 0
    many = _native.plan(cfg("H", maxclus=20), 100, 500)
    assert many["slots"] == [0] and many["lds_joint45"] == 0


@pytest.mark.parametrize("over,G,S,code", [
    (dict(clusmethod="gmm", maxclus=65), 100, 500, _native.KINC_E_INVALID),
    (dict(minclus=3, maxclus=2), 100, 500, _native.KINC_E_INVALID),
    ({}, 1, 500, _native.KINC_E_INVALID),
    ({}, 100, 32769, _native.KINC_E_UNSUPPORTED),
])
def test_plan_rejects_what_create_rejects(over, G, S, code):
    with pytest.raises(_native.KincError) as e:
        _native.plan(cfg("H", **over), G, S)
    assert e.value.code == code
