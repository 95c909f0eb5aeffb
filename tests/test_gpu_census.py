"""The phase census (kinc_selftest_census): the instrumented build of the
7-slot, 5-cluster pair kernel gives the product build's bits, and its event
counters equal what the kernel's own iteration records fix.  With 448 valid
samples per pair (7 full slots of 64 points) and BIC, a fit of K clusters
that ran `it` E steps and converged (c = 1) or not (c = 0) executes
  E-step slots 7 (it + c), pass A slots 7 it, pass B slots 7 (it - c),
  exps 7K it (E) + 14K it (pass A) + 7K c (final E), logs 7 it + 7 c,
and the fit with 1 cluster runs its own chain sums (7 (2 it - c) calls of
64 adds on 16 lanes) and M steps (it - c, 2 exps and 2 logs each), while
the fits with 2 and 3 clusters share theirs (max(it) pass-A chain calls on
17 lanes, max(it - c) pass-B calls on 16 lanes, and M steps), and so do the
fits with 4 and 5 clusters (pass A on 29 lanes, pass B on 27)."""
import numpy as np
import pytest
import torch

from kinc_amd import _native, gemgen

pytestmark = pytest.mark.gpu


def expected_counts(iters):
    it = (iters & 0x0FFF).astype(np.int64)
    c = ((iters >> 12) & 1).astype(np.int64)
    K = np.arange(1, 6)[None, :].astype(np.int64)
    single = [0]
    ms_single = (it - c)[:, single].sum()
    calls_single = (7 * (2 * it - c))[:, single].sum()
    calls = calls_single
    lanes = 16 * calls_single
    msteps = ms_single
    # the joint pairs of fits: (column a, column b, pass-A lanes, pass-B lanes)
    for ka, kb, la, lb in ((1, 2, 17, 16), (3, 4, 29, 27)):
        ja = np.maximum(it[:, ka], it[:, kb]).sum()
        jb = np.maximum(it[:, ka] - c[:, ka], it[:, kb] - c[:, kb]).sum()
        calls += 7 * (ja + jb)
        lanes += 7 * (la * ja + lb * jb)
        msteps += jb
    return {
        "fits": int(it.size),
        "iters": int(it.sum()),
        "estep_slots": int((7 * (it + c)).sum()),
        "a_slots": int((7 * it).sum()),
        "b_slots": int((7 * (it - c)).sum()),
        "chain_calls": int(calls),
        "chain_adds": int(64 * calls),
        "chain_lanes": int(64 * lanes),
        "exps": int((7 * K * it + 14 * K * it + 7 * K * c).sum() + 2 * msteps),
        "logs": int((7 * it + 7 * c).sum() + 2 * msteps),
    }


def test_census_counts_match_iteration_records_and_results_match():
    x = gemgen.modules(64, 448, seed=41, nan_rate=0.0)
    p = _native.make_params(clusmethod="gmm", corrmethod="spearman", crit="BIC",
                            preout=True, postout=True)
    eng = _native.Engine(x, p)
    ctr = torch.zeros(_native.CENSUS_COUNTERS, dtype=torch.int64, device="cuda:0")
    n = _native.pair_count(64)
    try:
        want = eng.execute(0, n)
        eng.set_census(ctr.data_ptr())
        got = eng.execute(0, n)
        eng.set_census(None)
        again = eng.execute(0, n)
    finally:
        eng.close()
    for a, b, name in zip(want, got, ("K", "labels", "corr", "iters")):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name
    for a, b in zip(want, again):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))

    c = _native.census_dict(ctr.cpu().numpy())
    print({k: v for k, v in c.items() if k != "cycles"}, flush=True)
    exp = expected_counts(want[3])
    assert c["pairs"] == n
    for k, v in exp.items():
        assert c[k] == v, (k, c[k], v)
    cyc = c["cycles"]
    named = sum(v for k, v in cyc.items() if k != "wave") + c["queue_cycles"]
    assert 0 < named <= cyc["wave"]
    assert named >= 0.8 * cyc["wave"]
    assert cyc["a_chain"] > 0 and cyc["b_chain"] > 0 and cyc["estep"] > 0


def test_census_refused_without_a_seven_slot_bin():
    x = gemgen.modules(20, 200, seed=42, nan_rate=0.0)
    eng = _native.Engine(x, _native.make_params(clusmethod="gmm", corrmethod="pearson"))
    ctr = torch.zeros(_native.CENSUS_COUNTERS, dtype=torch.int64, device="cuda:0")
    try:
        with pytest.raises(_native.KincError):
            eng.set_census(ctr.data_ptr())
        eng.set_census(None)
    finally:
        eng.close()
