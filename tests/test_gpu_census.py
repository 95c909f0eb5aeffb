"""The phase census (kinc_selftest_census): the instrumented build of the
7-slot, 5-cluster pair kernel gives the product build's bits, and its event
counters equal what the kernel's own iteration records fix: with 448 valid
samples per pair (7 full slots of 64 points) and BIC, a fit of K clusters
that ran `it` E steps and converged (c = 1) or not (c = 0) executes
  E-step slots 7 (it + c), pass A slots 7 it, pass B slots 7 (it - c),
  chain-sum adds 448 (2 it - c) on 16 active lanes each,
  exps 7K it (E) + 14K it (pass A) + 2 (it - c) (M step) + 7K c (final E),
  logs 7 it + 2 (it - c) + 7 c."""
import numpy as np
import pytest
import torch

from kinc_amd import _native, gemgen

pytestmark = pytest.mark.gpu


def expected_counts(iters, K_of_lane):
    it = (iters & 0x0FFF).astype(np.int64)
    c = ((iters >> 12) & 1).astype(np.int64)
    K = np.broadcast_to(K_of_lane, it.shape).astype(np.int64)
    return {
        "fits": int(it.size),
        "iters": int(it.sum()),
        "estep_slots": int((7 * (it + c)).sum()),
        "a_slots": int((7 * it).sum()),
        "b_slots": int((7 * (it - c)).sum()),
        "chain_adds": int((448 * (2 * it - c)).sum()),
        "chain_calls": int((7 * (2 * it - c)).sum()),
        "exps": int((7 * K * it + 14 * K * it + 2 * (it - c) + 7 * K * c).sum()),
        "logs": int((7 * it + 2 * (it - c) + 7 * c).sum()),
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
    iters = want[3]
    exp = expected_counts(iters, np.arange(1, 6)[None, :])
    assert c["pairs"] == n
    for k, v in exp.items():
        assert c[k] == v, (k, c[k], v)
    # every chain sum runs on 16 lanes (1 + 3K <= 16 chains for K <= 5)
    assert c["chain_lanes"] == 16 * c["chain_adds"]
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
