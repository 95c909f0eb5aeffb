"""The K-means assignment of the HIP path compares squared distances and
takes square roots only near a tie (kinc_amd/csrc/sim_pair_impl.h,
nearest_mean_q).  Its claim: when q_k > fl(q_min * (1 + 2^-20)), the
correctly rounded sqrt(q_k) is strictly above sqrt(q_min), so the first
minimum of q is the first minimum of sqrt(q) (the reference's rule,
pairwise_kmeans.cpp:37-54).  Checked here in float32 arithmetic (numpy's
sqrt is correctly rounded) on adversarial neighbours of every binade,
subnormals included."""
import numpy as np


def reference_rule(q):
    d = np.sqrt(q.astype(np.float32))
    best, bd = -1, np.float32(0)
    for k, v in enumerate(d):
        if best == -1 or v < bd:
            best, bd = k, v
    return best


def fast_rule(q):
    q = q.astype(np.float32)
    kq, qmin = 0, q[0]
    for k in range(1, q.size):
        if q[k] < qmin:
            kq, qmin = k, q[k]
    thr = np.float32(qmin * np.float32(1.0 + 2.0 ** -20))
    tie = any(k < kq and q[k] <= thr for k in range(q.size - 1))
    return kq, tie


def test_margin_separates_square_roots():
    rng = np.random.default_rng(0)
    exps = np.arange(-149, 128)
    base = np.concatenate([np.ldexp(rng.uniform(1, 2, 40), e).astype(np.float32) for e in exps])
    base = base[np.isfinite(base) & (base > 0)]
    with np.errstate(over="ignore"):
        thr = (base * np.float32(1.0 + 2.0 ** -20)).astype(np.float32)
    above = np.nextafter(thr, np.float32(np.inf))          # the smallest q_k with no tie flag
    ok = np.isfinite(above)
    assert np.all(np.sqrt(above[ok]) > np.sqrt(base[ok]))


def test_fast_rule_equals_reference_on_near_ties():
    rng = np.random.default_rng(1)
    for _ in range(20000):
        K = int(rng.integers(2, 6))
        e = int(rng.integers(-149, 100))
        q0 = np.float32(np.ldexp(rng.uniform(1, 2), e))
        q = np.empty(K, np.float32)
        for k in range(K):
            step = int(rng.integers(-3, 40))
            v = q0
            for _ in range(abs(step)):
                v = np.nextafter(v, np.float32(np.inf) if step > 0 else np.float32(0))
            q[k] = v
        if rng.random() < 0.1:
            q[int(rng.integers(0, K))] = np.nan
        kq, tie = fast_rule(q)
        if not tie:
            assert kq == reference_rule(q), q


def running_rule(q):
    """The device's form of fast_rule: the near-tie flag from the running
    minimum at the moment the choice last moved (the smallest q before kq)."""
    q = q.astype(np.float32)
    kq, qmin, before = 0, q[0], np.float32(np.nan)
    for k in range(1, q.size):
        if q[k] < qmin:
            before, kq, qmin = qmin, k, q[k]
    with np.errstate(over="ignore", invalid="ignore"):
        thr = np.float32(qmin * np.float32(1.0 + 2.0 ** -20))
    return kq, bool(before <= thr)


def test_running_minimum_flags_the_same_near_ties():
    """running_rule gives fast_rule's choice and flag on every draw: near ties
    around one value, NaN and +inf entries, exact duplicates, 1-64 entries."""
    rng = np.random.default_rng(2)
    for _ in range(40000):
        K = int(rng.integers(1, 65)) if rng.random() < 0.2 else int(rng.integers(1, 9))
        e = int(rng.integers(-149, 126))
        q0 = np.float32(np.ldexp(rng.uniform(1, 2), e))
        q = np.empty(K, np.float32)
        for k in range(K):
            step = int(rng.integers(-3, 40))
            v = q0
            for _ in range(abs(step)):
                v = np.nextafter(v, np.float32(np.inf) if step > 0 else np.float32(0))
            q[k] = v
        if K > 1 and rng.random() < 0.3:
            q[int(rng.integers(0, K))] = q[int(rng.integers(0, K))]
        for special in (np.nan, np.inf):
            if rng.random() < 0.1:
                q[int(rng.integers(0, K))] = special
        with np.errstate(over="ignore", invalid="ignore"):
            assert running_rule(q) == fast_rule(q), q
