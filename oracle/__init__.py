"""CPU oracle of the pairwise Similarity path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker or the timed CPU baseline.
The product (kinc_amd/) never imports, links or calls it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkinc_oracle.so")
_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32, i64, f32 = ctypes.c_int32, ctypes.c_int64, ctypes.c_float
        sig = {
            "oracle_pair_index": (None, [i64, P, P]),
            "oracle_pair_index_loop": (None, [i64, P, P]),
            "oracle_lcg": (ctypes.c_int, [P]),
            "oracle_expf": (f32, [f32]),
            "oracle_logf": (f32, [f32]),
            "oracle_bic": (f32, [ctypes.c_int, f32, ctypes.c_int, ctypes.c_int]),
            "oracle_icl": (f32, [ctypes.c_int, f32, ctypes.c_int, ctypes.c_int, f32]),
            "oracle_set_plant": (None, [ctypes.c_int]),
            "oracle_gmm_fit_flat": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, P, P, P, P]),
            "oracle_kmeans_fit_flat": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, P, P, P]),
            "oracle_mark_outliers_flat": (None, [P, ctypes.c_int, ctypes.c_int, P, ctypes.c_int,
                                                 ctypes.c_int]),
            "oracle_pearson_flat": (f32, [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "oracle_spearman_flat": (f32, [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "oracle_bitonic_sort": (None, [ctypes.c_int, P, P]),
            "oracle_execute_block": (ctypes.c_int, [P, P, i32, i32, i64, i64, P, P, P, P]),
            "oracle_corr_fp64_block": (ctypes.c_int, [P, P, i32, i32, i64, i64, P, P, P]),
            "oracle_math_array": (None, [ctypes.c_int, P, P, i64]),
            "oracle_execute_pairs": (ctypes.c_int, [P, P, i32, i32, P, P, i64, P, P, P, P]),
            "oracle_execute_pairs_range": (ctypes.c_int, [P, P, i32, i32, P, P, i64, i64, i64,
                                                          P, P, P, P]),
            "oracle_rmt_maximums": (None, [P, ctypes.c_int, ctypes.c_int, P]),
            "oracle_rmt_prune": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int, f32, P, P]),
            "oracle_rmt_eigenvalues": (None, [P, ctypes.c_int, P]),
            "oracle_akima": (ctypes.c_int, [P, P, ctypes.c_int, P, ctypes.c_int, P]),
            "oracle_rmt_chi_square": (f32, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int]),
            "oracle_rmt_run": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, f32, f32, f32, f32,
                                              f32, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, P, P, P, ctypes.c_int, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def _ptr(a):
    return a.ctypes.data if a is not None else None


def execute_block(params, emx, start, count, want_labels=True):
    """Serial restatement of Similarity::Serial::execute for one block."""
    emx = np.ascontiguousarray(emx, np.float32)
    G, S = emx.shape
    MC = params.max_clusters
    K = np.empty(count, np.int8)
    labels = np.empty((count, S), np.int8) if want_labels else None
    corr = np.empty((count, MC), np.float32)
    iters = np.empty((count, MC), np.uint16)
    rc = lib().oracle_execute_block(ctypes.byref(params), emx.ctypes.data, G, S, start, count,
                                    K.ctypes.data, _ptr(labels), corr.ctypes.data,
                                    iters.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle_execute_block failed: %d" % rc)
    return K, labels, corr, iters


def corr_fp64_block(params, emx, start, count, K, labels):
    emx = np.ascontiguousarray(emx, np.float32)
    G, S = emx.shape
    out = np.empty((count, params.max_clusters), np.float64)
    K = np.ascontiguousarray(K, np.int8)
    labels = np.ascontiguousarray(labels, np.int8)
    lib().oracle_corr_fp64_block(ctypes.byref(params), emx.ctypes.data, G, S, start, count,
                                 K.ctypes.data, labels.ctypes.data, out.ctypes.data)
    return out


def pair_index(k, loop=False):
    x = ctypes.c_int32()
    y = ctypes.c_int32()
    fn = lib().oracle_pair_index_loop if loop else lib().oracle_pair_index
    fn(k, ctypes.byref(x), ctypes.byref(y))
    return x.value, y.value


def lcg_sequence(n, seed=1):
    st = ctypes.c_uint64(seed)
    return [lib().oracle_lcg(ctypes.byref(st)) for _ in range(n)]


def expf(x):
    return lib().oracle_expf(x)


def logf(x):
    return lib().oracle_logf(x)


def gmm_fit(X, K):
    X = np.ascontiguousarray(X, np.float32)
    N = X.shape[0]
    labels = np.zeros(N, np.int8)
    logL = ctypes.c_float()
    E = ctypes.c_float()
    st = np.zeros(4, np.int32)
    ok = lib().oracle_gmm_fit_flat(X.ctypes.data, N, K, labels.ctypes.data, ctypes.byref(logL),
                                   ctypes.byref(E), st.ctypes.data)
    return bool(ok), labels, logL.value, E.value, st


def kmeans_fit(X, K):
    X = np.ascontiguousarray(X, np.float32)
    N = X.shape[0]
    labels = np.zeros(N, np.int8)
    logL = ctypes.c_float()
    st = np.zeros(4, np.int32)
    ok = lib().oracle_kmeans_fit_flat(X.ctypes.data, N, K, labels.ctypes.data,
                                      ctypes.byref(logL), st.ctypes.data)
    return bool(ok), labels, logL.value, st


def mark_outliers(X, labels, j, cluster, marker):
    X = np.ascontiguousarray(X, np.float32)
    lab = np.ascontiguousarray(labels, np.int8).copy()
    lib().oracle_mark_outliers_flat(X.ctypes.data, X.shape[0], j, lab.ctypes.data, cluster, marker)
    return lab


def pearson(data, labels, cluster, minsamp):
    data = np.ascontiguousarray(data, np.float32)
    labels = np.ascontiguousarray(labels, np.int8)
    return lib().oracle_pearson_flat(data.ctypes.data, labels.ctypes.data, labels.shape[0],
                                     cluster, minsamp)


def spearman(data, labels, cluster, minsamp):
    data = np.ascontiguousarray(data, np.float32)
    labels = np.ascontiguousarray(labels, np.int8)
    return lib().oracle_spearman_flat(data.ctypes.data, labels.ctypes.data, labels.shape[0],
                                      cluster, minsamp)


def bitonic_sort(key, payload):
    key = np.ascontiguousarray(key, np.float32).copy()
    payload = np.ascontiguousarray(payload, np.float32).copy()
    lib().oracle_bitonic_sort(key.shape[0], key.ctypes.data, payload.ctypes.data)
    return key, payload


def bic(K, logL, N, D=2):
    return lib().oracle_bic(K, logL, N, D)


def icl(K, logL, N, E, D=2):
    return lib().oracle_icl(K, logL, N, D, E)


def set_plant(which):
    """Switch a deliberate slip into the oracle's GMM entropy (0: none; 1:
    taken on gamma instead of log gamma after a converged fit; 2: log pi
    added to the log responsibilities).  Test hook only."""
    lib().oracle_set_plant(which)


def math_array(fn, x):
    """Host evaluation of exp (0), log (1), sqrt (2) or 1/x (3)."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().oracle_math_array(fn, x.ctypes.data, out.ctypes.data, x.size)
    return out


def execute_pairs(params, emx, pairs, seg_lo=None, want_labels=True):
    """Serial restatement over an explicit pair list; seg_lo marks work blocks."""
    emx = np.ascontiguousarray(emx, np.float32)
    G, S = emx.shape
    pairs = np.ascontiguousarray(pairs, np.int32)
    count = pairs.shape[0]
    MC = params.max_clusters
    K = np.empty(count, np.int8)
    labels = np.empty((count, S), np.int8) if want_labels else None
    corr = np.empty((count, MC), np.float32)
    iters = np.empty((count, MC), np.uint16)
    seg = None if seg_lo is None else np.ascontiguousarray(seg_lo, np.int32)
    lib().oracle_execute_pairs(ctypes.byref(params), emx.ctypes.data, G, S, pairs.ctypes.data,
                               _ptr(seg), count, K.ctypes.data, _ptr(labels), corr.ctypes.data,
                               iters.ctypes.data)
    return K, labels, corr, iters


# ------------------------------------------------------------------ RMT

def rmt_maximums(matrix):
    """RMT::computeMaximums on a dense [N][N][K] float32 matrix."""
    m = np.ascontiguousarray(matrix, np.float32)
    N, _, K = m.shape
    out = np.empty(N * K, np.float32)
    lib().oracle_rmt_maximums(m.ctypes.data, N, K, out.ctypes.data)
    return out


def rmt_prune(matrix, maximums, threshold):
    m = np.ascontiguousarray(matrix, np.float32)
    N, _, K = m.shape
    mx = np.ascontiguousarray(maximums, np.float32)
    n = int((mx >= np.float32(threshold)).sum())
    idx = np.empty(max(n, 1), np.int32)
    pr = np.empty((max(n, 1), max(n, 1)), np.float32)
    got = lib().oracle_rmt_prune(m.ctypes.data, mx.ctypes.data, N, K, threshold, idx.ctypes.data,
                                 pr.ctypes.data)
    assert got == n
    return idx[:n], pr[:n, :n]


def rmt_eigenvalues(prune):
    p = np.ascontiguousarray(prune, np.float32)
    n = p.shape[0]
    out = np.empty(n, np.float32)
    lib().oracle_rmt_eigenvalues(p.ctypes.data, n, out.ctypes.data)
    return out


def akima(x, y, xe):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    xe = np.ascontiguousarray(xe, np.float64)
    ye = np.empty_like(xe)
    rc = lib().oracle_akima(x.ctypes.data, y.ctypes.data, x.size, xe.ctypes.data, xe.size,
                            ye.ctypes.data)
    if rc:
        raise ValueError("akima init failed: %d" % rc)
    return ye


def rmt_chi_square(eigens, min_eigen=50, min_pace=10, max_pace=40, bins=60):
    e = np.ascontiguousarray(eigens, np.float32)
    return lib().oracle_rmt_chi_square(e.ctypes.data, e.size, min_eigen, min_pace, max_pace, bins)


def rmt_run(matrix, start=0.99, step=0.001, stop=0.5, chi1=99.607, chi2=200.0, min_eigen=50,
            min_pace=10, max_pace=40, bins=60, cap=1000):
    """RMT::process; returns (rc_or_steps, final_threshold, log rows (t, size, chi))."""
    m = np.ascontiguousarray(matrix, np.float32)
    N, _, K = m.shape
    lt = np.empty(cap, np.float32)
    ls = np.empty(cap, np.int32)
    lc = np.empty(cap, np.float32)
    ft = ctypes.c_float()
    steps = ctypes.c_int()
    rc = lib().oracle_rmt_run(m.ctypes.data, N, K, start, step, stop, chi1, chi2, min_eigen,
                              min_pace, max_pace, bins, lt.ctypes.data, ls.ctypes.data,
                              lc.ctypes.data, cap, ctypes.byref(ft), ctypes.byref(steps))
    n = min(steps.value, cap)
    return rc, ft.value, list(zip(lt[:n].tolist(), ls[:n].tolist(), lc[:n].tolist()))


def block_pairs(start, count):
    """(count, 2) int32 pair list (x, y) of the linear pairs [start, start + count)
    in index order (Pairwise::Index, src/core/pairwise_index.cpp:30-94)."""
    k = np.arange(start, start + count, dtype=np.int64)
    x = ((1 + np.sqrt(1 + 8 * k.astype(np.float64))) // 2).astype(np.int64)
    x -= (x * (x - 1) // 2 > k)
    x += ((x + 1) * x // 2 <= k)
    y = k - x * (x - 1) // 2
    return np.ascontiguousarray(np.stack([x, y], 1).astype(np.int32))


def execute_pairs_threads(params, emx, pairs, seg_lo=None, threads=8, want_labels=True):
    """execute_pairs split into contiguous ranges on host threads (each
    range replays the fetches of the earlier pairs of its work block, so the
    results equal one serial call's)."""
    from concurrent.futures import ThreadPoolExecutor
    emx = np.ascontiguousarray(emx, np.float32)
    G, S = emx.shape
    pairs = np.ascontiguousarray(pairs, np.int32)
    count = pairs.shape[0]
    MC = params.max_clusters
    K = np.empty(count, np.int8)
    labels = np.empty((count, S), np.int8) if want_labels else None
    corr = np.empty((count, MC), np.float32)
    iters = np.empty((count, MC), np.uint16)
    seg = None if seg_lo is None else np.ascontiguousarray(seg_lo, np.int32)
    # small ranges, handed out in order, balance pairs of uneven cost
    step = max(1, min(256, count // max(1, 4 * threads)))
    ranges = [(a, min(step, count - a)) for a in range(0, count, step)]

    def run(r):
        a, n = r
        rc = lib().oracle_execute_pairs_range(
            ctypes.byref(params), emx.ctypes.data, G, S, pairs.ctypes.data, _ptr(seg), count,
            a, n, K[a:].ctypes.data, labels[a:].ctypes.data if labels is not None else None,
            corr[a:].ctypes.data, iters[a:].ctypes.data)
        if rc != 0:
            raise RuntimeError("oracle_execute_pairs_range failed: %d" % rc)

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(run, ranges))
    return K, labels, corr, iters


def execute_block_threads(params, emx, start, count, threads=8, want_labels=True):
    """execute_block on host threads (see execute_pairs_threads)."""
    return execute_pairs_threads(params, emx, block_pairs(start, count), None, threads,
                                 want_labels)
