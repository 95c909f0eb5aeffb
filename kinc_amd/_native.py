"""ctypes binding of the C-ABI declared in include/kinc_sim.h.

The library is kinc_amd/libkinc.so, built in-tree by kinc_amd/build.py.
There is no fallback: if the library is missing or fails to load, every
entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkinc.so")

KINC_OK = 0
KINC_E_INVALID = -1
KINC_E_DEVICE = -2
KINC_E_NOMEM = -3
KINC_E_UNSUPPORTED = -4

CLUS = {"none": 0, "gmm": 1, "kmeans": 2}
CORR = {"pearson": 0, "spearman": 1}
CRIT = {"BIC": 0, "ICL": 1}
FORM = {"plain": 0, "persistent": 1}

WORK_BLOCK_SIZE = 32768
MAX_CLUSTER_SIZE = 64


class SimParams(ctypes.Structure):
    """Mirror of kinc_sim_params (include/kinc_sim.h)."""
    _fields_ = [
        ("clus_method", ctypes.c_int32),
        ("corr_method", ctypes.c_int32),
        ("min_expr", ctypes.c_float),
        ("min_samples", ctypes.c_int32),
        ("min_clusters", ctypes.c_int32),
        ("max_clusters", ctypes.c_int32),
        ("criterion", ctypes.c_int32),
        ("pre_out", ctypes.c_int32),
        ("post_out", ctypes.c_int32),
        ("min_corr", ctypes.c_float),
        ("max_corr", ctypes.c_float),
        ("ksize", ctypes.c_int32),
        ("form", ctypes.c_int32),
        ("batch_capacity", ctypes.c_int32),
        ("pearson_fp64", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 3),
    ]


_lib = None

# every entry point declared in include/kinc_sim.h: (name, restype, argtypes)
_P = ctypes.c_void_p
SIGNATURES = [
    ("kinc_sim_default_params", None, [ctypes.POINTER(SimParams)]),
    ("kinc_sim_create", ctypes.c_int, [ctypes.POINTER(SimParams), _P, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(_P)]),
    ("kinc_sim_execute", ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int64, _P, _P, _P, _P]),
    ("kinc_sim_execute_device", ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int64, _P, _P,
                                               _P, _P, _P]),
    ("kinc_sim_execute_pairs", ctypes.c_int, [_P, _P, _P, ctypes.c_int64, _P, _P, _P, _P]),
    ("kinc_sim_execute_pairs_device", ctypes.c_int, [_P, _P, _P, ctypes.c_int64, _P, _P, _P,
                                                     _P, _P]),
    ("kinc_sim_synchronize", ctypes.c_int, [_P]),
    ("kinc_sim_destroy", None, [_P]),
    ("kinc_sim_last_error", ctypes.c_char_p, [_P]),
    ("kinc_sim_set_profiling", ctypes.c_int, [_P, ctypes.c_int32]),
    ("kinc_sim_profile_read", ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int32]),
    ("kinc_selftest_math", ctypes.c_int, [ctypes.c_int32, _P, _P, ctypes.c_int64]),
    ("kinc_selftest_valu_probe", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]),
    ("kinc_selftest_census", ctypes.c_int, [_P, _P]),
    ("kinc_sim_autotune", ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int64, _P]),
    ("kinc_sim_autotune_group", ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, _P]),
    ("kinc_sim_get_tuning", ctypes.c_int, [_P, _P]),
    ("kinc_sim_set_tuning", ctypes.c_int, [_P, _P]),
    ("kinc_sim_plan_get", ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, _P]),
    ("kinc_sim_graph_begin", ctypes.c_int, [_P]),
    ("kinc_sim_graph_end", ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    ("kinc_sim_graph_launch", ctypes.c_int, [_P, _P]),
    ("kinc_sim_graph_destroy", None, [_P]),
    ("kinc_pair_count", ctypes.c_int64, [ctypes.c_int32]),
    ("kinc_pair_index", None, [ctypes.c_int64, ctypes.POINTER(ctypes.c_int32),
                               ctypes.POINTER(ctypes.c_int32)]),
]


def lib():
    """Load libkinc.so once; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("kinc_amd/libkinc.so is not built; run __graft_entry__.build()")
        l = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def default_params():
    p = SimParams()
    lib().kinc_sim_default_params(ctypes.byref(p))
    return p


def make_params(clusmethod="none", corrmethod="pearson", minexpr=float("-inf"), minsamp=30,
                minclus=1, maxclus=5, crit="ICL", preout=False, postout=False, mincorr=0.5,
                maxcorr=1.0, ksize=4096, form="persistent", batch_capacity=0, pearson_fp64=False):
    """Build kinc_sim_params from the reference's CLI flag names and defaults
    (src/core/similarity_input.cpp:96-254)."""
    p = default_params()
    p.clus_method = CLUS[clusmethod]
    p.corr_method = CORR[corrmethod]
    p.min_expr = minexpr
    p.min_samples = minsamp
    p.min_clusters = minclus
    p.max_clusters = maxclus
    p.criterion = CRIT[crit]
    p.pre_out = int(bool(preout))
    p.post_out = int(bool(postout))
    p.min_corr = mincorr
    p.max_corr = maxcorr
    p.ksize = ksize
    p.form = FORM[form]
    p.batch_capacity = batch_capacity
    p.pearson_fp64 = int(bool(pearson_fp64))
    return p


class Tuning(ctypes.Structure):
    """Mirror of kinc_sim_tuning (include/kinc_sim.h)."""
    _fields_ = [("pair_waves", ctypes.c_int32), ("aux_streams", ctypes.c_int32),
                ("joint45", ctypes.c_int32), ("candidates", ctypes.c_int32),
                ("sample_ms", ctypes.c_double)]

    def as_dict(self):
        return {"pair_waves": int(self.pair_waves), "aux_streams": int(self.aux_streams),
                "joint45": bool(self.joint45), "candidates": int(self.candidates),
                "sample_ms": float(self.sample_ms)}


class KincError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("kinc error %d: %s" % (code, msg))
        self.code = code


class Plan(ctypes.Structure):
    """Mirror of kinc_sim_plan (include/kinc_sim.h)."""
    _fields_ = [("bins", ctypes.c_int32), ("slots", ctypes.c_int32 * 12),
                ("lds_small", ctypes.c_int32), ("lds_joint45", ctypes.c_int32),
                ("pair_waves_allowed", ctypes.c_int32), ("tuning", Tuning)]


def plan(params, G, S):
    """The launch plan a handle would get (kinc_sim_plan_get; no device):
    bins' slot counts, per-wave LDS without / with the lockstep 4+5-cluster
    tile, the accepted pair_waves values and the starting configuration."""
    pl = Plan()
    rc = lib().kinc_sim_plan_get(ctypes.byref(params), G, S, ctypes.byref(pl))
    if rc != 0:
        raise KincError(rc, lib().kinc_sim_last_error(None).decode())
    return {"slots": [int(pl.slots[i]) for i in range(pl.bins)],
            "lds_small": int(pl.lds_small), "lds_joint45": int(pl.lds_joint45),
            "pair_waves_allowed": [w for w in range(32) if pl.pair_waves_allowed >> w & 1],
            "tuning": pl.tuning.as_dict()}


def pair_count(G):
    return lib().kinc_pair_count(G)


def pair_index(k):
    x = ctypes.c_int32()
    y = ctypes.c_int32()
    lib().kinc_pair_index(k, ctypes.byref(x), ctypes.byref(y))
    return x.value, y.value


class Engine:
    """One device handle with the expression matrix resident (kinc_sim_create)."""

    def __init__(self, emx, params, device=0):
        l = lib()
        emx = np.ascontiguousarray(emx, dtype=np.float32)
        self.G, self.S = emx.shape
        self.params = params
        h = _P()
        rc = l.kinc_sim_create(ctypes.byref(params), emx.ctypes.data, self.G, self.S, device,
                               ctypes.byref(h))
        if rc != KINC_OK:
            raise KincError(rc, l.kinc_sim_last_error(None).decode())
        self.h = h

    def _check(self, rc):
        if rc != KINC_OK:
            raise KincError(rc, lib().kinc_sim_last_error(self.h).decode())

    def execute(self, start, count, want_labels=True, want_iters=True):
        """Run one work block; returns (K, labels, corr, iters) numpy arrays."""
        MC = self.params.max_clusters
        K = np.empty(count, np.int8)
        corr = np.empty((count, MC), np.float32)
        labels = np.empty((count, self.S), np.int8) if want_labels else None
        iters = np.empty((count, MC), np.uint16) if want_iters else None
        rc = lib().kinc_sim_execute(self.h, start, count, K.ctypes.data,
                                    labels.ctypes.data if labels is not None else None,
                                    corr.ctypes.data,
                                    iters.ctypes.data if iters is not None else None)
        self._check(rc)
        return K, labels, corr, iters

    def execute_pairs(self, pairs, seg_lo=None, want_labels=True, want_iters=True):
        """Run an explicit pair list ((count, 2) int32 of x > y); seg_lo gives
        each pair's work-block start within the list (default: one block)."""
        pairs = np.ascontiguousarray(pairs, np.int32)
        count = pairs.shape[0]
        MC = self.params.max_clusters
        K = np.empty(count, np.int8)
        corr = np.empty((count, MC), np.float32)
        labels = np.empty((count, self.S), np.int8) if want_labels else None
        iters = np.empty((count, MC), np.uint16) if want_iters else None
        seg = None if seg_lo is None else np.ascontiguousarray(seg_lo, np.int32)
        rc = lib().kinc_sim_execute_pairs(self.h, pairs.ctypes.data,
                                          seg.ctypes.data if seg is not None else None, count,
                                          K.ctypes.data,
                                          labels.ctypes.data if labels is not None else None,
                                          corr.ctypes.data,
                                          iters.ctypes.data if iters is not None else None)
        self._check(rc)
        return K, labels, corr, iters

    def execute_device(self, start, count, K_ptr, labels_ptr, corr_ptr, iters_ptr, stream=None):
        rc = lib().kinc_sim_execute_device(self.h, start, count, K_ptr, labels_ptr, corr_ptr,
                                           iters_ptr, stream)
        self._check(rc)

    def set_profiling(self, on):
        self._check(lib().kinc_sim_set_profiling(self.h, 1 if on else 0))

    def profile_read(self):
        """{P: (launches, total_ms)} of the pair kernels since the last read."""
        cap = 32
        slots = np.zeros(cap, np.int32)
        launches = np.zeros(cap, np.int64)
        ms = np.zeros(cap, np.float64)
        n = lib().kinc_sim_profile_read(self.h, slots.ctypes.data, launches.ctypes.data,
                                        ms.ctypes.data, cap)
        if n < 0:
            self._check(n)
        return {int(slots[i]): (int(launches[i]), float(ms[i])) for i in range(n)}

    def synchronize(self):
        self._check(lib().kinc_sim_synchronize(self.h))

    def graph_begin(self):
        """Record the following execute_device calls (stream None) into a HIP
        graph instead of running them (kinc_sim_graph_begin)."""
        self._check(lib().kinc_sim_graph_begin(self.h))

    def graph_end(self):
        """Finish recording; returns a Graph that replays the recorded work."""
        g = _P()
        self._check(lib().kinc_sim_graph_end(self.h, ctypes.byref(g)))
        return Graph(g)

    def autotune(self, start, count):
        """Time the handle's launch configurations on pairs [start, start+count)
        and keep the fastest (kinc_sim_autotune); returns it as a dict."""
        t = Tuning()
        self._check(lib().kinc_sim_autotune(self.h, start, count, ctypes.byref(t)))
        return t.as_dict()

    def tuning(self):
        t = Tuning()
        self._check(lib().kinc_sim_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, pair_waves, aux_streams, joint45=None):
        """Set the launch configuration (kinc_sim_set_tuning); joint45 None
        keeps the handle's current choice."""
        t = Tuning()
        t.pair_waves = pair_waves
        t.aux_streams = aux_streams
        t.joint45 = (1 if self.tuning()["joint45"] else 0) if joint45 is None else (1 if joint45 else 0)
        self._check(lib().kinc_sim_set_tuning(self.h, ctypes.byref(t)))

    def set_census(self, counters_ptr):
        """Run the 7-slot, 5-cluster pairs on the phase-census build, adding to
        CENSUS_COUNTERS device uint64 counters at counters_ptr (zeroed by the
        caller), or on the product build again with None
        (kinc_selftest_census)."""
        self._check(lib().kinc_selftest_census(self.h, counters_ptr))

    def close(self):
        if getattr(self, "h", None):
            lib().kinc_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_group(engines, start, count):
    """Time the launch configurations with every engine of the group running
    the sample [start, start+count) at once, and set the fastest on all of them
    (kinc_sim_autotune_group; the engines must share device, matrix and
    parameters); returns it as a dict."""
    engines = list(engines)
    arr = (_P * len(engines))(*[e.h for e in engines])
    t = Tuning()
    rc = lib().kinc_sim_autotune_group(ctypes.cast(arr, _P), len(engines), start, count,
                                       ctypes.byref(t))
    if rc != KINC_OK:
        h = engines[0].h if engines else None
        raise KincError(rc, lib().kinc_sim_last_error(h).decode())
    return t.as_dict()


class Graph:
    """A recorded sequence of work blocks (kinc_sim_graph)."""

    def __init__(self, g):
        self.g = g

    def launch(self, stream=None):
        rc = lib().kinc_sim_graph_launch(self.g, stream)
        if rc != KINC_OK:
            raise KincError(rc, "kinc_sim_graph_launch failed")

    def close(self):
        if getattr(self, "g", None):
            lib().kinc_sim_graph_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_math(fn, x):
    """Evaluate exp (0), log (1), sqrt (2), 1/x (3), the log for finite
    x >= 1 or NaN (4) or the exp for x <= 0 or NaN (5) on the device."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    rc = lib().kinc_selftest_math(fn, x.ctypes.data, out.ctypes.data, x.size)
    if rc != KINC_OK:
        raise KincError(rc, "kinc_selftest_math failed")
    return out


# counters of the phase census (include/kinc_selftest.h): cycle phases 0..15,
# event counts 16..27
CENSUS_COUNTERS = 32
CENSUS_PHASES = ["fetch", "fit_init", "estep", "gmax", "a_terms", "a_chain", "mstep",
                 "b_terms", "b_chain", "cov", "labels", "entropy", "select", "corr",
                 "other_fit", "wave"]
CENSUS_EVENTS = ["pairs", "fits", "iters", "estep_slots", "a_slots", "b_slots", "chain_adds",
                 "chain_lanes", "chain_calls", "exps", "logs", "queue_cycles"]


def census_dict(values):
    """Name the KINC_CENSUS_COUNTERS values of a census run."""
    v = [int(x) for x in values]
    d = {"cycles": {n: v[i] for i, n in enumerate(CENSUS_PHASES)}}
    d.update({n: v[16 + i] for i, n in enumerate(CENSUS_EVENTS)})
    return d


# (independent chains, active lanes, instruction) of each kinc_selftest_valu_probe mode
VALU_PROBE_OPS = ["v_add_f32", "v_mul_f32", "v_fma_f32", "v_fmac_f32", "v_max_f32",
                  "v_rndne_f32", "v_ldexp_f32", "v_cvt_i32_f32", "v_cndmask_b32", "v_fmaak_f32",
                  "exp_le0", "exp"]
VALU_PROBE_MODES = [(8, 64, 0), (8, 16, 0), (8, 4, 0), (1, 64, 0), (1, 16, 0), (8, 32, 0),
                    (8, 64, 1), (8, 64, 2), (8, 64, 3), (8, 64, 4), (8, 64, 5), (8, 64, 6),
                    (8, 64, 7), (8, 64, 8), (8, 64, 9), (1, 64, 2), (8, 16, 2), (8, 64, 10),
                    (8, 64, 11)]


def valu_probe(mode, iters, workgroups, lds_bytes):
    """Duration in ms of one probe launch issuing 128 * iters operations per
    wave (one instruction each, or one exponential for the exp modes) on the
    lanes and chains of `mode` (kinc_selftest_valu_probe)."""
    ms = ctypes.c_float()
    rc = lib().kinc_selftest_valu_probe(mode, iters, workgroups, lds_bytes, ctypes.byref(ms))
    if rc != KINC_OK:
        raise KincError(rc, "kinc_selftest_valu_probe failed")
    return float(ms.value)
