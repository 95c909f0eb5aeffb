"""Python access to the query server of include/kinc_server.h."""
import ctypes

import numpy as np

from . import _native

_P = ctypes.c_void_p


class Hit(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32), ("cluster", ctypes.c_int32),
                ("corr", ctypes.c_float)]


HIT_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("cluster", "<i4"), ("corr", "<f4")])

SIGNATURES = {
    "kinc_server_create": (ctypes.c_int, [ctypes.POINTER(_native.SimParams), _P, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                          ctypes.c_int32, ctypes.POINTER(_P)]),
    "kinc_server_submit": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_float, _P]),
    "kinc_server_wait": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_double]),
    "kinc_server_result": (ctypes.c_int, [_P, ctypes.c_int64, _P, ctypes.c_int64, _P, _P, _P]),
    "kinc_server_stats": (ctypes.c_int, [_P, _P, _P, _P]),
    "kinc_server_timing": (ctypes.c_int, [_P, ctypes.c_int64, _P, _P]),
    "kinc_server_destroy": (None, [_P]),
    "kinc_server_last_error": (ctypes.c_char_p, [_P]),
}
_bound = False


def _bind():
    global _bound
    l = _native.lib()
    if not _bound:
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return l


def query_pairs(genes, G):
    """Pair list of a query, in the order the server evaluates it."""
    seen = {}
    uniq = []
    for g in genes:
        if g not in seen:
            seen[g] = len(uniq)
            uniq.append(g)
    out = []
    for i, a in enumerate(uniq):
        for g in range(G):
            if g == a or (g in seen and seen[g] < i):
                continue
            out.append((max(a, g), min(a, g)))
    return np.array(out, np.int32).reshape(-1, 2)


class Server:
    """Continuous-batching query server with the matrix resident on the GPU."""

    def __init__(self, emx, params, hbm_budget=1 << 30, batch_pairs=131072, device=0):
        l = _bind()
        emx = np.ascontiguousarray(emx, np.float32)
        self.G, self.S = emx.shape
        h = _P()
        rc = l.kinc_server_create(ctypes.byref(params), emx.ctypes.data, self.G, self.S, device,
                                  hbm_budget, batch_pairs, ctypes.byref(h))
        if rc != 0:
            raise _native.KincError(rc, l.kinc_server_last_error(None).decode())
        self.h = h

    def _check(self, rc):
        if rc < 0:
            raise _native.KincError(rc, _bind().kinc_server_last_error(self.h).decode())
        return rc

    def submit(self, genes, min_abs_corr=0.5):
        g = np.ascontiguousarray(genes, np.int32)
        qid = ctypes.c_int64()
        self._check(_bind().kinc_server_submit(self.h, g.ctypes.data, g.size, min_abs_corr,
                                               ctypes.byref(qid)))
        return qid.value

    def wait(self, qid, timeout=600.0):
        return self._check(_bind().kinc_server_wait(self.h, qid, timeout)) == 1

    def result(self, qid):
        """(hits as a structured array, latency seconds, pairs evaluated)."""
        l = _bind()
        n = ctypes.c_int64()
        lat = ctypes.c_double()
        npairs = ctypes.c_int64()
        self._check(l.kinc_server_result(self.h, qid, None, 0, ctypes.byref(n), ctypes.byref(lat),
                                         ctypes.byref(npairs)))
        out = np.empty(n.value, HIT_DTYPE)
        self._check(l.kinc_server_result(self.h, qid, out.ctypes.data, n.value, ctypes.byref(n),
                                         ctypes.byref(lat), ctypes.byref(npairs)))
        return out, lat.value, npairs.value

    def timing(self, qid):
        """(seconds from submit to the first block's launch, seconds from
        there to completion) of a finished query not yet released."""
        w = ctypes.c_double()
        r = ctypes.c_double()
        self._check(_bind().kinc_server_timing(self.h, qid, ctypes.byref(w), ctypes.byref(r)))
        return w.value, r.value

    def stats(self):
        pairs = ctypes.c_int64()
        batches = ctypes.c_int64()
        busy = ctypes.c_double()
        self._check(_bind().kinc_server_stats(self.h, ctypes.byref(pairs), ctypes.byref(batches),
                                              ctypes.byref(busy)))
        return pairs.value, batches.value, busy.value

    def close(self):
        if getattr(self, "h", None):
            _bind().kinc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
