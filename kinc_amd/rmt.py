"""Python access to the RMT threshold analytic of include/kinc_rmt.h."""
import ctypes

import numpy as np

from . import _native

_P = ctypes.c_void_p


class RmtParams(ctypes.Structure):
    _fields_ = [("threshold_start", ctypes.c_float), ("threshold_step", ctypes.c_float),
                ("threshold_stop", ctypes.c_float), ("chi_square1", ctypes.c_float),
                ("chi_square2", ctypes.c_float), ("min_eigenvalues", ctypes.c_int32),
                ("min_pace", ctypes.c_int32), ("max_pace", ctypes.c_int32),
                ("bins", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3)]


_PP = ctypes.POINTER(RmtParams)
SIGNATURES = {
    "kinc_rmt_default_params": (None, [_PP]),
    "kinc_rmt_create": (ctypes.c_int, [_PP, ctypes.c_int32, ctypes.c_int32, _P, _P, _P, _P,
                                       ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(_P)]),
    "kinc_rmt_eigenvalues": (ctypes.c_int, [_P, ctypes.c_float, _P, ctypes.c_int64, _P]),
    "kinc_rmt_chi_square": (ctypes.c_float, [_PP, _P, ctypes.c_int64]),
    "kinc_rmt_run": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int64, _P, _P]),
    "kinc_rmt_run_file": (ctypes.c_int, [_PP, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32,
                                         _P]),
    "kinc_rmt_destroy": (None, [_P]),
    "kinc_rmt_last_error": (ctypes.c_char_p, [_P]),
}
_bound = False

# command-line names of the reference (rmt_input.cpp:58-136)
FLAGS = {"tstart": "threshold_start", "tstep": "threshold_step", "tstop": "threshold_stop",
         "minpace": "min_pace", "maxpace": "max_pace", "bins": "bins"}


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


def make_params(**kw):
    p = RmtParams()
    _bind().kinc_rmt_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, FLAGS.get(k, k), v)
    return p


def chi_square(eigens, **kw):
    """RMT::computeChiSquare of |value|-ordered float eigenvalues (host)."""
    e = np.ascontiguousarray(eigens, np.float32)
    return _bind().kinc_rmt_chi_square(ctypes.byref(make_params(**kw)), e.ctypes.data, e.size)


def records_from_dense(matrix):
    """CMX-ordered records (x > y, cluster) of the non-zero entries of a
    dense [G][G][K] correlation matrix."""
    m = np.asarray(matrix, np.float32)
    G, _, K = m.shape
    xs, ys = np.tril_indices(G, -1)
    order = np.lexsort((ys, xs))            # by x then y: the pair index order
    xs, ys = xs[order], ys[order]
    vals = m[xs, ys, :]                      # [pairs][K]
    pi, ki = np.nonzero(vals)
    return (xs[pi].astype(np.int32), ys[pi].astype(np.int32), ki.astype(np.int8),
            vals[pi, ki].astype(np.float32))


class Rmt:
    """A correlation matrix resident on the GPU (kinc_rmt_create)."""

    def __init__(self, G, K, x, y, cluster, corr, device=0, **params):
        self.params = make_params(**params)
        self.x = np.ascontiguousarray(x, np.int32)
        self.y = np.ascontiguousarray(y, np.int32)
        self.k = np.ascontiguousarray(cluster, np.int8)
        self.r = np.ascontiguousarray(corr, np.float32)
        h = ctypes.c_void_p()
        l = _bind()
        rc = l.kinc_rmt_create(ctypes.byref(self.params), G, K, self.x.ctypes.data,
                               self.y.ctypes.data, self.k.ctypes.data, self.r.ctypes.data,
                               self.x.size, device, ctypes.byref(h))
        if rc != 0:
            raise _native.KincError(rc, l.kinc_rmt_last_error(None).decode())
        self.h = h

    @classmethod
    def from_dense(cls, matrix, device=0, **params):
        m = np.asarray(matrix, np.float32)
        return cls(m.shape[0], m.shape[2], *records_from_dense(m), device=device, **params)

    def _check(self, rc):
        if rc != 0:
            raise _native.KincError(rc, _bind().kinc_rmt_last_error(self.h).decode())

    def eigenvalues(self, threshold):
        n = ctypes.c_int64()
        self._check(_bind().kinc_rmt_eigenvalues(self.h, threshold, None, 0, ctypes.byref(n)))
        out = np.empty(max(n.value, 1), np.float32)
        self._check(_bind().kinc_rmt_eigenvalues(self.h, threshold, out.ctypes.data, out.size,
                                                 ctypes.byref(n)))
        return out[:n.value]

    def run(self, cap=2000):
        """RMT::process: (rc, final threshold, [(t, size, chi), ...])."""
        lt = np.empty(cap, np.float32)
        ls = np.empty(cap, np.int32)
        lc = np.empty(cap, np.float32)
        steps = ctypes.c_int64()
        ft = ctypes.c_float()
        rc = _bind().kinc_rmt_run(self.h, lt.ctypes.data, ls.ctypes.data, lc.ctypes.data, cap,
                                  ctypes.byref(steps), ctypes.byref(ft))
        n = min(steps.value, cap)
        return rc, ft.value, list(zip(lt[:n].tolist(), ls[:n].tolist(), lc[:n].tolist()))

    def last_error(self):
        return _bind().kinc_rmt_last_error(self.h).decode()

    def close(self):
        if self.h:
            _bind().kinc_rmt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_file(cmx, log, device=0, **params):
    """kinc run rmt: the sweep over a CMX file, the reference's log written."""
    p = make_params(**params)
    ft = ctypes.c_float()
    l = _bind()
    rc = l.kinc_rmt_run_file(ctypes.byref(p), cmx.encode(), log.encode(), device,
                             ctypes.byref(ft))
    if rc != 0:
        raise _native.KincError(rc, l.kinc_rmt_last_error(None).decode())
    return ft.value
