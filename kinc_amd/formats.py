"""Python access to the EMX / CCM / CMX files of include/kinc_formats.h.

Writing goes through the C++ library (kinc_emx_write, the Similarity writer);
reading is a small numpy parser of the same layout, used by tests and by
callers that want the records as arrays.
"""
import ctypes
import json
import struct

import numpy as np

from . import _native

MAGIC = b"KINCAMD1"
TYPES = {1: "emx", 2: "ccm", 3: "cmx"}


_P = ctypes.c_void_p
SIGNATURES = {
    "kinc_emx_write": (ctypes.c_int, [ctypes.c_char_p, _P, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_char_p]),
    "kinc_emx_info": (ctypes.c_int, [ctypes.c_char_p, _P, _P]),
    "kinc_emx_read": (ctypes.c_int, [ctypes.c_char_p, _P]),
    "kinc_file_metadata": (ctypes.c_int64, [ctypes.c_char_p, _P, ctypes.c_int64]),
    "kinc_similarity_writer_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p,
                                                   ctypes.c_int32, ctypes.c_int32,
                                                   ctypes.c_int32, ctypes.c_int32,
                                                   ctypes.c_float, ctypes.c_float,
                                                   ctypes.c_char_p, _P]),
    "kinc_similarity_writer_process": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int64, _P, _P,
                                                      _P, ctypes.c_int32]),
    "kinc_similarity_writer_close": (ctypes.c_int, [_P]),
    "kinc_formats_last_error": (ctypes.c_char_p, []),
    "kinc_emx_import_text": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                            _P, _P]),
    "kinc_emx_export_text": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
    "kinc_export_cmx_text": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
    "kinc_extract_network": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                            ctypes.c_char_p, ctypes.c_char_p, ctypes.c_float,
                                            ctypes.c_float]),
}
# ExpressionMatrix::TRANSFORM_NAMES (src/core/expressionmatrix.cpp:8-14)
TRANSFORMS = ["none", "natural logarithm", "logarithm base 2", "logarithm base 10"]
_bound = False


def _bind():
    global _bound
    l = _native.lib()
    if _bound:
        return l
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(l, name)
        fn.restype = res
        fn.argtypes = args
    _bound = True
    return l


def _check(rc):
    if rc != 0:
        raise _native.KincError(rc, _bind().kinc_formats_last_error().decode())


def write_emx(path, data, genes=None, samples=None, transform="none"):
    data = np.ascontiguousarray(data, np.float32)
    G, S = data.shape
    meta = {"genes": genes or ["g%d" % i for i in range(G)],
            "samples": samples or ["s%d" % i for i in range(S)],
            "transform": transform}
    _check(_bind().kinc_emx_write(path.encode(), data.ctypes.data, G, S,
                                  json.dumps(meta).encode()))


def import_text(text_path, emx_path, nan=None, samples=0, transform="none", threads=0):
    """ImportExpressionMatrix: whitespace-delimited text to an EMX file
    (kinc_emx_import_text).  Returns (genes, samples) sizes."""
    G = ctypes.c_int32()
    S = ctypes.c_int32()
    _check(_bind().kinc_emx_import_text(
        text_path.encode(), emx_path.encode(), nan.encode() if nan else None, samples,
        TRANSFORMS.index(transform), threads, ctypes.byref(G), ctypes.byref(S)))
    return G.value, S.value


def export_text(emx_path, text_path, nan=""):
    """ExportExpressionMatrix: an EMX file to text (kinc_emx_export_text)."""
    _check(_bind().kinc_emx_export_text(emx_path.encode(), text_path.encode(),
                                        (nan or "").encode()))


def export_cmx_text(ccm, cmx, out):
    """ExportCorrelationMatrix: CCM + CMX to text (kinc_export_cmx_text)."""
    _check(_bind().kinc_export_cmx_text(ccm.encode(), cmx.encode(), out.encode()))


def extract_network(emx, ccm, cmx, out, graphml=None, min_corr=0.85, max_corr=1.0):
    """Extract: the thresholded network as an edge list (+ GraphML)."""
    _check(_bind().kinc_extract_network(emx.encode(), ccm.encode(), cmx.encode(), out.encode(),
                                        graphml.encode() if graphml else None, min_corr,
                                        max_corr))


def read_emx_native(path):
    l = _bind()
    G = ctypes.c_int32()
    S = ctypes.c_int32()
    _check(l.kinc_emx_info(path.encode(), ctypes.byref(G), ctypes.byref(S)))
    out = np.empty((G.value, S.value), np.float32)
    _check(l.kinc_emx_read(path.encode(), out.ctypes.data))
    return out


class SimilarityWriter:
    """Similarity::process into CCM + CMX files (kinc_similarity_writer_*)."""

    def __init__(self, ccm, cmx, G, S, max_clusters, corr_method="pearson", min_corr=0.5,
                 max_corr=1.0, metadata=None):
        self.S = S
        self.MC = max_clusters
        h = ctypes.c_void_p()
        _check(_bind().kinc_similarity_writer_open(
            ccm.encode(), cmx.encode(), G, S, max_clusters, _native.CORR[corr_method],
            min_corr, max_corr, json.dumps(metadata or {}).encode(), ctypes.byref(h)))
        self.h = h

    def process(self, start, K, labels, corr):
        K = np.ascontiguousarray(K, np.int8)
        labels = np.ascontiguousarray(labels, np.int8)
        corr = np.ascontiguousarray(corr, np.float32)
        _check(_bind().kinc_similarity_writer_process(self.h, start, K.shape[0], K.ctypes.data,
                                                      labels.ctypes.data, corr.ctypes.data,
                                                      corr.shape[1]))

    def close(self):
        if self.h:
            _check(_bind().kinc_similarity_writer_close(self.h))
            self.h = None


def read_file(path):
    """Parse a KINCAMD1 file: returns (type, payload bytes, metadata dict)."""
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:8] != MAGIC:
        raise ValueError("%s: not a KINCAMD1 file" % path)
    typ, _, poff, psize, moff, msize = struct.unpack_from("<IIQQQQ", raw, 8)
    meta = json.loads(raw[moff:moff + msize].decode() or "{}")
    return TYPES[typ], raw[poff:poff + psize], meta


def read_emx(path):
    typ, payload, meta = read_file(path)
    assert typ == "emx"
    G, S = struct.unpack_from("<ii", payload, 0)
    return np.frombuffer(payload, np.float32, G * S, 8).reshape(G, S), meta


def read_pairwise(path):
    """Header fields and records of a CCM or CMX file.

    Returns (header dict, records) where records is a structured array with
    fields x, y, cluster and data (dataSize raw bytes)."""
    typ, payload, meta = read_file(path)
    G, MC, dsize, npairs, nclus, off = struct.unpack_from("<iiiqqh", payload, 0)
    hdr = dict(type=typ, genes=G, max_clusters=MC, data_size=dsize, pairs=npairs,
               clusters=nclus, offset=off, meta=meta)
    sub = payload[30:30 + off]
    if typ == "ccm":
        hdr["samples"] = struct.unpack_from("<i", sub, 0)[0]
    else:
        hdr["correlations"] = struct.unpack_from("<b", sub, 0)[0]
    dt = np.dtype([("x", "<i4"), ("y", "<i4"), ("cluster", "i1"), ("data", "u1", (dsize,))])
    rec = np.frombuffer(payload, dt, nclus, 30 + off)
    return hdr, rec


def ccm_masks(rec, S):
    """Unpack CCM nibble records into [n][S] sample values."""
    d = rec["data"]
    lo = d & 0x0F
    hi = (d >> 4) & 0x0F
    out = np.empty((d.shape[0], 2 * d.shape[1]), np.int8)
    out[:, 0::2] = lo
    out[:, 1::2] = hi
    return out[:, :S]


def cmx_values(rec):
    return rec["data"].copy().view("<f4")[:, 0]
