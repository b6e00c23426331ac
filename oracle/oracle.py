"""ctypes front of the CPU oracle (oracle/_build/liboracle.so) and of the
reference's permutation generator built from its own header
(oracle/_ref/libgsea_ref.so, rngpu/rngpu.hpp compiled by g++ as shipped).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg.  The product (cudagsea_amd/) never
imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GSEA_ORACLE_LIB selects another build of the oracle (used to check that the
# tests fail on an oracle with a planted mistake)
LIB_PATH = os.environ.get("GSEA_ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libgsea_ref.so")
REF_SOURCES = "/root/reference/cudaGSEA/src/include"

METRICS = [
    "naive_diff_of_classes", "naive_ratio_of_classes",
    "naive_log2_ratio_of_classes", "stable_diff_of_classes",
    "stable_ratio_of_classes", "stable_log2_ratio_of_classes",
    "onepass_signal2noise", "onepass_t_test",
    "twopass_signal2noise", "twopass_t_test",
    "stable_signal2noise", "stable_t_test",
    "overkill_signal2noise", "overkill_t_test",
]

_P = ctypes.c_void_p
_I64 = ctypes.c_int64


def build(ref=True):
    """Compile the oracle (and, where the reference sources exist, _ref)."""
    subprocess.check_call(["make", "-s", "-C", HERE])
    if ref and os.path.isdir(REF_SOURCES):
        subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "ref")])


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = ctypes.CDLL(LIB_PATH)
        L.orc_hash32.restype = ctypes.c_uint32
        L.orc_hash32.argtypes = [ctypes.c_uint32]
        L.orc_permute_labels.argtypes = [_P, _I64, ctypes.c_uint64, _P]
        L.orc_kiss_init.argtypes = [ctypes.c_uint32, _P]
        L.orc_kiss_next.restype = ctypes.c_uint32
        L.orc_kiss_next.argtypes = [_P]
        L.orc_log_f64.restype = ctypes.c_double
        L.orc_log_f64.argtypes = [ctypes.c_double]
        L.orc_log_f32.restype = ctypes.c_float
        L.orc_log_f32.argtypes = [ctypes.c_float]
        L.orc_metric_id.restype = ctypes.c_int
        L.orc_metric_id.argtypes = [ctypes.c_char_p]
        L.orc_num_threads.restype = ctypes.c_int
        for sfx in ("f32", "f64"):
            getattr(L, "orc_correlate_" + sfx).argtypes = [
                _P, _P, _I64, _I64, _I64, _I64, ctypes.c_int, ctypes.c_int, _P]
            getattr(L, "orc_pathway_sort_" + sfx).argtypes = [
                _P, _P, _I64, _I64, ctypes.c_int]
            getattr(L, "orc_scores_" + sfx).argtypes = [
                _P, _P, _P, _P, _I64, _I64, _I64, _P]
            getattr(L, "orc_scores_ext_" + sfx).argtypes = [
                _P, _P, _P, _P, _I64, _I64, _I64, _P, _P, _P]
            getattr(L, "orc_final_statistics_" + sfx).argtypes = [
                _P, _I64, _I64, _P]
            getattr(L, "orc_final_statistics_ext_" + sfx).argtypes = [
                _P, _I64, _I64, ctypes.c_int, _P]
            f = getattr(L, "orc_gsea_" + sfx)
            f.argtypes = [_P, _I64, _I64, _P, _P, _P, _I64, _I64,
                          ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P]
            f.restype = ctypes.c_int
        L.orc_sort_key_f32.restype = ctypes.c_uint32
        L.orc_sort_key_f32.argtypes = [ctypes.c_float, ctypes.c_int]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_PATH)
        R.ref_permute_labels.argtypes = [_P, _I64, ctypes.c_uint64, _P]
        R.ref_hash32.restype = ctypes.c_uint32
        R.ref_hash32.argtypes = [ctypes.c_uint32]
        R.ref_kiss_stream.argtypes = [ctypes.c_uint32, _P, _I64]
        _ref = R
    return _ref


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dt(double):
    return np.float64 if double else np.float32


def kiss_stream(seed, n, use_ref=False):
    """The first n draws of the generator seeded with `seed`."""
    out = np.empty(n, dtype=np.uint32)
    if use_ref:
        ref().ref_kiss_stream(seed, _ptr(out), n)
    else:
        st = (ctypes.c_uint32 * 5)()
        lib().orc_kiss_init(seed, st)
        for i in range(n):
            out[i] = lib().orc_kiss_next(st)
    return out


def permute_labels(labels, perm, use_ref=False):
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    out = np.empty_like(labels)
    L = ref() if use_ref else lib()
    fn = L.ref_permute_labels if use_ref else L.orc_permute_labels
    fn(_ptr(labels), labels.size, perm, _ptr(out))
    return out


def correlate(X, labels, P, metric, shift=0, sample_major=False, double=False):
    """Metric matrix [P][N].  X is [N][S] (gene-major) unless sample_major."""
    dt = _dt(double)
    X = np.ascontiguousarray(X, dtype=dt)
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    if sample_major:
        S, N = X.shape
    else:
        N, S = X.shape
    mid = METRICS.index(metric) if isinstance(metric, str) else int(metric)
    out = np.empty((P, N), dtype=dt)
    sfx = "f64" if double else "f32"
    getattr(lib(), "orc_correlate_" + sfx)(
        _ptr(X), _ptr(labels), N, S, P, shift, mid, int(sample_major), _ptr(out))
    return out


def pathway_sort(keys, ascending=False):
    """Returns (sorted keys [P][N], gene index at each rank [P][N])."""
    keys = np.array(keys, copy=True, order="C")
    double = keys.dtype == np.float64
    P, N = keys.shape
    idx = np.empty((P, N), dtype=np.uint32)
    getattr(lib(), "orc_pathway_sort_" + ("f64" if double else "f32"))(
        _ptr(keys), _ptr(idx), N, P, int(ascending))
    return keys, idx


def scores(sorted_keys, index, offsets, genes):
    sorted_keys = np.ascontiguousarray(sorted_keys)
    double = sorted_keys.dtype == np.float64
    P, N = sorted_keys.shape
    index = np.ascontiguousarray(index, dtype=np.uint32)
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    genes = np.ascontiguousarray(genes, dtype=np.int32)
    G = offsets.size - 1
    es = np.empty((G, P), dtype=sorted_keys.dtype)
    getattr(lib(), "orc_scores_" + ("f64" if double else "f32"))(
        _ptr(sorted_keys), _ptr(index), _ptr(offsets), _ptr(genes), G, N, P,
        _ptr(es))
    return es


def scores_ext(sorted_keys, index, offsets, genes):
    """(es, hi, lo) [G][P]: the scores and each walk's largest and smallest
    running sum."""
    sorted_keys = np.ascontiguousarray(sorted_keys)
    double = sorted_keys.dtype == np.float64
    P, N = sorted_keys.shape
    index = np.ascontiguousarray(index, dtype=np.uint32)
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    genes = np.ascontiguousarray(genes, dtype=np.int32)
    G = offsets.size - 1
    es, hi, lo = (np.empty((G, P), dtype=sorted_keys.dtype) for _ in range(3))
    getattr(lib(), "orc_scores_ext_" + ("f64" if double else "f32"))(
        _ptr(sorted_keys), _ptr(index), _ptr(offsets), _ptr(genes), G, N, P,
        _ptr(es), _ptr(hi), _ptr(lo))
    return es, hi, lo


def gsea_ext(X, labels, offsets, genes, P, metric, ascending=False, swap=False,
             double=False, signed_nes=False):
    """Whole CPU path with each walk's extremes: (stats, es, hi, lo)."""
    dt = _dt(double)
    X = np.ascontiguousarray(X, dtype=dt)
    lab = np.ascontiguousarray(labels, dtype=np.uint8)
    if swap:
        lab = (lab + 1) % 2
    r = correlate(X, lab, P, metric, sample_major=True, double=double)
    sk, idx = pathway_sort(r, ascending=ascending)
    es, hi, lo = scores_ext(sk, idx, offsets, genes)
    st, _ = final_statistics(es.copy(), signed_nes=signed_nes)
    return st, es, hi, lo


def final_statistics(es, signed_nes=False):
    """Returns (stats [5][G], normalised es).  signed_nes: Broad GSEA's
    per-sign normalisation (orc_final_statistics_ext)."""
    es = np.array(es, copy=True, order="C")
    double = es.dtype == np.float64
    G, P = es.shape
    out = np.empty((5, G), dtype=es.dtype)
    sfx = "f64" if double else "f32"
    getattr(lib(), "orc_final_statistics_ext_" + sfx)(_ptr(es), G, P,
                                                      4 if signed_nes else 0, _ptr(out))
    return out, es


def gsea(X, labels, offsets, genes, P, metric, ascending=False, swap=False,
         sample_major=True, double=False, want_es=True, signed_nes=False):
    """Whole reference CPU path.  Returns (stats [5][G], es [G][P] or None)."""
    dt = _dt(double)
    X = np.ascontiguousarray(X, dtype=dt)
    if sample_major:
        S, N = X.shape
    else:
        N, S = X.shape
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    genes = np.ascontiguousarray(genes, dtype=np.int32)
    G = offsets.size - 1
    mid = METRICS.index(metric) if isinstance(metric, str) else int(metric)
    stats = np.empty((5, G), dtype=dt)
    es = np.empty((G, P), dtype=dt) if want_es else None
    flags = (1 if ascending else 0) | (2 if swap else 0) | (4 if signed_nes else 0)
    rc = getattr(lib(), "orc_gsea_" + ("f64" if double else "f32"))(
        _ptr(X), N, S, _ptr(labels), _ptr(offsets), _ptr(genes), G, P, mid,
        flags, int(sample_major), _ptr(es) if want_es else None, _ptr(stats))
    if rc != 0:
        raise ValueError("oracle gsea failed with code %d" % rc)
    return stats, es
