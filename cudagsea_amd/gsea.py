"""Python mirror of the reference's R entry point and of the engine's C ABI.

`GSEA(...)` follows cudaGSEA/R/GSEA.R:1-5 and the Rcpp glue
cudaGSEA/src/cudaGSEA_wrapper.cpp:86-280: same argument names and meaning,
the same input checks (invalid data prints an "INVALID DATA" line and returns
False) and the same result columns ES, NES, NP, FWER with the set names as
row labels; the engine's FDR q-values are appended as a fifth column.
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import check, lib

METRICS = [
    "naive_diff_of_classes", "naive_ratio_of_classes",
    "naive_log2_ratio_of_classes", "stable_diff_of_classes",
    "stable_ratio_of_classes", "stable_log2_ratio_of_classes",
    "onepass_signal2noise", "onepass_t_test",
    "twopass_signal2noise", "twopass_t_test",
    "stable_signal2noise", "stable_t_test",
    "overkill_signal2noise", "overkill_t_test",
]

STAT_ROWS = ("ES", "NES", "NP", "FWER", "FDR")


def metric_id(metric):
    if isinstance(metric, (int, np.integer)):
        return int(metric)
    return lib().gsea_metric_id(metric.encode())


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _cstrings(strs):
    enc = [s.encode() if isinstance(s, str) else bytes(s) for s in strs]
    arr = (ctypes.c_char_p * max(1, len(enc)))(*enc)
    return arr, enc


def set_tuning(**kw):
    """Launch-shape tuning (gsea_set_tuning); no arguments = automatic.
    metric_gene_tile_kind / metric_perms_per_block_kind take six values, one
    per accumulator kind (0: the kind-independent knob)."""
    t = _lib.gsea_tuning()
    for k, v in kw.items():
        if k in _lib.gsea_tuning.KIND_KNOBS:
            vals = [int(x) for x in v]
            if len(vals) != 6:
                raise ValueError("%s takes 6 values, one per accumulator kind" % k)
            setattr(t, k, (ctypes.c_int16 * 6)(*vals))
        else:
            setattr(t, k, int(v))
    check(lib().gsea_set_tuning(ctypes.byref(t)))


def get_tuning():
    t = _lib.gsea_tuning()
    check(lib().gsea_get_tuning(ctypes.byref(t)))
    return t.as_dict()


# ---------------------------------------------------------------------------
# engine calls on host arrays
# ---------------------------------------------------------------------------
def map_sets(gene_symbols, gene_sets):
    """Gene-set symbol lists -> (set_offsets[G+1], set_genes) with the
    create_bitmaps semantics (create_bitmaps.cuh:13-65)."""
    gsym, _k1 = _cstrings(gene_symbols)
    flat = [s for gs in gene_sets for s in gs]
    fsym, _k2 = _cstrings(flat)
    offs = np.zeros(len(gene_sets) + 1, dtype=np.int64)
    offs[1:] = np.cumsum([len(gs) for gs in gene_sets])
    out_off = np.zeros(len(gene_sets) + 1, dtype=np.int32)
    total = ctypes.c_int64(0)
    L = lib()
    check(L.gsea_map_sets(ctypes.cast(gsym, ctypes.c_void_p), len(gene_symbols),
                          ctypes.cast(fsym, ctypes.c_void_p), _ptr(offs),
                          len(gene_sets), _ptr(out_off), None, 0,
                          ctypes.byref(total)))
    genes = np.zeros(max(1, total.value), dtype=np.int32)
    check(L.gsea_map_sets(ctypes.cast(gsym, ctypes.c_void_p), len(gene_symbols),
                          ctypes.cast(fsym, ctypes.c_void_p), _ptr(offs),
                          len(gene_sets), _ptr(out_off), _ptr(genes),
                          genes.size, ctypes.byref(total)))
    return out_off, genes[:total.value]


def _flags(ascending, swap_labels, signed_nes):
    return ((_lib.FLAG_ASCENDING if ascending else 0) |
            (_lib.FLAG_SWAP_LABELS if swap_labels else 0) |
            (_lib.FLAG_SIGNED_NES if signed_nes else 0))


def run(X, labels, set_offsets, set_genes, num_perms, metric,
        ascending=False, swap_labels=False, double=False, want_es=False,
        timings=False, signed_nes=False):
    """Whole path on host arrays.  X is sample-major [S][N].
    Returns stats [5][G] (and es [G][P] if want_es, timings dict if asked).
    signed_nes: Broad GSEA's per-sign normalisation of NES (GSEA_FLAG_SIGNED_NES)
    instead of the reference's single factor."""
    dt = np.float64 if double else np.float32
    X = np.ascontiguousarray(X, dtype=dt)
    S, N = X.shape
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    set_offsets = np.ascontiguousarray(set_offsets, dtype=np.int32)
    set_genes = np.ascontiguousarray(set_genes, dtype=np.int32)
    if set_genes.size == 0:
        set_genes = np.zeros(1, dtype=np.int32)
    G = set_offsets.size - 1
    mid = metric_id(metric)
    stats = np.empty((5, G), dtype=dt)
    es = np.empty((G, num_perms), dtype=dt) if want_es else None
    flags = _flags(ascending, swap_labels, signed_nes)
    t = _lib.gsea_timings()
    fn = lib().gsea_run_f64 if double else lib().gsea_run_f32
    check(fn(_ptr(X), N, S, _ptr(labels), _ptr(set_offsets), _ptr(set_genes),
             G, num_perms, mid, flags, _ptr(es), _ptr(stats), ctypes.byref(t)))
    out = [stats]
    if want_es:
        out.append(es)
    if timings:
        out.append(t.as_dict())
    return out[0] if len(out) == 1 else tuple(out)


# ---------------------------------------------------------------------------
# engine calls on device tensors (torch is only the allocator here)
# ---------------------------------------------------------------------------
def _dp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def run_device(X, labels, set_offsets, set_genes, num_perms, metric,
               stats_out, es_out=None, ascending=False, swap_labels=False,
               stream=None, signed_nes=False):
    """All arrays are device tensors except `labels` (host, S bytes).
    X [S][N] float32/float64, set CSR int32, stats_out [5][G], es_out [G][P]."""
    S, N = X.shape
    G = set_offsets.numel() - 1
    double = X.dtype.itemsize == 8
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    flags = _flags(ascending, swap_labels, signed_nes)
    t = _lib.gsea_timings()
    fn = lib().gsea_run_device_f64 if double else lib().gsea_run_device_f32
    st = ctypes.c_void_p(stream) if stream else None
    check(fn(_dp(X), N, S, _ptr(labels), _dp(set_offsets), _dp(set_genes), G,
             num_perms, metric_id(metric), flags, _dp(es_out), _dp(stats_out),
             st, ctypes.byref(t)))
    return t.as_dict()


def sweep_device(X, labels, set_offsets, set_genes, num_perms, metrics, stats_out,
                 ascending=False, swap_labels=False, graph=True, stream=None,
                 signed_nes=False):
    """Several metrics over the same device inputs (gsea_sweep_device_*):
    stats_out is a [len(metrics)][5][G] device tensor.  With graph=True the
    launch sequence is one HIP graph, captured on the first call and replayed
    by later identical calls.  Returns the timings dict (ms_total)."""
    S, N = X.shape
    G = set_offsets.numel() - 1
    double = X.dtype.itemsize == 8
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    mids = np.ascontiguousarray([metric_id(m) for m in metrics], dtype=np.int32)
    flags = _flags(ascending, swap_labels, signed_nes)
    t = _lib.gsea_timings()
    fn = lib().gsea_sweep_device_f64 if double else lib().gsea_sweep_device_f32
    st = ctypes.c_void_p(stream) if stream else None
    check(fn(_dp(X), N, S, _ptr(labels), _dp(set_offsets), _dp(set_genes), G, num_perms,
             _ptr(mids), len(mids), flags, int(bool(graph)), _dp(stats_out), st,
             ctypes.byref(t)))
    return t.as_dict()


def run_budget_device(X, labels, set_offsets, set_genes, num_perms, metric,
                      budget, stats_out, es_pg_out=None, extension=None,
                      ascending=False, swap_labels=False, stream=None, signed_nes=False):
    """Budgeted run: set s is scored for permutations 0..budget[s]-1 (host
    int array, each entry dividing num_perms).  extension = (fdr, z,
    nes_tol) or None (see gsea_run_budget_device_f32).  Returns (final budgets, timings).  es_pg_out, if
    given, is a [P][G] device tensor (entries beyond a set's budget are
    unspecified)."""
    S, N = X.shape
    G = set_offsets.numel() - 1
    double = X.dtype.itemsize == 8
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    bud = np.ascontiguousarray(budget, dtype=np.int32)
    if bud.shape != (G,):
        raise ValueError("budget must hold one entry per set")
    out = np.zeros(G, dtype=np.int32)
    ext = None
    if extension is not None:
        ext = _lib.gsea_extension(*[float(v) for v in extension])
    flags = _flags(ascending, swap_labels, signed_nes)
    t = _lib.gsea_timings()
    fn = lib().gsea_run_budget_device_f64 if double else lib().gsea_run_budget_device_f32
    st = ctypes.c_void_p(stream) if stream else None
    check(fn(_dp(X), N, S, _ptr(labels), _dp(set_offsets), _dp(set_genes), G,
             num_perms, metric_id(metric), flags, _ptr(bud),
             ctypes.byref(ext) if ext is not None else None, _ptr(out),
             _dp(es_pg_out), _dp(stats_out), st, ctypes.byref(t)))
    return out, t.as_dict()


def metric_device(X, labels, num_perms, metric, r_out, shift=0,
                  swap_labels=False, stream=None):
    """Stage (a): r_out[P][N] for global permutations shift..shift+P-1."""
    S, N = X.shape
    double = X.dtype.itemsize == 8
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    fn = lib().gsea_metric_f64 if double else lib().gsea_metric_f32
    st = ctypes.c_void_p(stream) if stream else None
    check(fn(_dp(X), N, S, _ptr(labels), num_perms, shift, metric_id(metric),
             _lib.FLAG_SWAP_LABELS if swap_labels else 0, _dp(r_out), st))


def rank_es_device(r, set_offsets, set_genes, es_out, ascending=False,
                   stream=None):
    """Stages (b)+(c): es_out[P][G] from r[P][N]."""
    P, N = r.shape
    G = set_offsets.numel() - 1
    double = r.dtype.itemsize == 8
    fn = lib().gsea_rank_es_f64 if double else lib().gsea_rank_es_f32
    st = ctypes.c_void_p(stream) if stream else None
    check(fn(_dp(r), N, P, _dp(set_offsets), _dp(set_genes), G,
             _lib.FLAG_ASCENDING if ascending else 0, _dp(es_out), st))


def argsort_device(keys, index_out, ascending=False, stream=None):
    """Stage (b): index_out[P][N] (int32 tensor) = gene at each rank."""
    P, N = keys.shape
    double = keys.dtype.itemsize == 8
    fn = lib().gsea_argsort_f64 if double else lib().gsea_argsort_f32
    st = ctypes.c_void_p(stream) if stream else None
    check(fn(_dp(keys), N, P, _lib.FLAG_ASCENDING if ascending else 0,
             _dp(index_out), st))


# ---------------------------------------------------------------------------
# R-level entry point
# ---------------------------------------------------------------------------
def _invalid(msg):
    print("INVALID DATA: " + msg)
    return False


def GSEA(exprsData, labelList, geneSets, numPermutations, metricString="",
         dumpFileName="", checkInput=True, doublePrecision=False):
    """Mirror of R `GSEA()` (R/GSEA.R:1-5, cudaGSEA_wrapper.cpp:86-280).

    exprsData: pandas.DataFrame, genes x samples, index = gene symbols,
               columns = sample names (the R data.frame).
    labelList: 0/1 per sample.
    geneSets:  dict name -> list of gene symbols (an R named list) or a list
               of lists (named "1".."n" with a warning, as the glue does).
    Returns a pandas.DataFrame with columns ES, NES, NP, FWER, FDR and the set
    names as index, or False for invalid input (checkInput=True).
    """
    import pandas as pd

    df = exprsData if isinstance(exprsData, pd.DataFrame) else pd.DataFrame(exprsData)
    rownames = [str(x) for x in df.index]
    colnames = [str(x) for x in df.columns]
    table = df.to_numpy(dtype=np.float64)

    if isinstance(geneSets, dict):
        names = [str(k) for k in geneSets.keys()]
        sets = [[str(g) for g in v] for v in geneSets.values()]
    else:
        sets = [[str(g) for g in v] for v in geneSets]
        names = [str(i + 1) for i in range(len(sets))]
        print("WARNING: We enumerated your gene set names from 1 to n.")
    labels = list(labelList)

    if checkInput:
        # cudaGSEA_wrapper.cpp:156-227
        if table.shape[0] != len(rownames):
            return _invalid("the vertical dimension and the number of row "
                            "names do not comply.")
        if table.shape[1] != len(colnames):
            return _invalid("the horizontal dimension and the number of column "
                            "names do not comply.")
        if len(labels) != table.shape[1]:
            return _invalid("Your label list has a different amount of entries "
                            "as the number of columns in your gene  expression "
                            "data frame.")
        for lab in labels:
            if lab != 0 and lab != 1:
                return _invalid("Label vector may only contain zeros and ones "
                                "since cudaGSEA only supports two classes of "
                                "phenotypes.")
        if np.isnan(table).any():
            return _invalid("expression table has NaN entries that cannot be "
                            "handled by the GSEA algorithm.")
        if table.size and table.min() < 0:
            print("WARNING: There are negative expression values in the "
                  "provided table. Some metrics cannot handle negative "
                  "amplitudes e.g. the log2_ratio_of_classes family.")
        if len(sets) == 0:
            return _invalid("you must at least specify one gene set.")
        for gs in sets:
            if len(gs) == 0:
                return _invalid("There is an empty gene set (no gene "
                                "identifiers) in your list of gene sets.")

    dt = np.float64 if doublePrecision else np.float32
    X = np.ascontiguousarray(table.T, dtype=dt)           # sample-major [S][N]
    S, N = X.shape
    lab = np.ascontiguousarray(np.asarray(labels, dtype=np.int64) & 0xff,
                               dtype=np.uint8)
    gsym, _k1 = _cstrings(rownames)
    flat = [s for gs in sets for s in gs]
    fsym, _k2 = _cstrings(flat)
    offs = np.zeros(len(sets) + 1, dtype=np.int64)
    offs[1:] = np.cumsum([len(gs) for gs in sets])
    G = len(sets)
    stats = np.empty((5, G), dtype=dt)
    check(lib().gsea_proxy_run(
        1 if doublePrecision else 0, _ptr(X), N, S,
        ctypes.cast(gsym, ctypes.c_void_p), _ptr(lab),
        ctypes.cast(fsym, ctypes.c_void_p), _ptr(offs), G,
        (metricString or "").encode(), int(numPermutations), 1, 0,
        (dumpFileName or "").encode(), _ptr(stats), None))
    return pd.DataFrame({k: stats[i].astype(np.float64)
                         for i, k in enumerate(STAT_ROWS)}, index=names)


def read_dump(path):
    """Read a <name>_<G>_<P>_<32|64>.es dump into a [G][P] array (the layout
    tools/reader.py of the reference parses)."""
    import os
    base = os.path.basename(path)
    stem = base[:-3] if base.endswith(".es") else base
    parts = stem.split("_")
    G, P, bits = int(parts[-3]), int(parts[-2]), int(parts[-1])
    dt = np.float64 if bits == 64 else np.float32
    return np.fromfile(path, dtype=dt).reshape(G, P)


def device_info():
    L = lib()
    n = ctypes.c_int(0)
    check(L.gsea_device_count(ctypes.byref(n)))
    out = []
    for i in range(n.value):
        buf = ctypes.create_string_buffer(256)
        mem = ctypes.c_size_t(0)
        check(L.gsea_device_name(i, buf, 256, ctypes.byref(mem)))
        out.append("device %d: %s with %d GiB RAM" % (i, buf.value.decode(),
                                                       mem.value >> 30))
    return out
