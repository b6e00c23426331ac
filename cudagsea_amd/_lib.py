"""ctypes binding of libgsea_hip.so (the C ABI of include/gsea_hip.h).

The HIP library is the only compute path: importing this module fails loudly
when the library has not been built, and there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgsea_hip.so")

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int

# error codes (include/gsea_hip.h)
GSEA_OK = 0
GSEA_ERR_UNKNOWN_METRIC = 1
GSEA_ERR_CANNOT_OPEN_FILE = 2
GSEA_ERR_NO_PATHS = 10
GSEA_ERR_EMPTY_DUMP_FILE_NAME = 19
GSEA_ERR_INVALID_ARGUMENT = 30
GSEA_ERR_UNSUPPORTED_SIZE = 31
GSEA_ERR_HIP = 32
GSEA_ERR_OUT_OF_MEMORY = 33

FLAG_ASCENDING = 1
FLAG_SWAP_LABELS = 2
FLAG_SIGNED_NES = 4

MAX_GENES = 65535
MAX_SAMPLES = 65535


class gsea_extension(ctypes.Structure):
    _fields_ = [("fdr", ctypes.c_double), ("z", ctypes.c_double),
                ("nes_tol", ctypes.c_double)]


class gsea_tuning(ctypes.Structure):
    _fields_ = [("batch_perms", ctypes.c_int64), ("metric_gene_tile", ctypes.c_int32),
                ("metric_perms_per_block", ctypes.c_int32), ("es_waves", ctypes.c_int32),
                ("es_lane_set_max", ctypes.c_int32), ("es_pair_set_max", ctypes.c_int32),
                ("metric_gene_tile_kind", ctypes.c_int16 * 6),
                ("metric_perms_per_block_kind", ctypes.c_int16 * 6),
                ("reserved", ctypes.c_int64 * 2)]
    KNOBS = ("batch_perms", "metric_gene_tile", "metric_perms_per_block", "es_waves",
             "es_lane_set_max", "es_pair_set_max")
    KIND_KNOBS = ("metric_gene_tile_kind", "metric_perms_per_block_kind")

    def as_dict(self):
        out = {k: getattr(self, k) for k in self.KNOBS}
        for k in self.KIND_KNOBS:
            out[k] = list(getattr(self, k))
        return out


class gsea_timings(ctypes.Structure):
    _fields_ = [("ms_lists", ctypes.c_double), ("ms_metric", ctypes.c_double),
                ("ms_rank_es", ctypes.c_double), ("ms_stats", ctypes.c_double),
                ("ms_transpose", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("perm_batches", ctypes.c_int64), ("ms_rank", ctypes.c_double),
                ("ms_es", ctypes.c_double), ("es_launches", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GseaError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("gsea error %d: %s" % (code, message))
        self.code = code


# exported symbols and their signatures: (restype, argtypes)
SIGNATURES = {
    "gsea_metric_id": (_I, [ctypes.c_char_p]),
    "gsea_metric_name": (ctypes.c_char_p, [_I]),
    "gsea_last_error": (ctypes.c_char_p, []),
    "gsea_version": (ctypes.c_char_p, []),
    "gsea_set_tuning": (_I, [_P]),
    "gsea_get_tuning": (_I, [_P]),
    "gsea_device_count": (_I, [ctypes.POINTER(_I)]),
    "gsea_get_device": (_I, [ctypes.POINTER(_I)]),
    "gsea_set_device": (_I, [_I]),
    "gsea_device_name": (_I, [_I, ctypes.c_char_p, ctypes.c_size_t,
                              ctypes.POINTER(ctypes.c_size_t)]),
    "gsea_run_f32": (_I, [_P, _I64, _I64, _P, _P, _P, _I64, _I64, _I, _I, _P,
                          _P, _P]),
    "gsea_run_f64": (_I, [_P, _I64, _I64, _P, _P, _P, _I64, _I64, _I, _I, _P,
                          _P, _P]),
    "gsea_run_device_f32": (_I, [_P, _I64, _I64, _P, _P, _P, _I64, _I64, _I,
                                 _I, _P, _P, _P, _P]),
    "gsea_run_device_f64": (_I, [_P, _I64, _I64, _P, _P, _P, _I64, _I64, _I,
                                 _I, _P, _P, _P, _P]),
    "gsea_sweep_device_f32": (_I, [_P, _I64, _I64, _P, _P, _P, _I64, _I64, _P, _I, _I, _I,
                                   _P, _P, _P]),
    "gsea_sweep_device_f64": (_I, [_P, _I64, _I64, _P, _P, _P, _I64, _I64, _P, _I, _I, _I,
                                   _P, _P, _P]),
    "gsea_run_budget_device_f32": (_I, [_P, _I64, _I64, _P, _P, _P, _I64, _I64, _I,
                                        _I, _P, _P, _P, _P, _P, _P, _P]),
    "gsea_run_budget_device_f64": (_I, [_P, _I64, _I64, _P, _P, _P, _I64, _I64, _I,
                                        _I, _P, _P, _P, _P, _P, _P, _P]),
    "gsea_metric_f32": (_I, [_P, _I64, _I64, _P, _I64, _I64, _I, _I, _P, _P]),
    "gsea_metric_f64": (_I, [_P, _I64, _I64, _P, _I64, _I64, _I, _I, _P, _P]),
    "gsea_metric_noalloc": (_I, [_I, _P, _I64, _I64, _P, _I64, _I64, _I, _I, _P, _P]),
    "gsea_rank_es_f32": (_I, [_P, _I64, _I64, _P, _P, _I64, _I, _P, _P]),
    "gsea_rank_es_f64": (_I, [_P, _I64, _I64, _P, _P, _I64, _I, _P, _P]),
    "gsea_argsort_f32": (_I, [_P, _I64, _I64, _I, _P, _P]),
    "gsea_argsort_f64": (_I, [_P, _I64, _I64, _I, _P, _P]),
    "gsea_map_sets": (_I, [_P, _I64, _P, _P, _I64, _P, _P, _I64,
                           ctypes.POINTER(_I64)]),
    "gsea_dump_es": (_I, [ctypes.c_char_p, _P, _I64, _I64, _I]),
    "gsea_mlp_num_params": (_I, []),
    "gsea_mlp_forward": (_I, [_P, _P, _I64, _I, _P, _P, _P]),
    "gsea_mlp_grad": (_I, [_P, _P, _P, _I64, _P, ctypes.POINTER(ctypes.c_float), _P]),
    "gsea_mlp_adam": (_I, [_P, _P, _P, _P, _I64, ctypes.c_float, ctypes.c_float, _P]),
    "gsea_set_features": (_I, [_P, _P, _I64, _P, _P, _I64, _P, _P, _P]),
    "gsea_null_moments": (_I, [_P, _I64, _I64, _P, _P]),
    "gsea_proxy_run": (_I, [_I, _P, _I64, _I64, _P, _P, _P, _P, _I64,
                            ctypes.c_char_p, _I64, _I, _I, ctypes.c_char_p, _P,
                            _P]),
}

_lib = None


def lib():
    """Load libgsea_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libgsea_hip.so is not built; run `python -m cudagsea_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != GSEA_OK:
        msg = lib().gsea_last_error()
        raise GseaError(rc, msg.decode() if msg else "")
    return rc
