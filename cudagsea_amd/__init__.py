"""cudagsea_amd -- permutation GSEA on AMD Instinct MI355X (gfx950).

Hand-written HIP kernels behind a C ABI (include/gsea_hip.h), reached through
`GSEA(...)`, the mirror of the reference's R entry point.
"""
from ._lib import GseaError, lib  # noqa: F401
from .gsea import (GSEA, METRICS, STAT_ROWS, argsort_device,  # noqa: F401
                   device_info, map_sets,
                   metric_device, metric_id, rank_es_device, read_dump, run,
                   run_budget_device, run_device, set_tuning, get_tuning,
                   sweep_device)

__version__ = "0.1.0"
