"""C ABI checks that need no GPU: the library loads, exports what
include/gsea_hip.h declares, and the host-side logic (metric names, set
mapping, dump files, argument errors, GSEA() input checks) behaves."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "gsea_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gsea_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import cudagsea_amd
    L = cudagsea_amd.lib()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    # and the ctypes table covers them all
    from cudagsea_amd._lib import SIGNATURES
    assert set(names) <= set(SIGNATURES), set(names) - set(SIGNATURES)


def test_proxy_abi_symbols_exported():
    # C++ proxy ABI of cudaGSEA_wrapper.cuh, looked up by mangled name
    import subprocess
    lib = os.path.join(ROOT, "cudagsea_amd", "libgsea_hip.so")
    out = subprocess.run(["nm", "-DC", lib], stdout=subprocess.PIPE, text=True).stdout
    for sym in ("gsea_single_proxy(", "gsea_double_proxy(", "listCudaDevices_proxy",
                "getCudaDevice_proxy", "setCudaDevice_proxy",
                "loadLabelsFromCLS_proxy", "loadGeneSetsFromGMT_proxy",
                "loadExpressionDataFromGCT_proxy"):
        assert sym in out, sym


CPP = os.path.join(ROOT, "tests", "cpp")
PROXY_DEVICES = os.path.join(CPP, "_build", "proxy_devices_test")


def test_proxy_device_functions_cpp_without_a_device():
    # tests/cpp/proxy_devices_test.cpp links the proxy ABI from C++ and checks
    # listCudaDevices_proxy / getCudaDevice_proxy / setCudaDevice_proxy
    # against cudaGSEA_wrapper.cu:87-138; with no device visible the list is
    # empty and every id but 0 falls back to device 0 with the warning
    import subprocess
    subprocess.run(["make", "-C", CPP], check=True, stdout=subprocess.PIPE)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    r = subprocess.run([PROXY_DEVICES], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, env=env, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "ok 0 device(s), 0 failure(s)"


RPKG = os.path.join(ROOT, "cudagsea_amd", "rpkg", "cudaGSEA")
R_EXPORTS = {"loadLabelsFromCLS", "loadGeneSetsFromGMT", "loadExpressionDataFromGCT",
             "listCudaDevices", "getCudaDevice", "setCudaDevice", "GSEA"}


def test_r_package_entry_points_are_consistent():
    # R is not installed here: check the package's pieces against each other
    # and against the reference package's exports (cudaGSEA/NAMESPACE): every
    # export is an R function whose .Call target is an RcppExport of the glue
    # taking as many SEXP arguments as the R call passes
    ns = open(os.path.join(RPKG, "NAMESPACE")).read()
    exports = set(re.findall(r"\w+", re.search(r"export\(([^)]*)\)", ns).group(1)))
    assert exports == R_EXPORTS
    assert re.search(r"useDynLib\(cudaGSEA\)", ns)
    desc = open(os.path.join(RPKG, "DESCRIPTION")).read()
    assert re.search(r"^Package: cudaGSEA$", desc, re.M)
    assert re.search(r"^LinkingTo: Rcpp$", desc, re.M)
    rsrc = open(os.path.join(RPKG, "R", "cudaGSEA.R")).read()
    funcs = dict(re.findall(r"(\w+) <- function\(([^)]*)\)", rsrc))
    assert set(funcs) == R_EXPORTS
    calls = {m[0]: m[1] for m in re.findall(r'\.Call\("(\w+)"((?:,[^,)]*)*?),\s*PACKAGE = "cudaGSEA"\)',
                                             rsrc.replace("\n", " "))}
    assert set(calls) == R_EXPORTS
    glue = open(os.path.join(RPKG, "src", "gsea_rcpp.cpp")).read()
    for name, args in calls.items():
        m = re.search(r"RcppExport SEXP %s\(([^)]*)\)" % name, glue)
        assert m, name
        n_c = len(re.findall(r"\bSEXP\b", m.group(1)))
        n_r = len([a for a in args.split(",") if a.strip()])
        assert n_c == n_r, (name, n_c, n_r)
    # GSEA()'s arguments and defaults (cudaGSEA/R/GSEA.R)
    assert re.sub(r"\s+", " ", funcs["GSEA"]) == (
        "exprsData, labelList, geneSets, numPermutations, metricString = \"\", "
        "dumpFileName = \"\", checkInput = TRUE, doublePrecision = FALSE")
    mk = open(os.path.join(RPKG, "src", "Makevars")).read()
    assert "-lgsea_hip" in mk and "GSEA_AMD_HOME" in mk


def test_metric_names_roundtrip():
    import cudagsea_amd as cg
    L = cg.lib()
    for i, m in enumerate(cg.METRICS):
        assert L.gsea_metric_id(m.encode()) == i
        assert L.gsea_metric_name(i).decode() == m
    assert L.gsea_metric_id(b"") == -1
    assert L.gsea_metric_id(b"pearson") == -1
    assert L.gsea_metric_name(14) == b""


def test_map_sets_semantics():
    import cudagsea_amd as cg
    genes = ["TP53", "BRCA1", "EGFR", "TP53", "MYC"]
    sets = [["TP53", "NOPE"], ["EGFR", "EGFR", "MYC"], ["ABSENT"], ["MYC", "BRCA1"]]
    off, g = cg.map_sets(genes, sets)
    got = [g[off[i]:off[i + 1]].tolist() for i in range(len(sets))]
    # brute force of create_bitmaps.cuh:38-60
    want = [[k for k, sym in enumerate(genes) if sym in set(s)] for s in sets]
    assert got == want


def test_dump_naming_and_content(tmp_path):
    import cudagsea_amd as cg
    es = np.arange(12, dtype=np.float32).reshape(3, 4)
    prefix = str(tmp_path / "run")
    cg._lib.check(cg.lib().gsea_dump_es(prefix.encode(), es.ctypes.data_as(ctypes.c_void_p),
                                        3, 4, 32))
    path = prefix + "_3_4_32.es"
    assert os.path.getsize(path) == 48
    assert (cg.read_dump(path) == es).all()
    rc = cg.lib().gsea_dump_es(b"", es.ctypes.data_as(ctypes.c_void_p), 3, 4, 32)
    assert rc == 19                      # CUDA_GSEA_EMPTY_DUMP_FILE_NAME_ERROR


def test_argument_errors_before_any_device_call():
    import cudagsea_amd as cg
    X = np.ones((4, 10), dtype=np.float32)
    lab = np.array([0, 0, 1, 1], dtype=np.uint8)
    off = np.array([0, 2], dtype=np.int32)
    genes = np.array([1, 3], dtype=np.int32)
    with pytest.raises(cg.GseaError) as e:
        cg.run(X, lab, off, genes, 8, 99)
    assert e.value.code == 1             # no known metric
    with pytest.raises(cg.GseaError) as e:
        cg.run(X, lab, off, np.array([3, 1], dtype=np.int32), 8, 0)
    assert e.value.code == 30            # unsorted set
    with pytest.raises(cg.GseaError) as e:
        cg.run(X, lab, off, np.array([1, 10], dtype=np.int32), 8, 0)
    assert e.value.code == 30            # gene out of range
    with pytest.raises(cg.GseaError) as e:
        cg.run(np.ones((4, 65536), dtype=np.float32), lab, off, genes, 8, 0)
    assert e.value.code == 31            # above GSEA_MAX_GENES


def test_tuning_roundtrip_and_range_checks():
    # gsea_set_tuning / gsea_get_tuning need no device: the per-kind metric
    # tiles round-trip, and values outside the documented ranges are refused
    import cudagsea_amd as cg
    from cudagsea_amd import _lib
    assert ctypes.sizeof(_lib.gsea_tuning) == 72
    try:
        cg.set_tuning(metric_gene_tile=32, metric_gene_tile_kind=[64, 0, 16, 0, 32, 0],
                      metric_perms_per_block_kind=[1024, 0, 0, 256, 0, 0], es_waves=12,
                      es_pair_set_max=200)
        t = cg.get_tuning()
        assert t["metric_gene_tile"] == 32 and t["es_waves"] == 12
        assert t["es_pair_set_max"] == 200
        assert t["metric_gene_tile_kind"] == [64, 0, 16, 0, 32, 0]
        assert t["metric_perms_per_block_kind"] == [1024, 0, 0, 256, 0, 0]
        for bad in ({"metric_gene_tile_kind": [0, 0, 24, 0, 0, 0]},
                    {"metric_perms_per_block_kind": [0, -1, 0, 0, 0, 0]},
                    {"metric_gene_tile": 48}, {"es_waves": 17}, {"es_lane_set_max": 100},
                    {"es_pair_set_max": 128}, {"es_pair_set_max": 257},
                    {"es_pair_set_max": -2}):
            with pytest.raises(cg.GseaError) as e:
                cg.set_tuning(**bad)
            assert e.value.code == 30, bad
        with pytest.raises(ValueError):
            cg.set_tuning(metric_gene_tile_kind=[16, 16])
    finally:
        cg.set_tuning()
    assert cg.get_tuning()["metric_gene_tile_kind"] == [0] * 6


def _frame():
    import pandas as pd
    rng = np.random.default_rng(0)
    return pd.DataFrame(rng.random((6, 4)) + 1, index=["G%d" % i for i in range(6)],
                        columns=["s%d" % i for i in range(4)])


def test_gsea_input_checks_return_false(capsys):
    import cudagsea_amd as cg
    df = _frame()
    sets = {"a": ["G1", "G2"]}
    assert cg.GSEA(df, [0, 1, 0], sets, 10, "onepass_t_test") is False
    assert cg.GSEA(df, [0, 1, 2, 1], sets, 10, "onepass_t_test") is False
    bad = df.copy()
    bad.iloc[2, 2] = np.nan
    assert cg.GSEA(bad, [0, 1, 0, 1], sets, 10, "onepass_t_test") is False
    assert cg.GSEA(df, [0, 1, 0, 1], {}, 10, "onepass_t_test") is False
    assert cg.GSEA(df, [0, 1, 0, 1], {"a": []}, 10, "onepass_t_test") is False
    assert capsys.readouterr().out.count("INVALID DATA") == 5


def test_gsea_unknown_metric_raises_instead_of_exit():
    import cudagsea_amd as cg
    df = _frame()
    with pytest.raises(cg.GseaError) as e:
        cg.GSEA(df, [0, 1, 0, 1], {"a": ["G1"]}, 10, "")
    assert e.value.code == 1


def test_autotuner_es_search_space():
    # every enrichment-stage candidate is a valid setting, each distinct run
    # appears once: the pair class only with the lane split at 128
    from cudagsea_amd import autotune
    import cudagsea_amd as cg
    cands = autotune.es_candidates()
    keys = {tuple(sorted(c.items())) for c in cands}
    assert len(keys) == len(cands)
    assert all(c["es_pair_set_max"] == -1 for c in cands if c["es_lane_set_max"] != 128)
    assert {c["es_pair_set_max"] for c in cands if c["es_lane_set_max"] == 128} == \
        set(autotune.ES_SPACE["es_pair_set_max"])
    try:
        for c in cands:
            cg.set_tuning(**c)
    finally:
        cg.set_tuning()
