"""HIP path vs the CPU oracle on the same seeded inputs (MI355X).

Bars: metrics and ranks bit-exact (integer/index work and the reference's
exact float arithmetic); ES and NES within 1e-5 relative (|d| <= 1e-5*|x| +
1e-6) in fp32 and fp64; NP and FWER are counts over permutations and must
agree exactly except where an ES comparison is within rounding of a tie
(each count difference needs an oracle comparison within the ES bar of a
tie); FDR within 1e-5 relative where NES agree.
"""
import numpy as np
import pytest

import oracle
from cudagsea_amd import workloads

pytestmark = pytest.mark.gpu

ES_RTOL = 1e-5
ES_ATOL = 1e-6


def bits_equal(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    u = np.uint64 if a.dtype == np.float64 else np.uint32
    return (a.view(u) == b.view(u)) | (np.isnan(a) & np.isnan(b))


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def es_close(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    ok = np.abs(a - b) <= ES_RTOL * np.abs(b) + ES_ATOL
    return ok | (np.isnan(a) & np.isnan(b))


def small(config, **kw):
    return workloads.make(config, **kw)


# ------------------------------------------------------------- metrics ----
@pytest.mark.parametrize("dbl", [False, True])
@pytest.mark.parametrize("config,N,S", [("c1", 2000, 40), ("c3", 700, 500),
                                        ("c4", 900, 300)])
def test_metric_bit_exact_all_metrics(gpu, config, N, S, dbl):
    import torch
    print("metric parity", config, N, S, "fp64" if dbl else "fp32", flush=True)
    w = small(config, num_genes=N, num_samples=S, num_sets=4, num_perms=24)
    dt = np.float64 if dbl else np.float32
    X = w["X"].astype(dt)
    Xd = dev(X)
    P = 24
    for m in oracle.METRICS:
        want = oracle.correlate(X, w["labels"], P, m, shift=5, sample_major=True,
                                double=dbl)
        r = torch.empty((P, N), dtype=torch.float64 if dbl else torch.float32,
                        device="cuda")
        gpu.metric_device(Xd, w["labels"], P, m, r, shift=5)
        got = r.cpu().numpy()
        eq = bits_equal(got, want)
        assert eq.all(), (m, int((~eq).sum()), np.argwhere(~eq)[:3])


def test_metric_log_bitwise_over_wide_range(gpu):
    # drive the log2-ratio metric through ratios spanning many binades,
    # subnormal and special results included
    import torch
    rng = np.random.default_rng(1)
    N, S = 4096, 2
    a = np.exp(rng.uniform(-80, 80, N)).astype(np.float32)
    b = np.exp(rng.uniform(-80, 80, N)).astype(np.float32)
    a[:8] = [0, 1, 1e-45, 3e38, 1, 5, 0, 2]
    b[:8] = [1, 0, 1, 1e-30, 1, 5, 0, 2]
    X = np.stack([a, b])                              # sample 0 class A, 1 class B
    lab = np.array([0, 1], dtype=np.uint8)
    for dbl in (False, True):
        Xt = X.astype(np.float64 if dbl else np.float32)
        want = oracle.correlate(Xt, lab, 1, "naive_log2_ratio_of_classes",
                                sample_major=True, double=dbl)
        r = torch.empty((1, N), dtype=torch.float64 if dbl else torch.float32,
                        device="cuda")
        gpu.metric_device(dev(Xt), lab, 1, "naive_log2_ratio_of_classes", r)
        assert bits_equal(r.cpu().numpy(), want).all()


def test_metric_nonfinite_inputs_naive(gpu):
    import torch
    X = np.ones((6, 64), dtype=np.float32)
    X[1, 3] = np.inf
    X[4, 9] = -np.inf
    X[2, 20] = np.nan
    lab = np.array([0, 0, 0, 1, 1, 1], dtype=np.uint8)
    for m in oracle.METRICS:
        want = oracle.correlate(X, lab, 6, m, sample_major=True)
        r = torch.empty((6, 64), dtype=torch.float32, device="cuda")
        gpu.metric_device(dev(X), lab, 6, m, r)
        assert bits_equal(r.cpu().numpy(), want).all(), m


# --------------------------------------------------------------- ranks ----
@pytest.mark.parametrize("dbl", [False, True])
def test_argsort_bit_exact_with_ties(gpu, dbl):
    import torch
    rng = np.random.default_rng(8)
    P, N = 12, 20000
    dt = np.float64 if dbl else np.float32
    keys = rng.standard_normal((P, N)).astype(dt)
    keys[0] = rng.integers(-2, 3, N)                    # massive ties
    keys[1, :300] = 0.0
    keys[1, 300:600] = -0.0
    keys[2, ::97] = np.nan
    keys[3, ::50] = np.inf
    keys[4] = 1.0                                       # all equal
    for asc in (False, True):
        _, want = oracle.pathway_sort(keys, ascending=asc)
        idx = torch.empty((P, N), dtype=torch.int32, device="cuda")
        gpu.argsort_device(dev(keys), idx, ascending=asc)
        got = idx.cpu().numpy().view(np.uint32)
        assert (got == want).all(), (asc, np.argwhere(got != want)[:3])


@pytest.mark.parametrize("N", [1, 2, 63, 64, 1000, 1025, 4096, 32768])
def test_argsort_sizes(gpu, N):
    import torch
    rng = np.random.default_rng(N)
    keys = rng.standard_normal((3, N)).astype(np.float32)
    _, want = oracle.pathway_sort(keys)
    idx = torch.empty((3, N), dtype=torch.int32, device="cuda")
    gpu.argsort_device(dev(keys), idx)
    assert (idx.cpu().numpy().view(np.uint32) == want).all()


# ------------------------------------------------------------ ES stage ----
def _es_stage_case(gpu, w, P, metric, dbl=False, asc=False):
    import torch
    dt = np.float64 if dbl else np.float32
    r = oracle.correlate(w["X"].astype(dt), w["labels"], P, metric,
                         sample_major=True, double=dbl)
    sk, idx = oracle.pathway_sort(r, ascending=asc)
    want = oracle.scores(sk, idx, w["set_offsets"], w["set_genes"])      # [G][P]
    es = torch.empty((P, w["G"]), dtype=torch.float64 if dbl else torch.float32,
                     device="cuda")
    genes = w["set_genes"] if w["set_genes"].size else np.zeros(1, np.int32)
    gpu.rank_es_device(dev(r), dev(w["set_offsets"]), dev(genes), es, ascending=asc)
    got = es.cpu().numpy().T
    return got, want


@pytest.mark.parametrize("dbl", [False, True])
@pytest.mark.parametrize("config", ["c1", "c3", "c4"])
def test_rank_es_matches_oracle(gpu, config, dbl):
    print("rank_es parity", config, dbl, flush=True)
    kw = dict(num_perms=48)
    if config != "c1":
        kw.update(num_genes=6000, num_sets=300)
    w = small(config, **kw)
    metric = w["metrics"][-1]
    got, want = _es_stage_case(gpu, w, 48, metric, dbl)
    ok = es_close(got, want)
    assert ok.mean() > 0.9999 and (~ok).sum() <= 2, (int((~ok).sum()),)
    assert (np.abs(got) <= 1 + 1e-6).all()


def test_rank_es_edge_sets(gpu):
    # empty set, single genes, every gene, first/last rank genes, duplicates
    # across sets; ascending order too
    rng = np.random.default_rng(3)
    N = 3000
    w = small("c1", num_genes=N, num_perms=16, num_sets=2)
    sets = [np.array([], np.int32), np.array([0], np.int32),
            np.arange(N, dtype=np.int32), np.array([N - 1], np.int32),
            np.sort(rng.choice(N, 2999, replace=False)).astype(np.int32),
            np.sort(rng.choice(N, 5, replace=False)).astype(np.int32)]
    # both sides of every size-class boundary of the set work split
    sets += [np.sort(rng.choice(N, k, replace=False)).astype(np.int32)
             for k in (8, 9, 16, 17, 31, 32, 33, 63, 64, 65, 128, 129)]
    sets += [np.arange(N - 64, N, dtype=np.int32), np.arange(64, dtype=np.int32)]
    off = np.zeros(len(sets) + 1, np.int32)
    off[1:] = np.cumsum([len(s) for s in sets])
    w["set_offsets"], w["set_genes"], w["G"] = off, np.concatenate(sets), len(sets)
    for asc in (False, True):
        got, want = _es_stage_case(gpu, w, 16, "stable_signal2noise", asc=asc)
        assert es_close(got, want).all(), (asc, got[:, :3], want[:, :3])


# ----------------------------------------------------------- whole path ----
def _tie_counts(es_want, P):
    """Per set, from the oracle's ES [G][P]: the permutations whose |ES|
    lies within the ES bar of the observed |ES| (NP compares those two), and
    those whose extreme normalised score over all sets lies within it of the
    set's NES (FWER compares those).  Where two values are only pinned to
    within the bar, their comparison may go either way."""
    e = es_want.astype(np.float64)
    a = np.abs(e)
    np_ties = (np.abs(a - a[:, :1]) <= ES_RTOL * a[:, :1] + ES_ATOL).sum(axis=1) - 1
    z = e * (P / a.sum(axis=1))[:, None]
    nes = z[:, 0]
    ext = np.where(nes[:, None] > 0, z.max(axis=0)[None, :], z.min(axis=0)[None, :])
    fw_ties = (np.abs(ext - nes[:, None]) <= ES_RTOL * np.abs(nes)[:, None] + ES_ATOL).sum(axis=1)
    return np_ties, fw_ties


def _compare_stats(got, want, P, es_want=None):
    ok_es = es_close(got[0], want[0])
    ok_nes = es_close(got[1], want[1])
    assert ok_es.all() and ok_nes.all(), (np.argwhere(~ok_es)[:3], np.argwhere(~ok_nes)[:3])
    # counts: exact except for comparisons that sit within the ES bar of a tie
    d_np = np.abs(got[2].astype(np.float64) - want[2]) * P
    d_fw = np.abs(got[3].astype(np.float64) - want[3]) * P
    if es_want is not None:
        np_ties, fw_ties = _tie_counts(es_want, P)
        assert (d_np <= np_ties + 0.5).all(), np.argwhere(d_np > np_ties + 0.5)[:5]
        assert (d_fw <= fw_ties + 0.5).all(), np.argwhere(d_fw > fw_ties + 0.5)[:5]
    else:
        assert (d_np <= 2.5).all() and (d_np > 0.5).mean() <= 0.01
        assert (d_fw <= 2.5).all() and (d_fw > 0.5).mean() <= 0.01
    fdr_ok = np.abs(got[4].astype(np.float64) - want[4]) <= 1e-3 + 1e-5 * np.abs(want[4])
    assert (fdr_ok | (np.isnan(got[4]) & np.isnan(want[4]))).mean() >= 0.99


@pytest.mark.parametrize("dbl", [False, True])
def test_full_path_c1(gpu, dbl):
    print("full path c1", dbl, flush=True)
    w = small("c1")
    P = w["P"]
    want, es_want = oracle.gsea(w["X"], w["labels"], w["set_offsets"], w["set_genes"],
                                P, "stable_signal2noise", double=dbl)
    got, es_got = gpu.run(w["X"], w["labels"], w["set_offsets"], w["set_genes"], P,
                          "stable_signal2noise", double=dbl, want_es=True)
    assert es_close(es_got, es_want).mean() > 0.9999
    _compare_stats(got, want, P, es_want)


@pytest.mark.parametrize("metric", oracle.METRICS)
def test_full_path_every_metric_small(gpu, metric):
    w = small("c3", num_genes=1500, num_sets=120, num_perms=200)
    want, es_want = oracle.gsea(w["X"], w["labels"], w["set_offsets"], w["set_genes"],
                                200, metric)
    got = gpu.run(w["X"], w["labels"], w["set_offsets"], w["set_genes"], 200, metric)
    _compare_stats(got, want, 200, es_want)


def test_full_path_flags(gpu):
    w = small("c1", num_perms=128)
    for asc in (False, True):
        for swap in (False, True):
            want, es_want = oracle.gsea(w["X"], w["labels"], w["set_offsets"],
                                        w["set_genes"], 128, "onepass_signal2noise",
                                        ascending=asc, swap=swap)
            got = gpu.run(w["X"], w["labels"], w["set_offsets"], w["set_genes"], 128,
                          "onepass_signal2noise", ascending=asc, swap_labels=swap)
            _compare_stats(got, want, 128, es_want)


def test_gsea_r_mirror_and_dump(gpu, tmp_path):
    import pandas as pd
    w = small("c1", num_perms=64)
    df = pd.DataFrame(w["X"].T.astype(np.float64), index=w["gene_symbols"],
                      columns=["s%d" % i for i in range(w["S"])])
    sets = {}
    for s in range(w["G"]):
        g = w["set_genes"][w["set_offsets"][s]:w["set_offsets"][s + 1]]
        sets[w["set_names"][s]] = [w["gene_symbols"][i] for i in g] + ["NOT_A_GENE"]
    prefix = str(tmp_path / "c1")
    res = gpu.GSEA(df, w["labels"].tolist(), sets, 64, "stable_signal2noise", prefix)
    assert list(res.columns) == ["ES", "NES", "NP", "FWER", "FDR"]
    assert list(res.index) == w["set_names"]
    want, es_want = oracle.gsea(w["X"], w["labels"], w["set_offsets"], w["set_genes"],
                                64, "stable_signal2noise")
    got = res.to_numpy().T.astype(np.float32)
    _compare_stats(got, want, 64, es_want)
    dump = gpu.read_dump(prefix + "_%d_64_32.es" % w["G"])
    assert es_close(dump, es_want).mean() > 0.9999


# ------------------------------------------------ size-independent checks ----
def test_c2_properties_full_size(gpu):
    import torch
    print("c2 full size properties", flush=True)
    w = small("c2")
    P, G = w["P"], w["G"]
    X = dev(w["X"])
    off, genes = dev(w["set_offsets"]), dev(w["set_genes"])
    st = torch.empty((5, G), dtype=torch.float32, device="cuda")
    es = torch.empty((G, P), dtype=torch.float32, device="cuda")
    gpu.run_device(X, w["labels"], off, genes, P, "onepass_t_test", st, es)
    st1, es1 = st.cpu().numpy(), es.cpu().numpy()
    assert (np.abs(es1) <= 1 + 1e-6).all()
    assert (st1[2] >= 1.0 / P).all() and (st1[2] <= 1).all()
    assert ((st1[3] >= 0) & (st1[3] <= 1)).all()
    assert ((st1[4] >= 0) & (st1[4] <= 1)).all()
    # determinism: a second run is bit-identical
    gpu.run_device(X, w["labels"], off, genes, P, "onepass_t_test", st, es)
    assert bits_equal(st.cpu().numpy(), st1).all()
    assert bits_equal(es.cpu().numpy(), es1).all()
    # seed addressability: permutations 1000..1063 recomputed alone match
    r = torch.empty((64, w["N"]), dtype=torch.float32, device="cuda")
    gpu.metric_device(X, w["labels"], 64, "onepass_t_test", r, shift=1000)
    es_pg = torch.empty((64, G), dtype=torch.float32, device="cuda")
    gpu.rank_es_device(r, off, genes, es_pg)
    assert bits_equal(es_pg.cpu().numpy().T, es1[:, 1000:1064]).all()
    # the observed column equals the oracle's (metric and ranks bit-exact)
    want_r = oracle.correlate(w["X"], w["labels"], 1, "onepass_t_test", sample_major=True)
    sk, idx = oracle.pathway_sort(want_r)
    want_es = oracle.scores(sk, idx, w["set_offsets"], w["set_genes"])[:, 0]
    assert es_close(es1[:, 0], want_es).all()
