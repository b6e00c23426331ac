"""The graph-captured sweep (gsea_sweep_device_*) against per-metric
gsea_run_device_* calls, and the statistics' device threshold sort against
its host fallback, through the C ABI (MI355X)."""
import numpy as np
import pytest

import oracle
import parity
from parity import bits_equal
from cudagsea_amd import workloads

pytestmark = pytest.mark.gpu

METRICS4 = ["stable_signal2noise", "onepass_t_test", "naive_log2_ratio_of_classes",
            "overkill_t_test"]


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _per_metric(gpu, X, w, P, metrics, tdt, signed=False):
    import torch
    out = []
    for m in metrics:
        st = torch.empty((5, w["G"]), dtype=tdt, device="cuda")
        gpu.run_device(X, w["labels"], dev(w["set_offsets"]), dev(w["set_genes"]), P, m, st,
                       signed_nes=signed)
        out.append(st.cpu().numpy())
    return np.stack(out)


@pytest.mark.parametrize("dbl", [False, True])
def test_sweep_graph_replays_match_run_device(gpu, dbl):
    import torch
    print("sweep graph", "fp64" if dbl else "fp32", flush=True)
    dt = np.float64 if dbl else np.float32
    tdt = torch.float64 if dbl else torch.float32
    w = workloads.make("c3", num_genes=1500, num_sets=120, num_perms=200, dtype=dt)
    P, G = 200, w["G"]
    X = dev(w["X"])
    off, genes = dev(w["set_offsets"]), dev(w["set_genes"])
    want = _per_metric(gpu, X, w, P, METRICS4, tdt)
    st = torch.full((len(METRICS4), 5, G), 3.0, dtype=tdt, device="cuda")
    for rep in range(4):
        t = gpu.sweep_device(X, w["labels"], off, genes, P, METRICS4, st, graph=True)
        assert t["ms_total"] > 0
        assert bits_equal(st.cpu().numpy(), want).all(), rep
        # the program's own eager work between replays
        junk = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        st.fill_(float(junk) * 0 + 5.0)
    # the same sequence queued without a graph
    gpu.sweep_device(X, w["labels"], off, genes, P, METRICS4, st, graph=False)
    assert bits_equal(st.cpu().numpy(), want).all()


def test_sweep_graph_follows_inputs_and_recaptures(gpu):
    import torch
    w = workloads.make("c3", num_genes=1200, num_sets=90, num_perms=128)
    P, G = 128, w["G"]
    X = dev(w["X"])
    off, genes = dev(w["set_offsets"]), dev(w["set_genes"])
    st = torch.empty((2, 5, G), dtype=torch.float32, device="cuda")
    gpu.sweep_device(X, w["labels"], off, genes, P, METRICS4[:2], st)
    # new expression values at the same address: the replay reads them
    X.mul_(1.5).add_(0.25)
    gpu.sweep_device(X, w["labels"], off, genes, P, METRICS4[:2], st)
    w2 = dict(w)
    w2["X"] = X.cpu().numpy()
    want = _per_metric(gpu, X, w2, P, METRICS4[:2], torch.float32)
    assert bits_equal(st.cpu().numpy(), want).all()
    # new set contents at the same address: the set tiers are rebuilt
    g2 = w["set_genes"].copy()
    for s in range(0, G, 3):
        a, b = w["set_offsets"][s], w["set_offsets"][s + 1]
        g2[a:b] = np.sort(np.random.default_rng(s).choice(w["N"], b - a, replace=False))
    genes.copy_(torch.from_numpy(g2))
    gpu.sweep_device(X, w["labels"], off, genes, P, METRICS4[:2], st)
    w3 = dict(w2)
    w3["set_genes"] = g2
    want3 = _per_metric(gpu, X, w3, P, METRICS4[:2], torch.float32)
    assert bits_equal(st.cpu().numpy(), want3).all()
    # another permutation count: captured again
    st4 = torch.empty((2, 5, G), dtype=torch.float32, device="cuda")
    gpu.sweep_device(X, w["labels"], off, genes, 96, METRICS4[:2], st4)
    want4 = _per_metric(gpu, X, w3, 96, METRICS4[:2], torch.float32)
    assert bits_equal(st4.cpu().numpy(), want4).all()


def test_sweep_graph_signed_nes(gpu):
    # the flag is part of the captured sequence: a signed sweep replays the
    # signed statistics, and switching the flag back recaptures
    import torch
    w = workloads.make("c3", num_genes=1200, num_sets=90, num_perms=160)
    P, G = 160, w["G"]
    X = dev(w["X"])
    off, genes = dev(w["set_offsets"]), dev(w["set_genes"])
    want_s = _per_metric(gpu, X, w, P, METRICS4[:2], torch.float32, signed=True)
    want_r = _per_metric(gpu, X, w, P, METRICS4[:2], torch.float32)
    st = torch.empty((2, 5, G), dtype=torch.float32, device="cuda")
    for rep in range(2):
        gpu.sweep_device(X, w["labels"], off, genes, P, METRICS4[:2], st, signed_nes=True)
        assert bits_equal(st.cpu().numpy(), want_s).all(), rep
    gpu.sweep_device(X, w["labels"], off, genes, P, METRICS4[:2], st)
    assert bits_equal(st.cpu().numpy(), want_r).all()
    assert not np.allclose(want_s[:, 1], want_r[:, 1])


def test_statistics_host_threshold_fallback_matches_oracle(gpu):
    # 40,000 sets: more observed NES than one workgroup sorts in LDS, so the
    # thresholds are sorted on the host; the statistics must still meet the
    # oracle's bars
    print("40,000 sets (host threshold sort)", flush=True)
    rng = np.random.default_rng(17)
    N, P = 600, 16
    w = workloads.make("c1", num_genes=N, num_perms=P, num_sets=2)
    sets = [np.sort(rng.choice(N, int(rng.integers(1, 12)), replace=False)).astype(np.int32)
            for _ in range(40000)]
    off = np.zeros(len(sets) + 1, np.int32)
    off[1:] = np.cumsum([len(s) for s in sets])
    genes = np.concatenate(sets)
    want, es_want, hi, lo = oracle.gsea_ext(w["X"], w["labels"], off, genes, P,
                                            "stable_signal2noise")
    got = gpu.run(w["X"], w["labels"], off, genes, P, "stable_signal2noise")
    parity.assert_stats(got, want, es_want, hi, lo)


def test_sweep_graph_stage_times_agree_with_events(gpu):
    # the graph's stage times come from device wall-clock stamps around each
    # launch; eager runs time the same stages with HIP events
    import torch
    w = workloads.make("c3", num_genes=4000, num_sets=600, num_perms=1024)
    P, G = 1024, w["G"]
    X = dev(w["X"])
    off, genes = dev(w["set_offsets"]), dev(w["set_genes"])
    st = torch.empty((2, 5, G), dtype=torch.float32, device="cuda")
    gpu.sweep_device(X, w["labels"], off, genes, P, METRICS4[:2], st, graph=True)
    tg = gpu.sweep_device(X, w["labels"], off, genes, P, METRICS4[:2], st, graph=True)
    te = gpu.sweep_device(X, w["labels"], off, genes, P, METRICS4[:2], st, graph=False)
    assert tg["es_launches"] == te["es_launches"] == 2
    for k in ("ms_metric", "ms_es", "ms_stats"):
        assert tg[k] > 0 and te[k] > 0, k
        assert abs(tg[k] - te[k]) <= 0.25 * te[k] + 0.05, (k, tg[k], te[k])
    assert tg["ms_total"] >= tg["ms_metric"] + tg["ms_es"]


def test_autotuner_short_search(gpu, tmp_path):
    # a short search on a small ragged workload: the knobs it returns are
    # from the search space, the tuned run is never slower than the defaults
    # as the tuner measured them, the cache entry is written, and the tuned
    # shapes give the default shapes' statistics bit for bit
    import torch
    from cudagsea_amd import autotune
    w = workloads.make("c3", num_genes=4000, num_samples=100, num_sets=600, num_perms=512)
    w["metrics"] = ["naive_diff_of_classes", "stable_t_test"]
    path = str(tmp_path / "cache.json")
    key, res = autotune.tune_workload(w, sample_perms=256, path=path)
    assert res["tuned_ms"] <= res["default_ms"]
    k = res["knobs"]
    if k:
        assert k["es_lane_set_max"] in autotune.ES_SPACE["es_lane_set_max"]
        assert k["es_pair_set_max"] in autotune.ES_SPACE["es_pair_set_max"]
        assert k["es_waves"] in autotune.ES_SPACE["es_waves"]
    assert autotune.load_cache(path)[key]["knobs"] == k
    X = dev(w["X"])
    st = []
    for knobs in ({}, k):
        gpu.set_tuning(**knobs)
        try:
            s = torch.empty((2, 5, w["G"]), dtype=torch.float32, device="cuda")
            gpu.sweep_device(X, w["labels"], dev(w["set_offsets"]), dev(w["set_genes"]),
                             512, w["metrics"], s)
            st.append(s.cpu().numpy())
        finally:
            gpu.set_tuning()
    assert bits_equal(st[0], st[1]).all()
