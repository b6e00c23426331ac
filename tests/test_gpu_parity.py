"""HIP path vs the CPU oracle on the same seeded inputs (MI355X), through the
C ABI.

Bars (tests/parity.py): metrics and ranks bit-exact (integer/index work and
the reference's exact float arithmetic); ES and NES within 1e-5 relative
(|d| <= 1e-5 |x| + 1e-6) in fp32 and fp64, element by element, with a sign
flip allowed only where the oracle's walk has two candidates of magnitudes
within that bar; NP and FWER exact up to comparisons within the bar of a
tie; FDR between the bounds those comparisons allow.
"""
import numpy as np
import pytest

import oracle
import parity
from parity import bits_equal
from cudagsea_amd import workloads

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def small(config, **kw):
    return workloads.make(config, **kw)


def csr(sets):
    off = np.zeros(len(sets) + 1, np.int32)
    off[1:] = np.cumsum([len(s) for s in sets])
    genes = np.concatenate(sets).astype(np.int32) if sets else np.zeros(0, np.int32)
    return off, genes


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


@pytest.mark.parametrize("dbl", [False, True])
def test_metric_bit_exact_many_samples(gpu, dbl):
    # 5,000 samples: no [S][16] tile fits the 64 KiB the kernel stages, so X
    # is read from global memory (one gene per lane); all 14 metrics
    import torch
    print("metric parity, 5000 samples", "fp64" if dbl else "fp32", flush=True)
    w = small("c2", num_genes=300, num_samples=5000, num_sets=4, num_perms=8)
    dt = np.float64 if dbl else np.float32
    X = w["X"].astype(dt)
    Xd = dev(X)
    for m in oracle.METRICS:
        want = oracle.correlate(X, w["labels"], 8, m, shift=2, sample_major=True, double=dbl)
        r = torch.empty((8, 300), dtype=torch.float64 if dbl else torch.float32,
                        device="cuda")
        gpu.metric_device(Xd, w["labels"], 8, m, r, shift=2)
        eq = bits_equal(r.cpu().numpy(), want)
        assert eq.all(), (m, int((~eq).sum()), np.argwhere(~eq)[:3])


def test_metric_tiles_per_kind_do_not_change_results(gpu):
    # every launch shape computes the same arithmetic: per-kind tiles and
    # permutation blocks give the default shapes' bits
    import torch
    w = small("c3", num_genes=3000, num_samples=120, num_perms=300)
    X = dev(w["X"])
    P, N = 300, w["N"]
    metrics = ["naive_log2_ratio_of_classes", "stable_t_test", "onepass_signal2noise",
               "twopass_t_test", "stable_signal2noise", "overkill_t_test"]
    want = {}
    for m in metrics:
        r = torch.empty((P, N), dtype=torch.float32, device="cuda")
        gpu.metric_device(X, w["labels"], P, m, r)
        want[m] = r.cpu().numpy()
    try:
        gpu.set_tuning(metric_gene_tile_kind=[64, 32, 16, 64, 32, 16],
                       metric_perms_per_block_kind=[1024, 96, 256, 32, 512, 2048])
        for m in metrics:
            r = torch.empty((P, N), dtype=torch.float32, device="cuda")
            gpu.metric_device(X, w["labels"], P, m, r)
            assert bits_equal(r.cpu().numpy(), want[m]).all(), m
    finally:
        gpu.set_tuning()


def test_metric_welford_fp32_wide_inputs(gpu):
    # the Welford metrics divide by the running count at every sample: many
    # samples (counts up to 3,000), values over 2^-20..2^40 with zeros,
    # cancelling pairs, near-constant and constant genes, and single values at
    # 2^-30, 2^50 and inf; bit-exact with the oracle
    import torch
    rng = np.random.default_rng(21)
    N, S, P = 640, 3000, 16
    sgn = rng.choice([-1.0, 1.0], (S, N))
    X = sgn * np.exp2(rng.uniform(-20, 40, (S, N)))
    X[:, 0:64] = 1e4 + rng.standard_normal((S, 64)) * 1e-2       # near-constant
    X[:, 64:128] = np.where(rng.random((S, 64)) < 0.7, 0.0, X[:, 64:128])
    X[::2, 128:192] = 3.0
    X[1::2, 128:192] = -3.0                                      # cancelling
    X[:, 192:224] = 2.0 ** -20
    X[:, 224:256] = -(2.0 ** 40)
    X[:, 256:272] = 7.25                                         # constant
    X[5, 300] = 2.0 ** -30                                       # out of range
    X[7, 340] = 2.0 ** 50
    X[9, 380] = np.inf
    X = X.astype(np.float32)
    lab = (rng.random(S) < 0.4).astype(np.uint8)
    Xd = dev(X)
    for m in ("stable_signal2noise", "stable_t_test"):
        want = oracle.correlate(X, lab, P, m, shift=3, sample_major=True)
        r = torch.empty((P, N), dtype=torch.float32, device="cuda")
        gpu.metric_device(Xd, lab, P, m, r, shift=3)
        eq = bits_equal(r.cpu().numpy(), want)
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


@pytest.mark.parametrize("N", [1, 2, 63, 64, 1000, 1025, 4096, 32767, 32768])
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
    """GPU rank+ES stage on the oracle's metric rows; returns the GPU ES and
    the oracle's (es, hi, lo), all [G][P]."""
    import torch
    dt = np.float64 if dbl else np.float32
    r = oracle.correlate(w["X"].astype(dt), w["labels"], P, metric,
                         sample_major=True, double=dbl)
    sk, idx = oracle.pathway_sort(r, ascending=asc)
    want, hi, lo = oracle.scores_ext(sk, idx, w["set_offsets"], w["set_genes"])
    es = torch.empty((P, w["G"]), dtype=torch.float64 if dbl else torch.float32,
                     device="cuda")
    genes = w["set_genes"] if w["set_genes"].size else np.zeros(1, np.int32)
    gpu.rank_es_device(dev(r), dev(w["set_offsets"]), dev(genes), es, ascending=asc)
    return es.cpu().numpy().T, want, hi, lo


@pytest.mark.parametrize("dbl", [False, True])
@pytest.mark.parametrize("config", ["c1", "c3", "c4"])
def test_rank_es_matches_oracle(gpu, config, dbl):
    print("rank_es parity", config, dbl, flush=True)
    kw = dict(num_perms=48)
    if config != "c1":
        kw.update(num_genes=6000, num_sets=300)
    w = small(config, **kw)
    got, want, hi, lo = _es_stage_case(gpu, w, 48, w["metrics"][-1], dbl)
    parity.assert_es(got, want, hi, lo)
    assert (np.abs(got) <= 1 + 1e-6).all()


@pytest.mark.parametrize("config,P", [("c3", 16), ("c4", 32)])
def test_rank_es_genome_scale_all_sets(gpu, config, P):
    # the configurations' own gene and set counts (20,000 genes; 10,000 or
    # 2,000 sets), a permutation subset
    print("rank_es genome scale", config, P, flush=True)
    w = small(config, num_perms=P)
    got, want, hi, lo = _es_stage_case(gpu, w, P, w["metrics"][-1])
    flips = parity.assert_es(got, want, hi, lo)
    print("  %d x %d scores, %d justified sign flips" % (w["G"], P, int(flips.sum())))


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
    w["set_offsets"], w["set_genes"] = csr(sets)
    w["G"] = len(sets)
    for asc in (False, True):
        got, want, hi, lo = _es_stage_case(gpu, w, 16, "stable_signal2noise", asc=asc)
        parity.assert_es(got, want, hi, lo)


def _es_rows_case(gpu, r, sets, asc=False):
    """GPU rank+ES stage on given metric rows r [P][N]; returns the GPU ES and
    the oracle's (es, hi, lo), all [G][P]."""
    import torch
    off, genes = csr(sets)
    sk, idx = oracle.pathway_sort(r, ascending=asc)
    want, hi, lo = oracle.scores_ext(sk, idx, off, genes)
    es = torch.empty((r.shape[0], len(sets)), dtype=torch.from_numpy(r).dtype,
                     device="cuda")
    gpu.rank_es_device(dev(r), dev(off), dev(genes), es, ascending=asc)
    return es.cpu().numpy().T, want, hi, lo


@pytest.mark.parametrize("dbl", [False, True])
def test_rank_es_pair_sets(gpu, dbl):
    # sets of 129..256 genes are scored by a pair of lanes (each sorts 128
    # ranks, one cross-lane compare-exchange splits them): both ends of the
    # class and its neighbours, sets holding the first and the last genes,
    # both orders; then the same sets with the pair class switched off
    print("pair sets", "fp64" if dbl else "fp32", flush=True)
    rng = np.random.default_rng(11)
    N, P = 3000, 24
    dt = np.float64 if dbl else np.float32
    r = rng.standard_normal((P, N)).astype(dt)
    sets = [np.sort(rng.choice(N, k, replace=False)).astype(np.int32)
            for k in (128, 129, 130, 160, 192, 200, 255, 256, 257)]
    sets += [np.concatenate([np.arange(70), np.arange(N - 70, N)]).astype(np.int32),
             np.arange(256, dtype=np.int32), np.arange(N - 129, N, dtype=np.int32)]
    sets += [np.sort(rng.choice(N, int(k), replace=False)).astype(np.int32)
             for k in rng.integers(129, 257, 40)]
    try:
        for pair_max in (0, 200, -1):
            gpu.set_tuning(es_pair_set_max=pair_max)
            for asc in (False, True):
                got, want, hi, lo = _es_rows_case(gpu, r, sets, asc=asc)
                parity.assert_es(got, want, hi, lo)
    finally:
        gpu.set_tuning()


@pytest.mark.parametrize("dbl", [False, True])
def test_rank_es_pair_sets_tie_rule(gpu, dbl):
    # |largest| == |smallest| running sum: the reference keeps the first.
    # Exact arithmetic: 256 genes whose |metric| is 2 (each hit adds
    # 2/512 = 1/256) among N = 2,304 (each miss subtracts 1/2048); equal
    # metrics rank by gene index, and gene g is given the g-th largest value.
    # (a) the top 128 and the bottom 128 ranks: +1/2 after hit 127, -1/2
    # just before hit 128, the odd lane's first; (b) the top 64 ranks, 1,024
    # misses, 128 hits each followed by 8 misses, then the bottom 64 ranks:
    # +1/4 after hit 63 and -1/4 just before hit 64, both on the even lane.
    N, k, P = 2304, 256, 4
    dt = np.float64 if dbl else np.float32
    a = np.full(N, -2.0)
    a[:128] = 2.0
    a[128:N - 128] = np.linspace(1.9, -1.9, N - 256)
    set_a = np.r_[0:128, N - 128:N]
    b = np.full(N, -2.0)
    b[:64] = 2.0
    b[64:1088] = np.linspace(1.9, -1.9, 1024)
    set_b = np.r_[0:64, 1088 + 9 * np.arange(128), N - 64:N]
    for base, genes, top in ((a, set_a, 0.5), (b, set_b, 0.25)):
        assert len(genes) == k and len(np.unique(genes)) == k
        r = np.tile(base.astype(dt), (P, 1))
        got, want, hi, lo = _es_rows_case(gpu, r, [genes.astype(np.int32)])
        assert (hi == top).all() and (lo == -top).all(), "the case must tie exactly"
        assert (want == top).all()
        assert bits_equal(got, want).all(), (top, got.ravel(), want.ravel())


@pytest.mark.parametrize("dbl", [False, True])
@pytest.mark.parametrize("k", [64, 200, 512])
def test_rank_es_tie_rule_lane_and_wavefront(gpu, k, dbl):
    # |largest| == |smallest| running sum on the one-lane path (64 genes), the
    # pair path (200) and the wavefront path (512): set genes have |metric| 2
    # (each hit adds 1/k) among N = k + 16k genes (each miss subtracts
    # 1/(16k)).  Equal metrics rank by gene index.
    # (max first) the top k/2 and bottom k/2 ranks are hits: +1/2 after hit
    # k/2-1, -1/2 just before hit k/2, so the ES is +1/2;
    # (min first) 8k misses (-1/2), then the k hits (+1/2), then 8k misses:
    # -1/2 comes first, so the ES is -1/2.
    # With k a power of two every step is dyadic and every sum exact: the GPU
    # must tie as the reference does and keep the first, bit for bit.  With
    # k = 200 the steps are not dyadic; the GPU walk carries c = |r|/N_R
    # unrounded in double, so its two extremes can differ in the last bits
    # where the reference's tie exactly, and either may win: the ES bar with
    # its sign flip at a tie applies (parity.es_check).
    dyadic = (k & (k - 1)) == 0

    def check(got, want, hi, lo, what):
        if dyadic:
            assert bits_equal(got, want).all(), (what, got.ravel(), want.ravel())
        else:
            parity.assert_es(got, want, hi, lo, what)
        # ascending order of -r ranks the genes as descending order of r
        # (equal values by gene index either way): the same scores, bit for bit
        mirror = _es_rows_case(gpu, -r, sets, asc=True)[0][:1]
        assert bits_equal(mirror, got).all(), (what, mirror.ravel(), got.ravel())

    N, P = 17 * k, 4
    dt = np.float64 if dbl else np.float32
    h = k // 2
    a = np.full(N, -2.0)
    a[:h] = 2.0
    a[h:N - h] = np.linspace(1.9, -1.9, N - 2 * h)
    set_a = np.r_[0:h, N - h:N]
    b = np.empty(N)
    b[:8 * k] = np.linspace(10.0, 3.0, 8 * k)
    b[8 * k:9 * k] = 2.0
    b[9 * k:] = np.linspace(1.9, -1.9, N - 9 * k)
    set_b = np.arange(8 * k, 9 * k)
    for base, genes, es in ((a, set_a, 0.5), (b, set_b, -0.5)):
        assert len(genes) == k
        r = np.tile(base.astype(dt), (P, 1))
        # 60 more sets in the same call (5..700 genes): they are scored by
        # the wavefronts and lanes that scored the tie, after it, so a tie
        # branch that left state behind (bitmap bits) shows in their scores
        rng = np.random.default_rng(k)
        extra = [np.sort(rng.choice(N, int(m), replace=False)).astype(np.int32)
                 for m in rng.integers(5, min(700, N), 60)]
        sets = [genes.astype(np.int32)] + extra
        got, want, hi, lo = _es_rows_case(gpu, r, sets)
        parity.assert_es(got[1:], want[1:], hi[1:], lo[1:], ("sets after the tie", k))
        got, want, hi, lo = got[:1], want[:1], hi[:1], lo[:1]
        assert (hi == 0.5).all() and (lo == -0.5).all(), "the case must tie exactly"
        assert (want == es).all(), (es, want.ravel())
        check(got, want, hi, lo, (k, es))
    # ties two ranks apart: all N = 2k metrics are 2, so ranks follow gene
    # index and a hit and a miss move the walk by the same step x = 1/k.
    # "hit, miss, miss, (hit, miss)*, hit" peaks at +x on rank 0 and first
    # falls to -x on rank 2 (ES +x); its mirror "miss, hit, hit, (miss,
    # hit)*, miss" gives -x first (ES -x)
    N = 2 * k
    r = np.full((P, N), 2.0, dtype=dt)
    hm = np.r_[[1, 0, 0], np.tile([1, 0], k - 2), [1]].astype(bool)
    for hits, sign in ((hm, 1.0), (~hm, -1.0)):
        genes = np.flatnonzero(hits).astype(np.int32)
        assert len(genes) == k
        sets = [genes]
        got, want, hi, lo = _es_rows_case(gpu, r, sets)
        assert (hi == -lo).all() and (hi > 0).all(), "the case must tie exactly"
        assert (np.sign(want) == sign).all(), (sign, want.ravel())
        check(got, want, hi, lo, (k, sign))


@pytest.mark.parametrize("N", [2047, 2048, 4095, 4096, 32767, 32768])
def test_rank_es_bitmap_capacity_boundaries(gpu, N):
    # the ES kernel's per-wavefront bitmap holds ranks 0..N (rank N is the
    # sentinel's): gene counts at and around multiples of 2,048 ranks (one
    # 32-bit word per lane) and the largest supported count, with small and
    # large sets that include the first and the last gene
    print("bitmap boundary N =", N, flush=True)
    rng = np.random.default_rng(N)
    w = small("c1", num_genes=N, num_perms=8, num_sets=2)
    sets = [np.sort(rng.choice(N, k, replace=False)).astype(np.int32)
            for k in (1, 5, 100, 129, 700, min(3000, N - 1))]
    sets += [np.array([0, N - 1], np.int32),
             np.unique(np.concatenate([[0, N - 1], rng.choice(N, 400, replace=False)]))
             .astype(np.int32),
             np.arange(N, dtype=np.int32)]
    w["set_offsets"], w["set_genes"] = csr(sets)
    w["G"] = len(sets)
    got, want, hi, lo = _es_stage_case(gpu, w, 8, "stable_signal2noise")
    parity.assert_es(got, want, hi, lo)


@pytest.mark.parametrize("knobs", [
    dict(es_lane_set_max=-1),                               # every set by a wavefront
    dict(es_lane_set_max=8, es_pair_set_max=-1, es_waves=12),
    dict(es_lane_set_max=32, es_waves=5),
    dict(es_lane_set_max=128, es_pair_set_max=256, es_waves=1),
])
@pytest.mark.parametrize("dbl", [False, True])
def test_rank_es_work_splits(gpu, knobs, dbl):
    # the ES stage's work split is a launch-shape choice: sets of every size
    # class, the empty set and single genes, through each path the knobs
    # select (tiny sets on the wavefront path, no lane pairs, few wavefronts
    # per workgroup), each per element against the oracle
    print("es work split", knobs, "fp64" if dbl else "fp32", flush=True)
    rng = np.random.default_rng(17)
    N = 5000
    w = small("c1", num_genes=N, num_perms=12, num_sets=2)
    sizes = [0, 1, 2, 7, 8, 9, 31, 33, 64, 65, 127, 128, 129, 200, 256, 257, 600, 2100]
    sets = [np.sort(rng.choice(N, k, replace=False)).astype(np.int32) for k in sizes]
    sets += [np.array([0], np.int32), np.array([N - 1], np.int32),
             np.arange(N - 300, N, dtype=np.int32)]
    w["set_offsets"], w["set_genes"] = csr(sets)
    w["G"] = len(sets)
    try:
        gpu.set_tuning(**knobs)
        for asc in (False, True):
            got, want, hi, lo = _es_stage_case(gpu, w, 12, "onepass_signal2noise", dbl=dbl,
                                               asc=asc)
            parity.assert_es(got, want, hi, lo)
    finally:
        gpu.set_tuning()


@pytest.mark.parametrize("dbl", [False, True])
def test_rank_es_large_set_bank_layouts(gpu, dbl):
    # large sets' gene ids are stored round by round, one id per LDS bank of
    # the 16-bit pos row ((id + 1) >> 1 mod 32): ids that all share one bank
    # (rounds of one id), two banks, every bank once (one full round), ragged
    # bank counts, and contiguous blocks; each set's ES against the oracle,
    # and a second identical call bit-equal to the first
    print("large-set bank layouts", dbl, flush=True)
    N = 20000
    rng = np.random.default_rng(11)
    ids = np.arange(N, dtype=np.int32)
    bank = ((ids + 1) >> 1) & 31
    sets = [ids[bank == 7][:600],
            ids[(bank == 0) | (bank == 31)][:900],
            np.sort(np.concatenate([rng.choice(ids[bank == b], 9, replace=False)
                                    for b in range(32)])),
            np.sort(np.concatenate([rng.choice(ids[bank == b], 1 + 3 * b, replace=False)
                                    for b in range(32)])),
            np.arange(5000, 7000, dtype=np.int32),
            np.arange(N - 257, N, dtype=np.int32),
            np.sort(rng.choice(N, 1999, replace=False)).astype(np.int32)]
    sets = [np.asarray(s, np.int32) for s in sets]
    w = small("c1", num_genes=N, num_perms=8, num_sets=2)
    w["set_offsets"], w["set_genes"] = csr(sets)
    w["G"] = len(sets)
    metric = "stable_signal2noise"
    got, want, hi, lo = _es_stage_case(gpu, w, 8, metric, dbl)
    parity.assert_es(got, want, hi, lo)
    again, _, _, _ = _es_stage_case(gpu, w, 8, metric, dbl)
    assert bits_equal(got, again).all()


# ----------------------------------------------------------- whole path ----
@pytest.mark.parametrize("dbl", [False, True])
def test_full_path_c1(gpu, dbl):
    print("full path c1", dbl, flush=True)
    w = small("c1")
    P = w["P"]
    want, es_want, hi, lo = oracle.gsea_ext(w["X"], w["labels"], w["set_offsets"],
                                            w["set_genes"], P, "stable_signal2noise",
                                            double=dbl)
    got, es_got = gpu.run(w["X"], w["labels"], w["set_offsets"], w["set_genes"], P,
                          "stable_signal2noise", double=dbl, want_es=True)
    parity.assert_es(es_got, es_want, hi, lo)
    parity.assert_stats(got, want, es_want, hi, lo)


@pytest.mark.parametrize("metric", oracle.METRICS)
def test_full_path_every_metric_small(gpu, metric):
    w = small("c3", num_genes=1500, num_sets=120, num_perms=200)
    want, es_want, hi, lo = oracle.gsea_ext(w["X"], w["labels"], w["set_offsets"],
                                            w["set_genes"], 200, metric)
    got, es_got = gpu.run(w["X"], w["labels"], w["set_offsets"], w["set_genes"], 200,
                          metric, want_es=True)
    parity.assert_es(es_got, es_want, hi, lo)
    parity.assert_stats(got, want, es_want, hi, lo)


def test_full_path_flags(gpu):
    w = small("c1", num_perms=128)
    for asc in (False, True):
        for swap in (False, True):
            want, es_want, hi, lo = oracle.gsea_ext(
                w["X"], w["labels"], w["set_offsets"], w["set_genes"], 128,
                "onepass_signal2noise", ascending=asc, swap=swap)
            got = gpu.run(w["X"], w["labels"], w["set_offsets"], w["set_genes"], 128,
                          "onepass_signal2noise", ascending=asc, swap_labels=swap)
            parity.assert_stats(got, want, es_want, hi, lo)


@pytest.mark.parametrize("dbl", [False, True])
def test_signed_nes_full_path(gpu, dbl):
    # GSEA_FLAG_SIGNED_NES: the GPU's ES against the oracle's walks, then the
    # GPU's statistics against the oracle's signed statistics of those same
    # ES, bit for bit (the same sums in the same order, the same products and
    # counts, the FDR ratio in double)
    print("signed NES", "fp64" if dbl else "fp32", flush=True)
    w = small("c3", num_genes=1500, num_sets=150, num_perms=256)
    P = 256
    args = (w["X"], w["labels"], w["set_offsets"], w["set_genes"], P, "stable_signal2noise")
    _, es_want, hi, lo = oracle.gsea_ext(*args, double=dbl, signed_nes=True)
    got, es_got = gpu.run(*args, double=dbl, want_es=True, signed_nes=True)
    parity.assert_es(es_got, es_want, hi, lo)
    want, _ = oracle.final_statistics(es_got, signed_nes=True)
    same = bits_equal(got, want)
    assert same.all(), (np.argwhere(~same)[:5].tolist(), got[~same][:5], want[~same][:5])
    # ES and NP are the reference's; NES follows the per-sign factors
    ref = gpu.run(*args, double=dbl)
    assert bits_equal(got[0], ref[0]).all() and bits_equal(got[2], ref[2]).all()
    assert not np.allclose(got[1], ref[1])


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
    want, es_want, hi, lo = oracle.gsea_ext(w["X"], w["labels"], w["set_offsets"],
                                            w["set_genes"], 64, "stable_signal2noise")
    got = res.to_numpy().T.astype(np.float32)
    parity.assert_stats(got, want, es_want, hi, lo)
    dump = gpu.read_dump(prefix + "_%d_64_32.es" % w["G"])
    parity.assert_es(dump, es_want, hi, lo)


# ------------------------------------------------- genome-scale configs ----
def test_proxy_device_functions_cpp(gpu):
    # tests/cpp/proxy_devices_test (built with the library): the device
    # proxies' strings, selection and fallback warning against
    # cudaGSEA_wrapper.cu:87-138, checked in C++ against hipGetDeviceProperties;
    # here also against torch's view of the same device
    import os
    import subprocess
    import torch
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build",
                       "proxy_devices_test")
    assert os.path.exists(exe), "build the library first (it builds tests/cpp)"
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    n = torch.cuda.device_count()
    assert lines[-1] == "ok %d device(s), 0 failure(s)" % n
    prop = torch.cuda.get_device_properties(0)
    gib = (prop.total_memory + (1 << 30) - 1) >> 30
    # the runtime's device name, or its ISA name where the runtime reports an
    # empty name (torch then substitutes a generic one)
    names = {prop.name, getattr(prop, "gcnArchName", prop.name)}
    assert lines[0] in {"device 0: %s with %d GiB RAM" % (n, gib) for n in names}, lines[0]


def test_c2_fp32_full_size_properties(gpu):
    import torch
    print("c2 fp32 full size", flush=True)
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
    # permutations 0..63 of the full run against the oracle, element by element
    want_r = oracle.correlate(w["X"], w["labels"], 64, "onepass_t_test", sample_major=True)
    sk, idx = oracle.pathway_sort(want_r)
    want_es, hi, lo = oracle.scores_ext(sk, idx, w["set_offsets"], w["set_genes"])
    parity.assert_es(es1[:, :64], want_es, hi, lo)


def test_c2_fp64_full_size(gpu):
    # doublePrecision = TRUE at the configuration's genes, samples and sets,
    # on a 64-permutation subset: metric rows and ranks bit-exact against
    # the fp64 oracle, statistics to the bars above
    import torch
    print("c2 fp64 full size, 64 permutations", flush=True)
    P = 64
    w = small("c2", dtype=np.float64, num_perms=P)
    N, G = w["N"], w["G"]
    X = dev(w["X"])
    r = torch.empty((P, N), dtype=torch.float64, device="cuda")
    gpu.metric_device(X, w["labels"], P, "onepass_t_test", r)
    want_r = oracle.correlate(w["X"], w["labels"], P, "onepass_t_test", sample_major=True,
                              double=True)
    assert bits_equal(r.cpu().numpy(), want_r).all()
    idx = torch.empty((P, N), dtype=torch.int32, device="cuda")
    gpu.argsort_device(r, idx)
    _, want_idx = oracle.pathway_sort(want_r)
    assert (idx.cpu().numpy().view(np.uint32) == want_idx).all()
    want, es_want, hi, lo = oracle.gsea_ext(w["X"], w["labels"], w["set_offsets"],
                                            w["set_genes"], P, "onepass_t_test", double=True)
    st = torch.empty((5, G), dtype=torch.float64, device="cuda")
    es = torch.empty((G, P), dtype=torch.float64, device="cuda")
    gpu.run_device(X, w["labels"], dev(w["set_offsets"]), dev(w["set_genes"]), P,
                   "onepass_t_test", st, es)
    parity.assert_es(es.cpu().numpy(), es_want, hi, lo)
    parity.assert_stats(st.cpu().numpy(), want, es_want, hi, lo)


def test_c4_fp32_against_fp64_oracle(gpu):
    # numerically hard ranking data: per-gene mean 1e4, sd 1e-2, 1% constant
    # genes.  Each precision is bit-exact with its own oracle (metric rows and
    # ranks) and within the ES bar of it; the fp32 path's error against the
    # fp64 oracle is a property of fp32 on this data (class means near 1e4
    # carry one ulp = 2^-10 of rounding, the size of the mean differences
    # themselves) and is stated here: see the figures printed and bounded.
    import torch
    P = 32
    w32 = small("c4", num_perms=P)
    X64 = w32["X"].astype(np.float64)
    N, G = w32["N"], w32["G"]
    const = np.all(w32["X"] == w32["X"][0], axis=0)
    rho_min = {"twopass_signal2noise": 0.3, "stable_t_test": 0.3, "overkill_t_test": 0.85}
    for m in w32["metrics"]:
        print("c4", m, flush=True)
        r32 = torch.empty((P, N), dtype=torch.float32, device="cuda")
        r64 = torch.empty((P, N), dtype=torch.float64, device="cuda")
        gpu.metric_device(dev(w32["X"]), w32["labels"], P, m, r32)
        gpu.metric_device(dev(X64), w32["labels"], P, m, r64)
        o32 = oracle.correlate(w32["X"], w32["labels"], P, m, sample_major=True)
        o64 = oracle.correlate(X64, w32["labels"], P, m, sample_major=True, double=True)
        assert bits_equal(r32.cpu().numpy(), o32).all()
        assert bits_equal(r64.cpu().numpy(), o64).all()
        i32 = torch.empty((P, N), dtype=torch.int32, device="cuda")
        i64 = torch.empty((P, N), dtype=torch.int32, device="cuda")
        gpu.argsort_device(r32, i32)
        gpu.argsort_device(r64, i64)
        g32 = i32.cpu().numpy().view(np.uint32)
        g64 = i64.cpu().numpy().view(np.uint32)
        assert (g32 == oracle.pathway_sort(o32)[1]).all()
        assert (g64 == oracle.pathway_sort(o64)[1]).all()
        # tie rule: the constant genes (metric exactly 0 in both precisions)
        # keep ascending gene order among themselves
        for g in (g32, g64):
            for p in range(P):
                z = g[p][np.isin(g[p], np.flatnonzero(const))]
                assert (np.diff(z.astype(np.int64)) > 0).all()
        # fp32 vs fp64: rank correlation and ES / NES error (stated bounds)
        pos32 = np.argsort(g32, axis=1)
        pos64 = np.argsort(g64, axis=1)
        rho = float(np.mean([np.corrcoef(pos32[p], pos64[p])[0, 1] for p in range(P)]))
        st32 = torch.empty((5, G), dtype=torch.float32, device="cuda")
        st64 = torch.empty((5, G), dtype=torch.float64, device="cuda")
        gpu.run_device(dev(w32["X"]), w32["labels"], dev(w32["set_offsets"]),
                       dev(w32["set_genes"]), P, m, st32)
        gpu.run_device(dev(X64), w32["labels"], dev(w32["set_offsets"]),
                       dev(w32["set_genes"]), P, m, st64)
        want64, es64, hi64, lo64 = oracle.gsea_ext(X64, w32["labels"], w32["set_offsets"],
                                                   w32["set_genes"], P, m, double=True)
        parity.assert_stats(st64.cpu().numpy(), want64, es64, hi64, lo64)
        d_es = np.abs(st32.cpu().numpy()[0].astype(np.float64) - want64[0])
        d_nes = np.abs(st32.cpu().numpy()[1].astype(np.float64) - want64[1])
        print("  fp32 vs fp64: rank correlation %.3f, observed ES max |d| %.3f, "
              "NES max |d| %.3f" % (rho, d_es.max(), d_nes.max()), flush=True)
        assert rho >= rho_min[m]
        assert d_es.max() <= 2.0 and np.isfinite(d_nes).all()


def test_multi_batch_bit_identical(gpu):
    # permutation batches (the reference's batch loop,
    # cudaGSEA_wrapper.cu:295-303): forcing 3 batches of at most 100
    # permutations gives the single batch's bits
    import torch
    for dbl in (False, True):
        w = small("c1", num_perms=256, dtype=np.float64 if dbl else np.float32)
        P, G = 256, w["G"]
        tdt = torch.float64 if dbl else torch.float32
        args = (dev(w["X"]), w["labels"], dev(w["set_offsets"]), dev(w["set_genes"]), P,
                "stable_signal2noise")
        st1 = torch.empty((5, G), dtype=tdt, device="cuda")
        es1 = torch.empty((G, P), dtype=tdt, device="cuda")
        t1 = gpu.run_device(*args, st1, es1)
        try:
            gpu.set_tuning(batch_perms=100)
            st2 = torch.empty((5, G), dtype=tdt, device="cuda")
            es2 = torch.empty((G, P), dtype=tdt, device="cuda")
            t2 = gpu.run_device(*args, st2, es2)
        finally:
            gpu.set_tuning()
        assert t1["perm_batches"] == 1 and t2["perm_batches"] == 3
        assert bits_equal(st1.cpu().numpy(), st2.cpu().numpy()).all()
        assert bits_equal(es1.cpu().numpy(), es2.cpu().numpy()).all()


# --------------------------------------------- beyond one LDS sort ----
@pytest.mark.parametrize("dbl", [False, True])
@pytest.mark.parametrize("N", [40000, 60000])
def test_large_gene_counts_rank_and_es(gpu, N, dbl):
    # more genes than one workgroup's LDS sort (32,768): the rank rows come
    # from the global-scratch radix passes, the ES kernel reads |metric|
    # from L2; ranks bit-exact, ES to the bar, with small and large sets
    import torch
    print("large N", N, "fp64" if dbl else "fp32", flush=True)
    rng = np.random.default_rng(N + dbl)
    dt = np.float64 if dbl else np.float32
    P = 6
    keys = rng.standard_normal((P, N)).astype(dt)
    keys[0, :2000] = 0.0                                # ties
    keys[1, ::101] = np.nan
    _, want_idx = oracle.pathway_sort(keys)
    idx = torch.empty((P, N), dtype=torch.int32, device="cuda")
    gpu.argsort_device(dev(keys), idx)
    assert (idx.cpu().numpy().view(np.uint32) == want_idx).all()
    w = small("c1", num_genes=N, num_perms=P, num_sets=2)
    sets = [np.sort(rng.choice(N, k, replace=False)).astype(np.int32)
            for k in (1, 7, 60, 128, 129, 500, 2000, 9000)]
    sets += [np.array([0, N - 1], np.int32), np.arange(N - 300, N, dtype=np.int32)]
    w["set_offsets"], w["set_genes"] = csr(sets)
    w["G"] = len(sets)
    got, want, hi, lo = _es_stage_case(gpu, w, P, "stable_signal2noise", dbl)
    parity.assert_es(got, want, hi, lo)


def test_large_gene_count_full_path(gpu):
    print("full path, 40,000 genes", flush=True)
    w = small("c1", num_genes=40000, num_perms=64, num_sets=40)
    want, es_want, hi, lo = oracle.gsea_ext(w["X"], w["labels"], w["set_offsets"],
                                            w["set_genes"], 64, "stable_signal2noise")
    got, es_got = gpu.run(w["X"], w["labels"], w["set_offsets"], w["set_genes"], 64,
                          "stable_signal2noise", want_es=True)
    parity.assert_es(es_got, es_want, hi, lo)
    parity.assert_stats(got, want, es_want, hi, lo)


def test_gene_limit_is_65535(gpu):
    import torch
    rng = np.random.default_rng(5)
    keys = rng.standard_normal((2, 65535)).astype(np.float32)
    _, want = oracle.pathway_sort(keys)
    idx = torch.empty((2, 65535), dtype=torch.int32, device="cuda")
    gpu.argsort_device(dev(keys), idx)
    assert (idx.cpu().numpy().view(np.uint32) == want).all()
    keys2 = np.zeros((1, 65536), np.float32)
    idx2 = torch.empty((1, 65536), dtype=torch.int32, device="cuda")
    with pytest.raises(gpu.GseaError) as e:
        gpu.argsort_device(dev(keys2), idx2)
    assert e.value.code == 31
