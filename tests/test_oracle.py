"""Pins the CPU oracle (oracle/).

The RNG is pinned against the reference's own generator: rngpu/rngpu.hpp
compiled as shipped by g++ (oracle/ref), whose outputs are committed in
tests/golden/ref_vectors.npz, plus the known answer that permutation 1 of
[0..9] is [5 1 2 3 6 4 8 9 0 7].  The rest of the
reference's path cannot be built here (its headers need nvcc's keywords, the
CUDA runtime or CUB), so the other oracle functions are pinned by closed forms
on exactly representable data, float64 brute force written from the
reference's definitions, accumulator-specific error behaviour on
cancellation-prone data, and invariants -- each chosen so that a plausible
mistake (a swapped accumulator, a dropped floor, > for >=, a skipped p = 0, the
wrong tail) fails one of them."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ref_vectors.npz"))
LOG2_METRICS = ("naive_log2_ratio_of_classes", "stable_log2_ratio_of_classes")


def bits_equal(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    u = np.uint64 if a.dtype == np.float64 else np.uint32
    same = a.view(u) == b.view(u)
    return same | (np.isnan(a) & np.isnan(b))


def ulps(got, want, dt):
    """|got - want| in units of the spacing of `want` in dtype dt."""
    w = np.asarray(want, dtype=np.float64)
    sp = np.spacing(np.abs(w).astype(dt)).astype(np.float64)
    return np.abs(np.asarray(got, dtype=np.float64) - w) / sp


# ---------------------------------------------------------------- RNG ------
def test_rng_known_answer(orc):
    # permutation 1 of [0..9] computed with the reference's rngpu.hpp
    got = orc.permute_labels(np.arange(10), 1)
    assert got.tolist() == [5, 1, 2, 3, 6, 4, 8, 9, 0, 7]


@pytest.mark.parametrize("S", [2, 3, 10, 40, 200, 500])
def test_rng_matches_reference_vectors(orc, S):
    base = (np.arange(S) % 251).astype(np.uint8) if S > 255 else np.arange(S, dtype=np.uint8)
    want = GOLD["perm_S%d" % S]
    for k, p in enumerate(GOLD["perm_ids"]):
        assert (orc.permute_labels(base, int(p)) == want[k]).all(), (S, p)


def test_rng_streams_and_hash_match_reference(orc):
    for k, seed in enumerate(GOLD["kiss_seeds"]):
        assert (orc.kiss_stream(int(seed), 64) == GOLD["kiss"][k]).all(), seed
    for x, want in zip(GOLD["hash_in"], GOLD["hash_out"]):
        assert orc.lib().orc_hash32(int(x)) == int(want)


@pytest.mark.skipif(not __import__("oracle").ref_available(),
                    reason="reference generator (oracle/_ref) not built")
def test_rng_matches_reference_build_random(orc):
    rng = np.random.default_rng(12)
    for S in (7, 64, 150, 333):
        lab = rng.integers(0, 2, S).astype(np.uint8)
        for p in rng.integers(1, 2 ** 31, 20):
            assert (orc.permute_labels(lab, int(p)) ==
                    orc.permute_labels(lab, int(p), use_ref=True)).all()


def test_rng_is_a_permutation_and_perm0_identity(orc):
    lab = np.arange(200, dtype=np.uint8)
    assert (orc.permute_labels(lab, 0) == lab).all()
    for p in range(1, 50):
        assert sorted(orc.permute_labels(lab, p).tolist()) == list(range(200))


def test_rng_uniformity_chi2(orc):
    # position of element 0 after shuffling 8 items, over 8000 permutations
    lab = np.arange(8, dtype=np.uint8)
    counts = np.zeros(8)
    for p in range(1, 8001):
        counts[int(np.argmax(orc.permute_labels(lab, p) == 0))] += 1
    chi2 = ((counts - 1000) ** 2 / 1000).sum()
    assert chi2 < 30, counts        # 7 dof, p ~ 1e-4


# ------------------------------------------------------------ metrics ------
# Closed forms, functors.cuh:10-80 (combiners, fixlow = true, biased = true as
# correlate_genes.cuh:121-127 instantiates them) over class data whose sums,
# means, running (Welford) means and biased variances are exact in binary
# floating point, so every accumulator yields the exact moments and only the
# combiner's last operations round.  Class A = label 0 (functors.cuh:110).
#   gene 0: A = 1 3 2 6      (mean 3, var 3.5)  B = 10 30 20 60 (30, 350)
#   gene 1: A = 100 102 101 105 (102, 3.5): both floors active on A
#           B = 4 12 8 24    (12, 56)
#   gene 2: A = -1 -3 -2 -6  (-3, 3.5)          B = 1 3 2 6 (3, 3.5)
#   gene 3: constant 5; gene 4: all zero
_CF_A = [[1, 3, 2, 6], [100, 102, 101, 105], [-1, -3, -2, -6], [5] * 4, [0] * 4]
_CF_B = [[10, 30, 20, 60], [4, 12, 8, 24], [1, 3, 2, 6], [5] * 4, [0] * 4]
_CF_M = [(3.0, 3.5, 30.0, 350.0), (102.0, 3.5, 12.0, 56.0), (-3.0, 3.5, 3.0, 3.5),
         (5.0, 0.0, 5.0, 0.0), (0.0, 0.0, 0.0, 0.0)]


def _cf_expected(metric, mA, vA, mB, vB, n=4.0):
    if "diff" in metric:
        return mA - mB
    if "log2" in metric:
        q = mA / mB if mB != 0 else np.nan
        return np.log(q) * 1.44269504089 if q >= 0 else np.nan
    if "ratio" in metric:
        return mA / mB if mB != 0 else np.nan
    if "signal2noise" in metric:
        sA = max(np.sqrt(vA), 0.2 * abs(mA))
        sB = max(np.sqrt(vB), 0.2 * abs(mB))
        return (mA - mB) / (sA + sB) if sA + sB else np.nan
    vA = max(vA, 0.04 * abs(mA))
    vB = max(vB, 0.04 * abs(mB))
    den = np.sqrt(vA / n + vB / n)
    return (mA - mB) / den if den else np.nan


@pytest.mark.parametrize("dbl", [False, True])
def test_metric_closed_forms_every_metric(orc, dbl):
    dt = np.float64 if dbl else np.float32
    X = np.array([a + b for a, b in zip(_CF_A, _CF_B)], dtype=dt)   # [N][S]
    lab = np.array([0] * 4 + [1] * 4, dtype=np.uint8)
    for m in orc.METRICS:
        got = orc.correlate(X, lab, 1, m, double=dbl)[0]
        for g, (mA, vA, mB, vB) in enumerate(_CF_M):
            want = _cf_expected(m, mA, vA, mB, vB)
            if np.isnan(want):
                assert np.isnan(got[g]), (m, g, got[g])
            elif want == 0:
                assert got[g] == 0, (m, g, got[g])
            else:
                assert ulps(got[g], want, dt) <= 2.0, (m, g, got[g], want)
    # the closed forms tell the formulas apart: on gene 0 no two combiners
    # agree, and the floors change gene 1
    assert _cf_expected("twopass_signal2noise", 102, 3.5, 12, 56) != \
        (102 - 12) / (np.sqrt(3.5) + np.sqrt(56))
    assert _cf_expected("twopass_t_test", 102, 3.5, 12, 56) != \
        (102 - 12) / np.sqrt(3.5 / 4 + 56 / 4)


def test_metric_plain_vs_compensated_sum_closed_form(orc):
    # class A = 2^24, 1, 1, 1 in fp32: a plain sum rounds every +1 away
    # (2^24 + 1 ties to 2^24), mean 2^22; the Kahan sum of
    # functors.cuh:144-172 carries them, rounds 2^24 + 3 to 2^24 + 4 and gives
    # mean 2^22 + 1; Welford's running mean (functors.cuh:396-428) ends at
    # fl(2^22 + 0.75) = 2^22 + 1.  Class B = 1 1 1 1.
    X = np.array([[2.0 ** 24, 1, 1, 1, 1, 1, 1, 1]], dtype=np.float32)
    lab = np.array([0] * 4 + [1] * 4, dtype=np.uint8)
    r = lambda m: float(orc.correlate(X, lab, 1, m)[0, 0])
    assert r("naive_diff_of_classes") == 2.0 ** 22 - 1
    assert r("naive_ratio_of_classes") == 2.0 ** 22
    assert r("stable_diff_of_classes") == 2.0 ** 22
    assert r("stable_ratio_of_classes") == 2.0 ** 22 + 1
    # fp64 carries every unit: 2^22 + 0.75 either way
    X64 = X.astype(np.float64)
    for m in ("naive_ratio_of_classes", "stable_ratio_of_classes"):
        assert float(orc.correlate(X64, lab, 1, m, double=True)[0, 0]) == 2.0 ** 22 + 0.75


@pytest.mark.parametrize("dbl", [False, True])
def test_metric_float64_agreement_every_metric(orc, dbl):
    rng = np.random.default_rng(5)
    N, S, nA = 300, 30, 12
    X = np.exp(5 + rng.standard_normal((N, S)))
    X[:40, :nA] *= 1.8
    lab = np.array([0] * nA + [1] * (S - nA), dtype=np.uint8)
    dt = np.float64 if dbl else np.float32
    Xt = X.astype(dt).astype(np.float64)
    A, B = Xt[:, lab == 0], Xt[:, lab == 1]
    mA, mB, vA, vB = A.mean(1), B.mean(1), A.var(1), B.var(1)
    tol = 1e-12 if dbl else 5e-5
    mag = np.abs(mA) + np.abs(mB)       # size of the operands of mA - mB
    for m in orc.METRICS:
        if "diff" in m:
            want, scale = mA - mB, mag
        elif "log2" in m:
            want = np.log(mA / mB) * 1.44269504089
            scale = np.abs(want) + 1e-3
        elif "ratio" in m:
            want, scale = mA / mB, mA / mB
        elif "signal2noise" in m:
            den = (np.maximum(np.sqrt(vA), 0.2 * np.abs(mA)) +
                   np.maximum(np.sqrt(vB), 0.2 * np.abs(mB)))
            want, scale = (mA - mB) / den, mag / den
        else:
            den = np.sqrt(np.maximum(vA, 0.04 * np.abs(mA)) / nA +
                          np.maximum(vB, 0.04 * np.abs(mB)) / (S - nA))
            want, scale = (mA - mB) / den, mag / den
        got = orc.correlate(X.astype(dt), lab, 1, m, double=dbl)[0].astype(np.float64)
        assert (np.abs(got - want) <= tol * scale).all(), m


def test_metric_onepass_variance_cancels_and_the_others_do_not(orc):
    # t-test with the variance floor 0.04|mean| inactive (mean 1e4, sd 25):
    # E[x^2] - E[x]^2 in fp32 (functors.cuh:197-218) loses ~mean^2/var * eps
    # of the variance, the two-pass, Welford and compensated two-pass
    # variances (functors.cuh:251-290, 396-428, 317-370) do not
    rng = np.random.default_rng(8)
    N, S = 400, 40
    X = 1e4 + 25.0 * rng.standard_normal((N, S))
    lab = np.array([0, 1] * (S // 2), dtype=np.uint8)
    Xt = X.astype(np.float32).astype(np.float64)
    A, B = Xt[:, lab == 0], Xt[:, lab == 1]
    want = (A.mean(1) - B.mean(1)) / np.sqrt(A.var(1) / 20 + B.var(1) / 20)

    def med_rel_err(m):
        got = orc.correlate(X.astype(np.float32), lab, 1, m)[0].astype(np.float64)
        return float(np.median(np.abs(got - want) / np.abs(want)))
    # one-pass: relative variance error ~ mean^2/var * u = 1.6e5 * 6e-8 ~ 1e-2;
    # the others: the class means' rounding (~ulp(1e4) over a mean
    # difference ~8) ~ 1e-4
    assert med_rel_err("onepass_t_test") > 3e-3
    for m in ("twopass_t_test", "stable_t_test", "overkill_t_test"):
        assert med_rel_err(m) < 1e-3, m


def test_metric_accumulators_told_apart_on_cancellation_data(orc):
    # c4-style data: per-gene mean 1e4, sd 1e-2, 1% constant genes.  The
    # floors (0.2|mean| on sd, 0.04|mean| on variance) dominate, so every
    # s2n / t metric is the class-mean difference over a known denominator
    # and the class means decide: plain recursive sums (naive, one-pass,
    # two-pass), Kahan sums (stable diff/ratio, overkill) and Welford running
    # means (stable s2n / t) round differently.  Kahan's bound
    # |err| <= 2u sum|x| + O(n u^2) puts each Kahan mean within 1.22 + 0.5
    # ulps of the exact mean (ulp = 2^-10 at 1e4), so a mean difference within
    # 3.5 ulps; plain sums and Welford's n divisions have no such bound.
    from cudagsea_amd import workloads
    w = workloads.make("c4", num_genes=4000, num_perms=1)
    X = w["X"].astype(np.float32)
    lab = w["labels"]
    nA = int((lab == 0).sum())
    nB = int((lab == 1).sum())
    Xd = X.astype(np.float64)
    mA, mB = Xd[lab == 0].mean(0), Xd[lab == 1].mean(0)
    ulp = 2.0 ** -10

    def mean_diff(m):
        r = orc.correlate(X, lab, 1, m, sample_major=True)[0].astype(np.float64)
        if "signal2noise" in m:
            return r * 0.2 * (np.abs(mA) + np.abs(mB))
        if "t_test" in m:
            return r * np.sqrt(0.04 * np.abs(mA) / nA + 0.04 * np.abs(mB) / nB)
        return r

    err = {m: (mean_diff(m) - (mA - mB)) / ulp for m in
           ("naive_diff_of_classes", "onepass_signal2noise", "twopass_signal2noise",
            "twopass_t_test", "stable_diff_of_classes", "overkill_signal2noise",
            "overkill_t_test", "stable_signal2noise", "stable_t_test")}
    rms = {m: float(np.sqrt(np.mean(e ** 2))) for m, e in err.items()}
    comp = ("stable_diff_of_classes", "overkill_signal2noise", "overkill_t_test")
    plain = ("naive_diff_of_classes", "onepass_signal2noise", "twopass_signal2noise",
             "twopass_t_test")
    welf = ("stable_signal2noise", "stable_t_test")
    for m in comp:
        assert np.abs(err[m]).max() <= 3.5, (m, np.abs(err[m]).max())
    # one set of means per group: the group members agree to far below an ulp
    for grp in (comp, plain, welf):
        for m in grp[1:]:
            assert np.abs(err[m] - err[grp[0]]).max() < 1e-2, (grp[0], m)
    assert min(rms[m] for m in plain) > 1.5 * max(rms[m] for m in comp), rms
    assert min(rms[m] for m in welf) > 1.5 * max(rms[m] for m in plain), rms
    # constant genes give exactly 0 with every accumulator
    const = np.all(X == X[0], axis=0)
    assert const.sum() >= 40
    for m in err:
        assert (orc.correlate(X, lab, 1, m, sample_major=True)[0][const] == 0).all(), m


@pytest.mark.parametrize("dbl", [False, True])
def test_metric_log_is_correctly_rounded(orc, dbl):
    # the log2-ratio metrics use the oracle's own logarithm on both CPU and
    # GPU: in fp64 it is faithful (within 1 ulp of the long double log), in
    # fp32 (fp64 log rounded once) it is the correctly rounded log on every
    # sampled binade but for double-rounding cases within 1 ulp
    rng = np.random.default_rng(4)
    x = np.exp(rng.uniform(-700 if dbl else -85, 700 if dbl else 85, 20000))
    x = np.concatenate([x, [1.0, 2.0, 0.5, np.e, 1e-300 if dbl else 1e-40]])
    if dbl:
        got = np.array([orc.lib().orc_log_f64(float(v)) for v in x])
        want = np.log(x.astype(np.longdouble))
        err = np.abs(got.astype(np.longdouble) - want) / np.spacing(np.abs(got))
        assert float(err.max()) < 1.0
        assert float((err <= 0.5 + 1e-3).mean()) > 0.95
    else:
        xf = x.astype(np.float32)
        got = np.array([orc.lib().orc_log_f32(float(v)) for v in xf], dtype=np.float32)
        want = np.log(xf.astype(np.float64)).astype(np.float32)
        assert (got == want).mean() > 0.9999
        assert (ulps(got, np.log(xf.astype(np.float64)), np.float32) <= 1.0).all()


# --------------------------------------------------------------- sort ------
def _stated_order(keys, ascending=False):
    k = np.asarray(keys)
    f = np.where(np.isnan(k), np.inf, k)
    # descending: larger first; +0 before -0; NaN first; ties by gene id
    sign = np.signbit(k) & (k == 0)
    nan = np.isnan(k)
    if ascending:
        return np.lexsort((np.arange(k.size), ~sign, f, nan))
    return np.lexsort((np.arange(k.size), sign, -f, ~nan))


def test_sort_stated_tie_rule(orc):
    rng = np.random.default_rng(11)
    keys = rng.integers(-3, 4, size=(6, 500)).astype(np.float32)
    keys[0, :40] = 0.0
    keys[0, 40:80] = -0.0
    keys[1, 5] = np.nan
    keys[1, 9] = np.inf
    keys[2] = rng.standard_normal(500)
    for asc in (False, True):
        sk, idx = orc.pathway_sort(keys, ascending=asc)
        for p in range(6):
            want = _stated_order(keys[p], asc)
            assert (idx[p] == want).all(), (asc, p)
            assert bits_equal(sk[p], keys[p][want]).all()


def test_sort_key_is_monotone(orc):
    vals = np.array([-np.inf, -3.5, -1e-30, -0.0, 0.0, 1e-30, 2.0, np.inf, np.nan],
                    dtype=np.float32)
    k = [orc.lib().orc_sort_key_f32(float(v), 1) for v in vals]
    assert all(a < b for a, b in zip(k, k[1:]))


# ----------------------------------------------------------------- ES ------
def _es_walk64(r_sorted, hits):
    """Brute-force running sum (enrichment_scores.cuh:289-330) in float64."""
    N = r_sorted.size
    a = np.abs(r_sorted.astype(np.float64))
    NR = a[hits].sum()
    nh = hits.sum()
    steps = np.where(hits, a / NR if NR else np.nan, -1.0 / (N - nh) if N > nh else 0.0)
    res = np.cumsum(steps)
    best = 0.0
    for v in res:
        if abs(v) > abs(best):
            best = v
    return best


def _one_set(orc, r_sorted, genes_at_rank, members):
    off = np.array([0, len(members)], dtype=np.int32)
    return orc.scores(r_sorted[None, :], genes_at_rank[None, :], off,
                      np.array(sorted(members), dtype=np.int32))[0, 0]


def test_es_closed_forms(orc):
    N = 100
    r = np.linspace(2, -2, N).astype(np.float32)
    idx = np.arange(N, dtype=np.uint32)
    # empty set: N steps of -1/N
    assert abs(_one_set(orc, r, idx, []) + 1.0) < 1e-6
    # single hit at 0-based rank q: max(1 - q/(N-1), q/(N-1)) with sign
    for q in (0, 10, 49, 50, 77, 99):
        want = 1 - q / (N - 1) if 1 - q / (N - 1) > q / (N - 1) else -q / (N - 1)
        assert abs(_one_set(orc, r, idx, [q]) - want) < 2e-6, q
    # every gene in the set: the walk ends at 1
    assert abs(_one_set(orc, r, idx, list(range(N))) - 1.0) < 1e-6
    # single hit at rank 0 of [3,2,1,0] gives 2/3 (first strict max)
    assert abs(_one_set(orc, np.array([3, 2, 1, 0], np.float32),
                        np.arange(4, dtype=np.uint32), [0]) - 1.0) < 1e-7


def test_es_bruteforce_float64(orc):
    rng = np.random.default_rng(21)
    N, P = 400, 6
    r = rng.standard_normal((P, N)).astype(np.float32)
    sk, idx = orc.pathway_sort(r)
    sets = [np.sort(rng.choice(N, size=k, replace=False)) for k in (1, 5, 17, 60, 200, 399)]
    off = np.zeros(len(sets) + 1, dtype=np.int32)
    off[1:] = np.cumsum([len(s) for s in sets])
    es = orc.scores(sk, idx, off, np.concatenate(sets).astype(np.int32))
    for p in range(P):
        for s, mem in enumerate(sets):
            hits = np.isin(idx[p], mem)
            want = _es_walk64(sk[p], hits)
            assert abs(es[s, p] - want) < 1e-5 * max(1, abs(want)), (s, p)
            assert abs(es[s, p]) <= 1.0 + 1e-6


def test_es_tile_invariance(orc):
    # the same set at different positions of the set list (different 64-set
    # bitmap tiles) gives the same score
    rng = np.random.default_rng(2)
    N = 300
    r = rng.standard_normal((2, N)).astype(np.float32)
    sk, idx = orc.pathway_sort(r)
    mem = np.sort(rng.choice(N, 30, replace=False))
    sets = [mem] + [np.sort(rng.choice(N, 10, replace=False)) for _ in range(70)] + [mem]
    off = np.zeros(len(sets) + 1, dtype=np.int32)
    off[1:] = np.cumsum([len(s) for s in sets])
    es = orc.scores(sk, idx, off, np.concatenate(sets).astype(np.int32))
    assert bits_equal(es[0], es[-1]).all()


# -------------------------------------------------------------- stats ------
def _stats_grid_case():
    """ES matrix [G][P] on the grid k/64 (|k| <= 64).  Every non-zero row is
    a permutation of one multiset, so every row has the same sum of |ES| (an
    exact fp32 sum) and the same normalising factor; comparisons between
    normalised scores are then comparisons between grid values.  Rows:
    0  ES unique maximum of its row            (NP = 1/P: p = 0 counts)
    1  ES tied with three other permutations   (>= vs >)
    2  negative ES                             (lower tail for FWER)
    3  ES = 0 (NES = 0, lower-tail branch)
    4  all-zero row                            (factor inf, NES NaN)
    5  positive ES below other rows' maxima    (FWER from other sets)"""
    P = 16
    base = np.array([40, -30, 25, -25, 20, 20, 20, -10, 5, 0, -5, 12, -38, 33, -2, 8],
                    dtype=np.float64)
    rng = np.random.default_rng(3)
    rows = []

    def row(first):
        rest = list(base)
        rest.remove(first)
        rng.shuffle(rest)
        return [first] + rest
    rows.append(row(40))
    rows.append(row(20))
    rows.append(row(-30))
    rows.append(row(0))
    rows.append([0.0] * P)
    rows.append(row(12))
    return np.array(rows) / 64.0


def _stats_brute(es):
    """statistics.cuh:46-117 in float64: NES = ES * P / sum_p |ES_p| (all p,
    p = 0 included); NP = #{p : |ES_p| >= |ES_0|} / P (p = 0 included);
    normalised z = ES * factor; FWER = #{p : max_s z[s,p] >= NES} / P for
    NES > 0, else #{p : min_s z[s,p] <= NES} / P, NaN never counted."""
    G, P = es.shape
    with np.errstate(divide="ignore", invalid="ignore"):
        fac = P / np.abs(es).sum(axis=1)
        z = es * fac[:, None]
    nes = z[:, 0]
    np_ = (np.abs(es) >= np.abs(es[:, :1])).sum(axis=1) / P
    zmax = np.nanmax(np.where(np.isnan(z), -np.inf, z), axis=0)
    zmin = np.nanmin(np.where(np.isnan(z), np.inf, z), axis=0)
    fw = np.empty(G)
    for s in range(G):
        if nes[s] > 0:
            fw[s] = (zmax >= nes[s]).sum() / P
        else:
            fw[s] = (zmin <= nes[s]).sum() / P      # NaN: no count
    return es[:, 0], nes, np_, fw, z


@pytest.mark.parametrize("dbl", [False, True])
def test_statistics_bruteforce_float64(orc, dbl):
    es = _stats_grid_case()
    dt = np.float64 if dbl else np.float32
    st, norm = orc.final_statistics(es.astype(dt))
    e0, nes, np_, fw, z = _stats_brute(es)
    assert bits_equal(st[0], e0.astype(dt)).all()
    fin = np.isfinite(nes)
    assert (ulps(st[1][fin], nes[fin], dt) <= 1.0).all()
    assert np.isnan(st[1][~fin]).all() and (~fin).sum() == 1
    assert (st[2] == np_).all(), (st[2], np_)
    assert (st[3] == fw).all(), (st[3], fw)
    zf = np.isfinite(z)
    assert (ulps(norm[zf], z[zf], dt) <= 1.0).all() and np.isnan(norm[~zf]).all()
    # the cases the rows were built for
    P = es.shape[1]
    assert np_[0] == 1 / P                 # p = 0 counts itself
    strict = int((np.abs(es[1]) > abs(es[1, 0])).sum())
    assert np_[1] * P == strict + 3        # itself and two ties counted (>=)
    assert np_[4] == 1.0 and fw[4] == 0.0  # zero row: NaN NES, no FWER count
    assert 0 < fw[2] < 1 and fw[5] > 1 / P


def _fdr_brute(norm, nes):
    G, P = norm.shape
    null = norm[:, 1:].ravel()
    null = null[~np.isnan(null)]
    obs = nes[~np.isnan(nes)]
    out = np.empty(G)
    for s in range(G):
        v = nes[s]
        if np.isnan(v):
            out[s] = np.nan
            continue
        if v > 0:
            dn = (null >= 0).sum()
            num = (null >= v).sum() / dn if dn else 0.0
            den = (obs >= v).sum() / (obs >= 0).sum()
        else:
            dn = (null <= 0).sum()
            num = (null <= v).sum() / dn if dn else 0.0
            den = (obs <= v).sum() / (obs <= 0).sum()
        out[s] = min(1.0, num / den)
    return out


def test_fdr_bruteforce(orc):
    rng = np.random.default_rng(9)
    es = (rng.standard_normal((40, 50)) * 0.3).astype(np.float32)
    es[:, 0] += np.linspace(-0.5, 0.5, 40).astype(np.float32)
    es[7, 3] = es[7, 0]                           # ties
    st, norm = orc.final_statistics(es)
    want = _fdr_brute(norm.astype(np.float64), st[1].astype(np.float64))
    assert np.allclose(st[4], want, rtol=1e-6, equal_nan=True)
    assert ((st[4] >= 0) & (st[4] <= 1)).all()


def test_statistics_invariants(orc):
    rng = np.random.default_rng(4)
    es = (rng.standard_normal((30, 200)) * 0.3).astype(np.float32)
    st, _ = orc.final_statistics(es)
    assert (st[2] >= 1 / 200).all()               # observed counts itself
    assert ((st[3] >= 0) & (st[3] <= 1)).all()
    assert np.allclose(st[0], es[:, 0])


# ------------------------------------------- signed NES (Broad GSEA) -------
# The option divides a score >= 0 by the mean of its set's null scores >= 0
# and a score < 0 by the mean magnitude of its null scores < 0 (Subramanian et
# al., PNAS 2005, "normalized enrichment score"), null = permutations 1..P-1.
@pytest.mark.parametrize("dbl", [False, True])
def test_signed_nes_known_answers(orc, dbl):
    dt = np.float64 if dbl else np.float32
    es = np.array([[0.5, 1.0, 2.0, -1.0, -3.0],     # null + {1, 2}, - {1, 3}
                   [-2.0, 1.0, 2.0, -1.0, -3.0],
                   [0.0, 1.0, 2.0, -1.0, -3.0],
                   [-1.0, 1.0, 2.0, 3.0, 0.0],      # no null score < 0
                   [4.0, -8.0, 2.0, -4.0, -4.0]],   # the observed is not in the mean
                  dtype=dt)
    st, norm = orc.final_statistics(es, signed_nes=True)
    nes = st[1].astype(np.float64)
    assert nes[0] == pytest.approx(0.5 / 1.5, rel=1e-7)
    assert nes[1] == -1.0
    assert nes[2] == 0.0
    assert np.isnan(nes[3])
    assert nes[4] == 2.0
    assert bits_equal(st[0], es[:, 0]).all()
    # NP is the reference's two-sided count whatever the normalisation
    ref, _ = orc.final_statistics(es)
    assert bits_equal(st[2], ref[2]).all()
    # the null scores of row 0 normalised per sign
    assert np.allclose(norm[0], [1 / 3, 2 / 3, 4 / 3, -0.5, -1.5], rtol=1e-7)


def test_signed_nes_null_means_are_one(orc):
    # the defining property: per set and sign, the normalised null scores
    # average to +1 / -1; the observed score (p = 0) is not in the mean
    rng = np.random.default_rng(21)
    es = rng.standard_normal((25, 300)) * 0.4 + rng.normal(0, 0.2, (25, 1))
    es[:, 0] += 3.0
    st, norm = orc.final_statistics(es, signed_nes=True)
    null = norm[:, 1:]
    for s in range(es.shape[0]):
        v = null[s]
        assert v[v >= 0].mean() == pytest.approx(1.0, rel=1e-12)
        assert v[v < 0].mean() == pytest.approx(-1.0, rel=1e-12)
    # and the observed row goes through the same factors
    assert np.allclose(st[1], norm[:, 0], rtol=0, atol=0)


def test_signed_nes_statistics_follow_the_normalised_scores(orc):
    # FWER and FDR of the signed option: the same definitions over the
    # per-sign normalised scores it returns (brute force in float64)
    rng = np.random.default_rng(5)
    es = (rng.standard_normal((40, 120)) * 0.3 + rng.normal(0, 0.1, (40, 1)))
    es[:, 0] += np.linspace(-0.6, 0.6, 40)
    st, norm = orc.final_statistics(es, signed_nes=True)
    nes = st[1]
    P = es.shape[1]
    zmax, zmin = norm.max(axis=0), norm.min(axis=0)
    fw = np.where(nes > 0, (zmax[None, :] >= nes[:, None]).sum(axis=1),
                  (zmin[None, :] <= nes[:, None]).sum(axis=1)) / P
    assert (st[3] == fw).all()
    assert np.allclose(st[4], _fdr_brute(norm, nes), rtol=1e-12, equal_nan=True)
    # the option changes the answers (a factor per sign, not one per set)
    ref, _ = orc.final_statistics(es)
    assert not np.allclose(ref[1], st[1])


@pytest.mark.parametrize("signed", [False, True])
def test_statistics_sign_and_scale_symmetry(orc, signed):
    # ES -> -ES swaps the two sides: NES changes sign, FWER / FDR / NP stay;
    # ES -> 4 ES (exact) changes nothing but ES
    rng = np.random.default_rng(8)
    es = (rng.standard_normal((30, 90)) * 0.3).astype(np.float32)
    es[:, 0] += np.linspace(-0.5, 0.5, 30).astype(np.float32)
    assert (es != 0).all()
    st, _ = orc.final_statistics(es, signed_nes=signed)
    neg, _ = orc.final_statistics(-es, signed_nes=signed)
    big, _ = orc.final_statistics(es * 4, signed_nes=signed)
    assert bits_equal(neg[1], -st[1]).all()
    for r in (2, 3, 4):
        assert bits_equal(neg[r], st[r]).all(), r
        assert bits_equal(big[r], st[r]).all(), r
    assert bits_equal(big[1], st[1]).all()


# --------------------------------------------------------- whole path ------
def test_gsea_composes_stages(orc):
    from cudagsea_amd import workloads
    w = workloads.make("c1", num_perms=64)
    stats, es = orc.gsea(w["X"], w["labels"], w["set_offsets"], w["set_genes"],
                         64, "stable_signal2noise")
    r = orc.correlate(w["X"], w["labels"], 64, "stable_signal2noise", sample_major=True)
    sk, idx = orc.pathway_sort(r)
    es2 = orc.scores(sk, idx, w["set_offsets"], w["set_genes"])
    assert bits_equal(es, es2).all()
    st2, _ = orc.final_statistics(es2)
    assert bits_equal(stats, st2).all()
