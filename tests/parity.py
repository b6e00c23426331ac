"""Per-element parity bars between the HIP path and the CPU oracle.

ES: within half the tie window (below) absolute, 2^-21 in fp32 and 2^-33 in
fp64: the GPU walk's drift from the reference's (bounded in tie_window) plus
a one-ulp difference of N_R and the Kahan error of the reference's sums in T
(an ulp or two) stay below it.  NES and the ES-derived statistics: within
1e-5 relative (|d| <= 1e-5 |x| + 1e-6), fp32 and fp64.
The one allowed exception is a sign flip where the oracle's walk has two
candidates of nearly equal magnitude: its largest positive running sum `hi`
and its most negative one `lo` (oracle.scores_ext).  The reference keeps the
one of larger magnitude (the first of two equal ones); when ||hi| - |lo|| is
within tie_window (2^-20 in fp32, 2^-32 in fp64, derived there) the two
implementations' roundings may pick either, so the GPU ES must then be
within the bar of the other candidate.

NP and FWER are counts over permutations: exact, except for comparisons of two
values that lie within one tie window (NP, on ES) or within the NES bar
(FWER, on normalised scores) of each other, plus, for FWER, the null scores
that may flip as above.

FDR: the counts it is made of (null scores beyond NES, null scores on NES's
side of 0, observed NES beyond NES, observed NES on its side of 0) are
recomputed from the oracle's ES with every comparison inside the band counted
both ways; the GPU FDR must lie between the resulting bounds.
"""
import numpy as np

ES_RTOL = 1e-5
ES_ATOL = 1e-6


def bits_equal(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    u = np.uint64 if a.dtype == np.float64 else np.uint32
    return (a.view(u) == b.view(u)) | (np.isnan(a) & np.isnan(b))


def close(a, b, rtol=ES_RTOL, atol=ES_ATOL):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ok = np.abs(a - b) <= rtol * np.abs(b) + atol
    return ok | (np.isnan(a) & np.isnan(b))




def tie_window(dtype):
    """The largest gap between the two candidates' magnitudes at which the
    GPU's walk may pick the other one.  It carries c = |r|/N_R unrounded in
    double; the reference rounds c to T and sums in T with Kahan.  Their
    walks differ by at most sum |c - rounded c| <= 2^-24 in fp32 (the c of a
    set sum to 1), plus the double roundings (< 2^-35) and an ulp or two of
    T in each rounded candidate (magnitudes <= 1): below 2^-21 in all.  In
    fp64 the c differ by at most 2.5 ulp and the walks by < 2^-35.  The
    windows are twice those bounds or more."""
    return 2.0 ** -20 if np.dtype(dtype) == np.float32 else 2.0 ** -32


def flippable(hi, lo):
    """Walks whose two candidates have magnitudes within tie_window: the
    scores whose sign may differ from the oracle's."""
    win = tie_window(np.asarray(hi).dtype)
    hi = np.asarray(hi, dtype=np.float64)
    lo = np.asarray(lo, dtype=np.float64)
    m = np.maximum(np.abs(hi), np.abs(lo))
    return (np.abs(np.abs(hi) - np.abs(lo)) <= win) & (m > 0)


def es_check(got, want, hi, lo):
    """(ok, flipped): ok per element; flipped marks the elements that passed
    as a justified sign flip: the oracle's two candidates lie within
    tie_window of each other in magnitude, and the GPU value is within the
    ES bar of the other candidate."""
    win = tie_window(np.asarray(hi).dtype)
    h = np.asarray(hi, dtype=np.float64)
    l = np.asarray(lo, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    alt = np.where(want == h, l, h)
    direct = close(got, want, rtol=0.0, atol=win / 2)
    flip = ~direct & flippable(hi, lo) & close(got, alt, rtol=0.0, atol=win / 2)
    return direct | flip, flip


def assert_es(got, want, hi, lo, what="ES"):
    ok, flip = es_check(got, want, hi, lo)
    bad = np.argwhere(~ok)
    assert ok.all(), (what, int((~ok).sum()), bad[:5].tolist(),
                      [(float(np.asarray(got)[tuple(i)]), float(np.asarray(want)[tuple(i)]))
                       for i in bad[:3]])
    return flip


def _normalised(es):
    """statistics.cuh:61-90: factor = P / sum_p |ES_p|, z = ES * factor."""
    e = np.asarray(es, dtype=np.float64)
    P = e.shape[1]
    with np.errstate(divide="ignore", invalid="ignore"):
        fac = P / np.abs(e).sum(axis=1)
        z = e * fac[:, None]
    return fac, z


def _band(x):
    # a normalised score carries the ES bar once through ES and once through
    # the factor (a sum of terms each within the bar)
    return 2.0 * (ES_RTOL * np.abs(x) + ES_ATOL) + 1e-7


def _count_ge(sorted_vals, t):
    return sorted_vals.size - np.searchsorted(sorted_vals, t, side="left")


def _count_gt(sorted_vals, t):
    return sorted_vals.size - np.searchsorted(sorted_vals, t, side="right")


def _count_le(sorted_vals, t):
    return np.searchsorted(sorted_vals, t, side="right")


def _count_lt(sorted_vals, t):
    return np.searchsorted(sorted_vals, t, side="left")


def stats_bounds(es, hi, lo):
    """Bounds implied by the oracle's ES [G][P] for NP, FWER and FDR counts.
    Returns dict of per-set arrays np_lo, np_hi, fw_lo, fw_hi, fdr_lo, fdr_hi
    (NP and FWER as counts), and `flip_obs`, the sets whose observed ES may
    flip sign."""
    e = np.asarray(es, dtype=np.float64)
    G, P = e.shape
    fl = flippable(hi, lo)
    alt = np.where(e == np.asarray(hi, dtype=np.float64), lo, hi).astype(np.float64)
    fac, z = _normalised(e)
    z_alt = alt * fac[:, None]
    nes = z[:, 0]

    # NP: |ES_p| >= |ES_0| (p = 0 included); each GPU ES is within half the
    # tie window of the oracle's, so comparisons within one window may go
    # either way
    a = np.abs(e)
    b = tie_window(np.asarray(hi).dtype)
    np_lo = (a > a[:, :1] + b).sum(axis=1) + 1
    np_hi = (a >= a[:, :1] - b).sum(axis=1)

    # FWER: per-permutation extremes over sets; a flippable score may take
    # either value
    zz = np.where(np.isnan(z), np.nan, z)
    zmin_c = np.where(fl, np.fmin(zz, z_alt), zz)
    zmax_c = np.where(fl, np.fmax(zz, z_alt), zz)
    with np.errstate(invalid="ignore"):
        mx_lo = np.nanmax(np.where(np.isnan(zmin_c), -np.inf, zmin_c), axis=0)
        mx_hi = np.nanmax(np.where(np.isnan(zmax_c), -np.inf, zmax_c), axis=0)
        mn_lo = np.nanmin(np.where(np.isnan(zmin_c), np.inf, zmin_c), axis=0)
        mn_hi = np.nanmin(np.where(np.isnan(zmax_c), np.inf, zmax_c), axis=0)
    fw_lo = np.zeros(G)
    fw_hi = np.zeros(G)
    for s in range(G):
        v = nes[s]
        if np.isnan(v):
            continue
        bv = _band(v)
        if v > bv:
            fw_lo[s] = (mx_lo > v + bv).sum()
            fw_hi[s] = (mx_hi >= v - bv).sum()
        elif v < -bv:
            fw_lo[s] = (mn_hi < v - bv).sum()
            fw_hi[s] = (mn_lo <= v + bv).sum()
        else:
            # NES near 0: either branch of the reference
            fw_lo[s] = min((mx_lo > v + bv).sum(), (mn_hi < v - bv).sum())
            fw_hi[s] = max((mx_hi >= v - bv).sum(), (mn_lo <= v + bv).sum())

    # FDR: null = normalised scores of permutations 1..P-1, obs = NES
    null = z[:, 1:][~np.isnan(z[:, 1:])]
    null_alt = z_alt[:, 1:][fl[:, 1:] & ~np.isnan(z_alt[:, 1:])]
    nflip = null_alt.size
    null_s = np.sort(null)
    obs_s = np.sort(nes[~np.isnan(nes)])
    fdr_lo = np.full(G, np.nan)
    fdr_hi = np.full(G, np.nan)

    def ratio(a_, dn, b_, dobs):
        num = np.where(dn > 0, a_ / np.maximum(dn, 1), 0.0)
        den = b_ / np.maximum(dobs, 1)
        with np.errstate(divide="ignore", invalid="ignore"):
            q = np.where(den > 0, num / den, np.inf)
        return np.minimum(1.0, q)

    b0 = 1e-7
    for s in range(G):
        v = nes[s]
        if np.isnan(v):
            continue
        bv = _band(v)
        if v > 0:
            a_lo = _count_gt(null_s, v + bv) - nflip
            a_hi = _count_ge(null_s, v - bv) + nflip
            dn_lo = _count_gt(null_s, b0) - nflip
            dn_hi = _count_ge(null_s, -b0) + nflip
            bb_lo = _count_gt(obs_s, v + bv) + 1      # the set's own NES counts
            bb_hi = _count_ge(obs_s, v - bv)
            do_lo = _count_gt(obs_s, b0)
            do_hi = _count_ge(obs_s, -b0)
        else:
            a_lo = _count_lt(null_s, v - bv) - nflip
            a_hi = _count_le(null_s, v + bv) + nflip
            dn_lo = _count_lt(null_s, -b0) - nflip
            dn_hi = _count_le(null_s, b0) + nflip
            bb_lo = _count_lt(obs_s, v - bv) + 1
            bb_hi = _count_le(obs_s, v + bv)
            do_lo = _count_lt(obs_s, -b0)
            do_hi = _count_le(obs_s, b0)
        a_lo, dn_lo, do_lo = max(a_lo, 0), max(dn_lo, 0), max(do_lo, 1)
        fdr_lo[s] = ratio(np.float64(a_lo), np.float64(dn_hi), np.float64(bb_hi),
                          np.float64(do_lo))
        fdr_hi[s] = ratio(np.float64(a_hi), np.float64(max(dn_lo, 1) if a_hi else dn_lo),
                          np.float64(bb_lo), np.float64(do_hi))
    return {"np_lo": np_lo, "np_hi": np_hi, "fw_lo": fw_lo, "fw_hi": fw_hi,
            "fdr_lo": fdr_lo, "fdr_hi": fdr_hi, "flip_obs": fl[:, 0],
            "nes_alt": z_alt[:, 0]}


def assert_stats(got, want, es, hi, lo):
    """got / want [5][G] (ES, NES, NP, FWER, FDR); es, hi, lo [G][P] from
    oracle.scores_ext / gsea_ext."""
    G, P = np.asarray(es).shape
    bd = stats_bounds(es, hi, lo)
    fo = bd["flip_obs"]
    # ES and NES: direct, or the justified flip of the observed walk
    ok_es = close(got[0], want[0]) | (fo & close(got[0], np.where(
        np.asarray(es)[:, 0] == np.asarray(hi)[:, 0], np.asarray(lo)[:, 0],
        np.asarray(hi)[:, 0])))
    ok_nes = close(got[1], want[1]) | (fo & close(got[1], bd["nes_alt"]))
    assert ok_es.all(), ("ES", np.argwhere(~ok_es)[:5].tolist())
    assert ok_nes.all(), ("NES", np.argwhere(~ok_nes)[:5].tolist())
    # counts: a flipped observed walk leaves |ES| (NP) within the bar; its own
    # FWER / FDR follow the other sign and are not compared
    c_np = np.rint(np.asarray(got[2], dtype=np.float64) * P)
    c_fw = np.rint(np.asarray(got[3], dtype=np.float64) * P)
    ok_np = (c_np >= bd["np_lo"]) & (c_np <= bd["np_hi"])
    ok_fw = ((c_fw >= bd["fw_lo"]) & (c_fw <= bd["fw_hi"])) | fo | np.isnan(want[1])
    assert ok_np.all(), ("NP", np.argwhere(~ok_np)[:5].tolist())
    assert ok_fw.all(), ("FWER", np.argwhere(~ok_fw)[:5].tolist(),
                         c_fw[~ok_fw][:5], bd["fw_lo"][~ok_fw][:5], bd["fw_hi"][~ok_fw][:5])
    f = np.asarray(got[4], dtype=np.float64)
    ok_fdr = ((f >= bd["fdr_lo"] * (1 - 1e-6) - 1e-12) & (f <= bd["fdr_hi"] * (1 + 1e-6) + 1e-12))
    ok_fdr |= fo | (np.isnan(f) & np.isnan(np.asarray(want[4], dtype=np.float64)))
    assert ok_fdr.all(), ("FDR", np.argwhere(~ok_fdr)[:5].tolist(), f[~ok_fdr][:5],
                          bd["fdr_lo"][~ok_fdr][:5], bd["fdr_hi"][~ok_fdr][:5])
    return int(fo.sum())
