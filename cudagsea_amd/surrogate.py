"""Learned permutation allocation.

A 16 -> 64 -> 64 -> 4 MLP (GELU) predicts, for every gene set, the moments of
its null enrichment scores (mean and std of the positive null ES, mean and std
of |negative null ES|) from 16 features of the observed ranking.  From those
moments and the observed ES the allocation derives each set's NES precision at
a given number of permutations and gives the set the smallest budget (a
divisor of the permutation count) at which that precision meets the target.
The engine's budgeted run then scores every set for its budget, checks the
measured statistics and scores for all permutations every set whose FDR lies
near the threshold or whose measured NES precision misses the target.  The
plain full run stays the reference the allocation is checked against.

Everything numerical runs in the HIP library: features (gsea_set_features),
training targets (gsea_null_moments), the network's forward/backward passes and
Adam steps (gsea_mlp_*), bf16 inference and the budgeted run.  torch only
holds device memory here.
"""
import ctypes
import math
import os

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .gsea import (_dp, argsort_device, metric_device, run_budget_device,
                   run_device)

NUM_FEATURES = 16
HIDDEN = 64
NUM_OUT = 4
NUM_PARAMS = HIDDEN * NUM_FEATURES + HIDDEN + HIDDEN * HIDDEN + HIDDEN + \
    NUM_OUT * HIDDEN + NUM_OUT
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                          "surrogate_model.npz")
# smallest budget the allocation hands out
MIN_BUDGET = 512


# ---------------------------------------------------------------------------
# kernel wrappers
# ---------------------------------------------------------------------------
def _st(stream):
    return ctypes.c_void_p(stream) if stream else None


def observed_ranking(X, labels, metric, ascending=False, swap_labels=False):
    """Observed metric r0[N] (float32) and the gene at each rank [N] (int32),
    both device tensors."""
    S, N = X.shape
    r0 = torch.empty((1, N), dtype=X.dtype, device=X.device)
    metric_device(X, labels, 1, metric, r0, swap_labels=swap_labels)
    idx = torch.empty((1, N), dtype=torch.int32, device=X.device)
    argsort_device(r0, idx, ascending=ascending)
    return r0[0].float().contiguous(), idx[0].contiguous()


def set_features(r0, rank_genes, set_offsets, set_genes, stream=None):
    """features[G][16] (device float32) of every set."""
    N = r0.numel()
    G = set_offsets.numel() - 1
    feat = torch.empty((G, NUM_FEATURES), dtype=torch.float32, device=r0.device)
    glob = torch.empty(4, dtype=torch.float32, device=r0.device)
    check(lib().gsea_set_features(_dp(r0), _dp(rank_genes), N, _dp(set_offsets),
                                  _dp(set_genes), G, _dp(glob), _dp(feat), _st(stream)))
    return feat


def null_moments(es_pg, stream=None):
    """[G][5] (device float32): mean/std of the positive null ES, mean/std of
    |negative null ES|, fraction positive, over permutations 1..P-1 of
    es_pg[P][G] (float32)."""
    P, G = es_pg.shape
    out = torch.empty((G, 5), dtype=torch.float32, device=es_pg.device)
    check(lib().gsea_null_moments(_dp(es_pg), P, G, _dp(out), _st(stream)))
    return out


def mlp_forward(params, x, bf16=False, norm=None, stream=None):
    """y[B][4] = mlp(x); norm: optional device [40] standardisation
    (x_mean, x_inv_std, y_scale, y_shift) applied inside the kernel."""
    B = x.shape[0]
    y = torch.empty((B, NUM_OUT), dtype=torch.float32, device=x.device)
    check(lib().gsea_mlp_forward(_dp(params), _dp(x), B, 1 if bf16 else 0, _dp(norm),
                                 _dp(y), _st(stream)))
    return y


def mlp_grad(params, x, t, stream=None):
    """(gradient of mean squared error [NUM_PARAMS] device, loss float)."""
    B = x.shape[0]
    g = torch.empty(NUM_PARAMS, dtype=torch.float32, device=x.device)
    loss = ctypes.c_float(0.0)
    check(lib().gsea_mlp_grad(_dp(params), _dp(x), _dp(t), B, _dp(g),
                              ctypes.byref(loss), _st(stream)))
    return g, float(loss.value)


def mlp_adam(params, m, v, grad, step, lr, weight_decay=0.0, stream=None):
    check(lib().gsea_mlp_adam(_dp(params), _dp(m), _dp(v), _dp(grad), int(step),
                              ctypes.c_float(lr), ctypes.c_float(weight_decay),
                              _st(stream)))


def init_params(seed=0):
    """He (fan-in) normal weights, zero biases, packed in the C ABI's order."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_out, fan_in in ((HIDDEN, NUM_FEATURES), (HIDDEN, HIDDEN), (NUM_OUT, HIDDEN)):
        parts.append(rng.standard_normal(fan_out * fan_in) * math.sqrt(2.0 / fan_in))
        parts.append(np.zeros(fan_out))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == NUM_PARAMS
    return p


# ---------------------------------------------------------------------------
# targets: log null moments, standardised
# ---------------------------------------------------------------------------
_FLOOR = 1e-6


def moments_to_targets(mom):
    """[G][5] null moments -> [G][4] log(mean+, std+, mean-, std-)."""
    m = np.asarray(mom, dtype=np.float64)[:, :4]
    return np.log(np.maximum(m, _FLOOR))


def targets_to_moments(t):
    return np.exp(np.asarray(t, dtype=np.float64))


class Surrogate:
    """Trained network plus its feature / target standardisation."""

    def __init__(self, params, feat_mean, feat_std, tgt_mean, tgt_std, meta=None):
        self.params = np.asarray(params, dtype=np.float32)
        self.feat_mean = np.asarray(feat_mean, dtype=np.float32)
        self.feat_std = np.asarray(feat_std, dtype=np.float32)
        self.tgt_mean = np.asarray(tgt_mean, dtype=np.float64)
        self.tgt_std = np.asarray(tgt_std, dtype=np.float64)
        self.meta = dict(meta or {})
        self._dev = None

    def device_params(self, device="cuda"):
        """(parameters, standardisation [40]) as device tensors."""
        if self._dev is None or self._dev[0].device != torch.device(device):
            norm = np.concatenate([self.feat_mean, 1.0 / self.feat_std,
                                   self.tgt_std, self.tgt_mean]).astype(np.float32)
            self._dev = (torch.from_numpy(self.params).to(device),
                         torch.from_numpy(norm).to(device))
        return self._dev

    def save(self, path=MODEL_PATH):
        meta = {k: np.asarray(v) for k, v in self.meta.items()}
        np.savez(path, params=self.params, feat_mean=self.feat_mean,
                 feat_std=self.feat_std, tgt_mean=self.tgt_mean,
                 tgt_std=self.tgt_std, **{"meta_" + k: v for k, v in meta.items()})

    @classmethod
    def load(cls, path=MODEL_PATH):
        z = np.load(path, allow_pickle=False)
        meta = {k[5:]: z[k] for k in z.files if k.startswith("meta_")}
        return cls(z["params"], z["feat_mean"], z["feat_std"], z["tgt_mean"],
                   z["tgt_std"], meta)

    def predict(self, feat, bf16=True):
        """Null moments [G][4] (numpy float64) from device features [G][16]."""
        params, norm = self.device_params(feat.device)
        y = mlp_forward(params, feat.contiguous(), bf16, norm)
        return targets_to_moments(y.cpu().numpy().astype(np.float64))


def train(features, targets, epochs=8, batch=4096, lr=3e-3, weight_decay=0.0,
          seed=0, val_frac=0.05, device="cuda", log=None):
    """Fit the network on (features [M][16], log-moment targets [M][4]) with
    the library's forward/backward and Adam kernels.  Returns the Surrogate
    and a dict of training / validation losses."""
    feats = np.asarray(features, dtype=np.float32)
    tg = np.asarray(targets, dtype=np.float64)
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(feats))
    nval = max(1, int(len(feats) * val_frac))
    val, tr = order[:nval], order[nval:]
    fmean = feats[tr].mean(0)
    fstd = feats[tr].std(0)
    fstd = np.where(fstd > 1e-8, fstd, 1.0).astype(np.float32)
    tmean = tg[tr].mean(0)
    tstd = tg[tr].std(0)
    tstd = np.where(tstd > 1e-8, tstd, 1.0)
    xs = torch.from_numpy(((feats - fmean) / fstd).astype(np.float32)).to(device)
    ts = torch.from_numpy(((tg - tmean) / tstd).astype(np.float32)).to(device)
    tr_d = torch.from_numpy(tr).to(device)
    xv, tv = xs[torch.from_numpy(val).to(device)].contiguous(), \
        ts[torch.from_numpy(val).to(device)].contiguous()
    params = torch.from_numpy(init_params(seed)).to(device)
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    step = 0
    hist = {"train": [], "val": []}
    steps_per_epoch = max(1, len(tr) // batch)
    total = epochs * steps_per_epoch
    for ep in range(epochs):
        perm = tr_d[torch.from_numpy(rng.permutation(len(tr))).to(device)]
        run = 0.0
        for i in range(steps_per_epoch):
            idx = perm[i * batch:(i + 1) * batch]
            xb, tb = xs[idx].contiguous(), ts[idx].contiguous()
            g, loss = mlp_grad(params, xb, tb)
            step += 1
            # cosine decay of the step size
            lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
            mlp_adam(params, m, v, g, step, lr_t, weight_decay)
            run += loss
        _, vloss = mlp_grad(params, xv, tv)
        hist["train"].append(run / steps_per_epoch)
        hist["val"].append(vloss)
        if log:
            log("epoch %d train %.4f val %.4f" % (ep, hist["train"][-1], vloss))
    model = Surrogate(params.cpu().numpy(), fmean, fstd, tmean, tstd,
                      {"epochs": epochs, "samples": len(tr)})
    return model, hist


def size_baseline(features, targets, val_idx=None):
    """The method without the network: log-moments as a quadratic function of
    log set size alone (least squares).  Returns coefficients [3][4]."""
    f = np.asarray(features, dtype=np.float64)
    x = f[:, 0]                      # log(k) / log(N)
    A = np.stack([np.ones_like(x), x, x * x], 1)
    coef, *_ = np.linalg.lstsq(A, np.asarray(targets, dtype=np.float64), rcond=None)
    return coef


def size_baseline_predict(coef, features):
    x = np.asarray(features, dtype=np.float64)[:, 0]
    A = np.stack([np.ones_like(x), x, x * x], 1)
    return targets_to_moments(A @ coef)


# ---------------------------------------------------------------------------
# allocation
# ---------------------------------------------------------------------------
def null_abs_stats(mom):
    """mean and coefficient of variation of |null ES| from [G][4] moments
    (positive and negative tails weighted equally)."""
    mom = np.asarray(mom, dtype=np.float64)
    mp, sp, mn, sn = mom[:, 0], mom[:, 1], mom[:, 2], mom[:, 3]
    mean = 0.5 * (mp + mn)
    m2 = 0.5 * (sp * sp + mp * mp + sn * sn + mn * mn)
    cv = np.sqrt(np.maximum(m2 - mean * mean, 0.0)) / np.maximum(mean, _FLOOR)
    return mean, cv


def budgets(mom, es0, num_perms, nes_tol=0.02, z=4.5, margin=0.15,
            min_budget=MIN_BUDGET):
    """Per-set permutation budget: the smallest num_perms / 2^i (>= min_budget)
    at which z * cv * |NES| / sqrt(budget) <= nes_tol, with |NES| and cv from
    the predicted moments inflated by `margin`."""
    mean, cv = null_abs_stats(mom)
    nes = np.abs(np.asarray(es0, dtype=np.float64)) / np.maximum(mean, _FLOOR)
    need = (z * cv * (1.0 + margin) * nes * (1.0 + margin) / nes_tol) ** 2
    need = np.where(np.isfinite(need), need, float(num_perms))
    levels = []
    b = int(num_perms)
    while b >= max(2, min_budget):
        levels.append(b)
        if b % 2:
            break
        b //= 2
    levels = np.array(sorted(levels))
    pick = np.searchsorted(levels, need, side="left")
    pick = np.minimum(pick, len(levels) - 1)
    return levels[pick].astype(np.int32)


def adaptive_run(X, labels, set_offsets, set_genes, num_perms, metric, model,
                 nes_tol=0.02, z=4.5, margin=0.15, fdr=0.25,
                 check_nes=True, min_budget=MIN_BUDGET, stats_out=None,
                 ascending=False, swap_labels=False, predictor=None):
    """Surrogate-guided run.  Returns (stats [5][G] device, final budgets,
    initial budgets, timing dict in ms).  predictor(features) -> moments
    replaces the network (the allocation without it)."""
    G = set_offsets.numel() - 1
    dev = X.device
    torch.cuda.synchronize(dev)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    r0, rank_genes = observed_ranking(X, labels, metric, ascending, swap_labels)
    feat = set_features(r0, rank_genes, set_offsets, set_genes)
    if predictor is None:
        mom = model.predict(feat, bf16=True)
    else:
        mom = predictor(feat.cpu().numpy())
    # the observed ES of every set: one budgeted-free pass over permutation 0
    es0 = observed_es(r0.to(X.dtype).view(1, -1), set_offsets, set_genes, ascending)
    b0 = budgets(mom, es0, num_perms, nes_tol, z, margin, min_budget)
    t1.record()
    torch.cuda.synchronize(dev)
    ms_alloc = t0.elapsed_time(t1)
    if stats_out is None:
        stats_out = torch.empty((5, G), dtype=X.dtype, device=dev)
    ext = (fdr, z, nes_tol if check_nes else 0.0)
    bfin, tm = run_budget_device(X, labels, set_offsets, set_genes, num_perms, metric,
                                 b0, stats_out, extension=ext, ascending=ascending,
                                 swap_labels=swap_labels)
    tm = dict(tm)
    tm["ms_alloc"] = ms_alloc
    return stats_out, bfin, b0, tm


def observed_es(r0, set_offsets, set_genes, ascending=False):
    """ES of every set for the observed ranking (host float64 [G])."""
    from .gsea import rank_es_device
    G = set_offsets.numel() - 1
    es = torch.empty((1, G), dtype=r0.dtype, device=r0.device)
    rank_es_device(r0.contiguous(), set_offsets, set_genes, es, ascending=ascending)
    return es[0].double().cpu().numpy()


# ---------------------------------------------------------------------------
# evaluation against the plain full run
# ---------------------------------------------------------------------------
def compare(stats_plain, stats_adapt, fdr=0.25):
    """Significant-set calls (FDR < fdr) and NES differences of two runs
    (host [5][G] arrays)."""
    sp = np.asarray(stats_plain, dtype=np.float64)
    sa = np.asarray(stats_adapt, dtype=np.float64)
    cp = sp[4] < fdr
    ca = sa[4] < fdr
    d = np.abs(sp[1] - sa[1])
    d = d[np.isfinite(d)]
    return {"calls_plain": int(cp.sum()), "calls_adaptive": int(ca.sum()),
            "calls_identical": bool(np.array_equal(cp, ca)),
            "calls_differing": int((cp != ca).sum()),
            "max_abs_dnes": float(d.max()) if d.size else 0.0}


def evaluate(w, model, num_perms=None, metric=None, repeats=2, predictor=None,
             plain=None, **kw):
    """Plain full run vs surrogate-guided run on workload dict `w` (one
    metric).  Times are the best of `repeats` after one warm-up, each
    bracketed by device synchronisation.  Returns a dict (and the plain
    run's host statistics and time, reusable as `plain`)."""
    import time
    P = int(num_perms or w["P"])
    metric = metric or w["metrics"][0]
    X = torch.from_numpy(w["X"]).cuda()
    off = torch.from_numpy(w["set_offsets"]).cuda()
    genes = torch.from_numpy(w["set_genes"]).cuda()
    G = w["G"]
    stats = torch.empty((5, G), dtype=X.dtype, device=X.device)

    def timed(fn):
        best = None
        out = None
        for i in range(repeats + 1):
            torch.cuda.synchronize()
            t = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            if i > 0:
                best = dt if best is None else min(best, dt)
        return out, best

    if plain is None:
        _, t_plain = timed(lambda: run_device(X, w["labels"], off, genes, P, metric, stats))
        sp = stats.cpu().numpy()
    else:
        sp, t_plain = plain
    res, t_ad = timed(lambda: adaptive_run(X, w["labels"], off, genes, P, metric, model,
                                           predictor=predictor, **kw))
    sa, bfin, b0, tm = res
    out = compare(sp, sa.cpu().numpy())
    work = float(np.sum(bfin.astype(np.float64))) / (G * P)
    out.update({"metric": metric, "num_perms": P, "num_sets": G,
                "s_plain": t_plain, "s_adaptive": t_ad, "speedup": t_plain / t_ad,
                "es_work_fraction": work,
                "es_work_fraction_initial": float(np.sum(b0.astype(np.float64))) / (G * P),
                "sets_extended": int((bfin != b0).sum()),
                "budget_hist": {int(k): int(v) for k, v in
                                zip(*np.unique(bfin, return_counts=True))}})
    return out, (sp, t_plain)


# ---------------------------------------------------------------------------
# training data: c3-shaped datasets with fresh seeds
# ---------------------------------------------------------------------------
def make_training_data(num_datasets, perms=1024, seed=1000, metrics=None,
                       num_sets=None, log=None, device="cuda"):
    """(features [M][16], targets [M][4]) from `num_datasets` c3-shaped
    datasets (seeds seed, seed+1, ...), each scored with every metric in
    `metrics` for `perms` permutations by the engine itself."""
    from . import workloads
    from .gsea import METRICS
    metrics = list(metrics or ["stable_signal2noise", "onepass_t_test"])
    feats, tgts = [], []
    for i in range(num_datasets):
        w = workloads.make("c3", num_sets=num_sets, seed=seed + i)
        X = torch.from_numpy(w["X"]).to(device)
        off = torch.from_numpy(w["set_offsets"]).to(device)
        genes = torch.from_numpy(w["set_genes"]).to(device)
        G = w["G"]
        for metric in metrics:
            assert metric in METRICS
            r0, rg = observed_ranking(X, w["labels"], metric)
            f = set_features(r0, rg, off, genes)
            stats = torch.empty((5, G), dtype=torch.float32, device=device)
            es = torch.empty((G, perms), dtype=torch.float32, device=device)
            run_device(X, w["labels"], off, genes, perms, metric, stats, es_out=es)
            mom = null_moments(es.t().contiguous())
            feats.append(f.cpu().numpy())
            tgts.append(moments_to_targets(mom.cpu().numpy()))
        if log:
            log("dataset %d/%d" % (i + 1, num_datasets))
    return np.concatenate(feats), np.concatenate(tgts)


# ---------------------------------------------------------------------------
# python -m cudagsea_amd.surrogate train [--datasets N] [--out PATH]
# ---------------------------------------------------------------------------
TRAIN_METRICS = ["stable_signal2noise", "onepass_t_test", "naive_diff_of_classes",
                 "stable_log2_ratio_of_classes"]


def main(argv=None):
    import argparse
    import time
    ap = argparse.ArgumentParser(prog="python -m cudagsea_amd.surrogate")
    ap.add_argument("command", choices=["train"])
    ap.add_argument("--datasets", type=int, default=50)
    ap.add_argument("--perms", type=int, default=1024)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=10_000)
    ap.add_argument("--out", default=MODEL_PATH)
    a = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    t0 = time.time()
    feats, tg = make_training_data(a.datasets, perms=a.perms, seed=a.seed,
                                   metrics=TRAIN_METRICS, log=log)
    log("training samples: %d (%.1f s)" % (len(feats), time.time() - t0))
    t1 = time.time()
    model, hist = train(feats, tg, epochs=a.epochs, batch=a.batch, lr=a.lr,
                        seed=a.seed, log=log)
    log("trained in %.1f s" % (time.time() - t1))
    coef = size_baseline(feats, tg)
    model.meta.update({"size_baseline": coef, "datasets": a.datasets,
                       "perms": a.perms, "val_loss": hist["val"][-1],
                       "train_seed": a.seed})
    model.save(a.out)
    log("saved %s" % a.out)


if __name__ == "__main__":
    main()
