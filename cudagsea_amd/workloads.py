"""Seeded synthetic workloads of the benchmark configurations.

Each returns a dict with X (sample-major [S][N]), labels (S, 0/1), CSR gene
sets (set_offsets, set_genes), gene symbols "G%05d", set names, the number of
permutations and the metric list.  numpy PCG64 with seed 20261203 + config id.

  c1: 2,000 genes x 40 samples, 20/20, 50 sets of 15-200, 1,024 permutations
  c2: 20,000 x 200, 100/100, 1,000 sets of 15-500, 4,096 permutations
  c3: 20,000 x 500, 150/350, 10,000 sets log-uniform in [5, 2000], 16,384
  c4: 20,000 x 300, 150/150, N(1e4, 1e-2) per gene, 1% constant genes,
      2,000 sets of 15-500, 4,096 permutations
  c5: c3's shape with its own seed, 400 shifted genes and 3% of the sets
      carrying a planted 2-30% share of them (so that FDR < 0.25 calls
      exist), 16,384 permutations, stable_signal2noise
"""
import numpy as np

SEED = 20261203
ALL_METRICS = [
    "naive_diff_of_classes", "naive_ratio_of_classes",
    "naive_log2_ratio_of_classes", "stable_diff_of_classes",
    "stable_ratio_of_classes", "stable_log2_ratio_of_classes",
    "onepass_signal2noise", "onepass_t_test",
    "twopass_signal2noise", "twopass_t_test",
    "stable_signal2noise", "stable_t_test",
    "overkill_signal2noise", "overkill_t_test",
]


def _labels(rng, nA, nB):
    lab = np.array([0] * nA + [1] * nB, dtype=np.uint8)
    rng.shuffle(lab)
    return lab


def _csr(sets):
    off = np.zeros(len(sets) + 1, dtype=np.int32)
    off[1:] = np.cumsum([len(s) for s in sets])
    genes = (np.concatenate(sets).astype(np.int32) if sets
             else np.zeros(0, dtype=np.int32))
    return off, genes


def _sets_uniform(rng, G, N, lo, hi):
    return [np.sort(rng.choice(N, size=int(rng.integers(lo, hi + 1)),
                               replace=False)) for _ in range(G)]


def _sets_loguniform(rng, G, N, lo, hi):
    sizes = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=G)))
    sizes = np.clip(sizes.astype(np.int64), lo, min(hi, N))
    return [np.sort(rng.choice(N, size=int(k), replace=False)) for k in sizes]


def _lognormal(rng, N, S, lab, shifted, shift):
    z = rng.standard_normal((S, N))
    logx = 5.0 + z
    b = lab == 1
    logx[np.ix_(b, shifted)] += shift
    return np.exp(logx)


def make(config, dtype=np.float32, num_genes=None, num_samples=None,
         num_sets=None, num_perms=None, seed=None):
    """Build one workload; the optional sizes shrink it for quick checks.
    c5 is c3's shape drawn with its own seed (the dataset the learned
    permutation allocation is evaluated on); `seed` replaces the default
    seed (fresh c3-shaped datasets for training the allocation model)."""
    cid = {"c1": 1, "c2": 2, "c3": 3, "c4": 4, "c5": 5}[config]
    rng = np.random.default_rng(SEED + cid if seed is None else seed)
    if config == "c5":
        config = "c3"
        c5 = True
    else:
        c5 = False
    if config == "c1":
        N, S, nA, G, P = 2000, 40, 20, 50, 1024
        metrics = ["stable_signal2noise"]
    elif config == "c2":
        N, S, nA, G, P = 20000, 200, 100, 1000, 4096
        metrics = ["onepass_t_test"]
    elif config == "c3":
        N, S, nA, G, P = 20000, 500, 150, 10000, 16384
        metrics = list(ALL_METRICS)
    else:
        N, S, nA, G, P = 20000, 300, 150, 2000, 4096
        metrics = ["twopass_signal2noise", "stable_t_test", "overkill_t_test"]
    if num_genes:
        N = num_genes
    if num_samples:
        nA = max(1, int(round(nA * num_samples / S)))
        S = num_samples
    if num_sets:
        G = num_sets
    if num_perms:
        P = num_perms
    nB = S - nA
    lab = _labels(rng, nA, nB)
    if config in ("c1", "c3"):
        nshift = min(100 if not c5 else 400, N)
        shifted = rng.choice(N, size=nshift, replace=False)
        X = _lognormal(rng, N, S, lab, shifted, 0.5)
    elif config == "c2":
        shifted = rng.choice(N, size=max(1, N // 20), replace=False)
        X = np.exp(5.0 + rng.standard_normal((S, N)))
        b = lab == 1
        X[np.ix_(b, shifted)] *= np.exp(rng.normal(0.8, 0.2, size=shifted.size))
    else:
        X = 1e4 + 1e-2 * rng.standard_normal((S, N))
        const = rng.choice(N, size=max(1, N // 100), replace=False)
        X[:, const] = 1e4
    hi = {"c1": 200, "c2": 500, "c4": 500}.get(config, 2000)
    hi = min(hi, N)
    lo = min(15 if config != "c3" else 5, hi)
    if config == "c3":
        sets = _sets_loguniform(rng, G, N, lo, hi)
        if c5:
            # 3% of the sets carry a planted share (2-30%) of shifted genes
            for s in rng.choice(G, size=max(1, (3 * G) // 100), replace=False):
                k = sets[s].size
                m = min(shifted.size, max(1, int(round(k * rng.uniform(0.02, 0.3)))))
                keep = rng.choice(sets[s], size=k - m, replace=False)
                add = rng.choice(shifted, size=m, replace=False)
                sets[s] = np.unique(np.concatenate([keep, add]))
    else:
        sets = _sets_uniform(rng, G, N, lo, hi)
    off, genes = _csr(sets)
    if c5:
        config, metrics = "c5", ["stable_signal2noise"]
    return {
        "config": config, "X": np.ascontiguousarray(X, dtype=dtype),
        "labels": lab, "set_offsets": off, "set_genes": genes,
        "gene_symbols": ["G%05d" % g for g in range(N)],
        "set_names": ["SET%05d" % s for s in range(G)],
        "num_perms": P, "metrics": metrics,
        "N": N, "S": S, "G": G, "P": P,
    }
