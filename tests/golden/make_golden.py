"""Write tests/golden/ref_vectors.npz from the reference's own arithmetic.

Runs only where the reference sources exist: oracle/ref/Makefile compiles
ref_driver.cpp against the reference headers under /root/reference into
oracle/_ref/libgsea_ref.so; this script calls that library and stores inputs
and outputs (data only) so the tests can pin the oracle without the
reference present.

Vectors:
  perm_S<S>      label permutations of arange(S) for perm ids PERMS (RNG +
                 Fisher-Yates, rngpu.hpp)
  metric_X_f32/f64, metric_labels, metric_<m>_f32/f64
                 the 14 metrics (functors.cuh) over permutations 0..7 of a
                 small gene-major table, reference loop order
  stats_in_f32/f64, stats_out_f32/f64, stats_norm_f32/f64
                 final_statistics (statistics.cuh) on a random ES matrix
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

PERMS = [1, 2, 3, 17, 255, 4096, 65535, 1000003]
SIZES = [2, 3, 10, 40, 200, 500]


def main():
    oracle.build(ref=True)
    out = {"perm_ids": np.array(PERMS, dtype=np.uint64)}
    for S in SIZES:
        base = np.arange(S, dtype=np.uint8) if S <= 255 else (
            np.arange(S) % 251).astype(np.uint8)
        out["perm_S%d" % S] = np.stack(
            [oracle.permute_labels(base, p, use_ref=True) for p in PERMS])

    rng = np.random.default_rng(7)
    N, S = 48, 14
    X = np.exp(5 + rng.standard_normal((N, S)))
    X[3, :] = 11.0                        # constant gene
    X[5, :] = 0.0                         # all-zero gene (0/0 metrics)
    X[7, ::2] = 1e4 + 1e-2 * rng.standard_normal(S // 2)  # cancellation
    X[7, 1::2] = 1e4
    labels = np.array([0, 1] * (S // 2), dtype=np.uint8)
    out["metric_labels"] = labels
    for dbl, sfx in ((False, "f32"), (True, "f64")):
        Xt = X.astype(np.float64 if dbl else np.float32)
        out["metric_X_" + sfx] = Xt
        for m in oracle.METRICS:
            out["metric_%s_%s" % (m, sfx)] = oracle.correlate(
                Xt, labels, 8, m, double=dbl, use_ref=True)

    es = (rng.standard_normal((12, 64)) * 0.3)
    es[2, :] = 0.0                        # all-zero set row: NaN factor
    es[4, 0] = 0.9
    for dbl, sfx in ((False, "f32"), (True, "f64")):
        e = es.astype(np.float64 if dbl else np.float32)
        st, norm = oracle.final_statistics(e, use_ref=True)
        out["stats_in_" + sfx] = e
        out["stats_out_" + sfx] = st
        out["stats_norm_" + sfx] = norm
    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
