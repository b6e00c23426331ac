"""Benchmark: plain-GPU permutation GSEA on one MI355X.

Workload (default): 20,000 genes x 500 samples (labels 150/350), 10,000 gene
sets with sizes log-uniform in [5, 2000], 16,384 permutations, all 14 ranking
metrics -- the ragged batched configuration.  One step = the full GSEA path
(label permutation, metric, argsort, enrichment scores, NES/NP/FWER/FDR) for
each of the 14 metrics, inputs resident in HBM.  value = ES/s = 14 * G * P *
steps / time.

Prints one JSON line with roofline and cpu_baseline objects.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(kernel, N, S, G, P, sumk, w):
    """Per-launch algorithmic HBM bytes, materialised-stage model:
    metric:  P*(N*w + S) + N*S*w          (write r, read labels; X once)
    rank_es: P*(N*(3w+8) + G*w) + N*G/8 + 4*sumk
             (read r, write+read ranked keys and 4-byte ids, write ES;
              membership once)."""
    if kernel == "metric":
        return P * (N * w + S) + N * S * w
    return P * (N * (3 * w + 8) + G * w) + N * G / 8 + 4 * sumk


def cpu_baseline(w, metric, target_s):
    """The oracle (C restatement of the reference CPU path, OpenMP over
    permutations) on a bounded permutation sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.build(ref=False)
    cores = oracle.lib().orc_num_threads()
    P0 = 4
    t = time.perf_counter()
    oracle.gsea(w["X"], w["labels"], w["set_offsets"], w["set_genes"], P0,
                metric, want_es=False)
    dt = time.perf_counter() - t
    P1 = int(max(P0, min(4096, P0 * target_s / max(dt, 1e-3))))
    t = time.perf_counter()
    oracle.gsea(w["X"], w["labels"], w["set_offsets"], w["set_genes"], P1,
                metric, want_es=False)
    dt = time.perf_counter() - t
    return {"value": w["G"] * P1 / dt, "unit": "ES/s", "cores": cores,
            "kind": "port",
            "sample": "%s, %d genes x %d samples, %d sets, %d permutations "
                      "(of %d), %.1f s on %d threads"
                      % (metric, w["N"], w["S"], w["G"], P1, w["P"], dt, cores)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--metrics", default="all",
                    help="comma list or 'all' (the configuration's metrics)")
    ap.add_argument("--perms", type=int, default=0, help="override permutations")
    ap.add_argument("--double", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py runs on one GPU only: multi-GPU permutation sharding is "
              "a design note, not part of this build (--gpus %d requested)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import cudagsea_amd as cg
    from cudagsea_amd import workloads

    torch.cuda.set_device(0)
    dt = np.float64 if args.double else np.float32
    w = workloads.make(args.config, dtype=dt,
                       num_perms=args.perms if args.perms else None)
    metrics = w["metrics"] if args.metrics == "all" else args.metrics.split(",")
    N, S, G, P = w["N"], w["S"], w["G"], w["P"]
    sumk = int(w["set_offsets"][-1])
    X = torch.from_numpy(w["X"]).cuda()
    off = torch.from_numpy(w["set_offsets"]).cuda()
    genes = torch.from_numpy(w["set_genes"]).cuda()
    stats = torch.empty((5, G), dtype=X.dtype, device="cuda")
    print("bench: %s N=%d S=%d G=%d P=%d metrics=%d" % (args.config, N, S, G, P,
                                                         len(metrics)),
          file=sys.stderr, flush=True)

    def step():
        tims = []
        for m in metrics:
            tims.append(cg.run_device(X, w["labels"], off, genes, P, m, stats))
        return tims

    for i in range(args.warmup):
        step()
        print("warmup %d done" % i, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    all_t = []
    for i in range(args.steps):
        all_t.extend(step())
        print("step %d done" % i, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    n_runs = len(all_t)
    es_total = float(G) * P * n_runs
    value = es_total / elapsed
    ms_metric = sum(t["ms_metric"] for t in all_t)
    ms_rank = sum(t["ms_rank_es"] for t in all_t)
    ms_stats = sum(t["ms_stats"] for t in all_t)
    batches = sum(t["perm_batches"] for t in all_t)
    wbytes = np.dtype(dt).itemsize
    # dominant kernel and its roofline
    if ms_rank >= ms_metric:
        kern, ms, name = "rank_es", ms_rank, "k_rank_es"
    else:
        kern, ms, name = "metric", ms_metric, "k_metric"
    per_launch_ms = ms / batches
    per_launch_P = P / (batches / n_runs)
    abytes = algorithmic_bytes(kern, N, S, G, per_launch_P, sumk, wbytes)
    achieved = abytes / (per_launch_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s_%s.txt" % (name, args.config))
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("bytes_per_launch")
        except (ValueError, OSError):
            traffic = None
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic, "kernel": name,
            "launch_ms": round(per_launch_ms, 4),
            "algorithmic_bytes": int(abytes)}
    stage_ms = {"metric": round(ms_metric / n_runs, 3),
                "rank_es": round(ms_rank / n_runs, 3),
                "stats": round(ms_stats / n_runs, 3)}

    cpu = None
    if not args.no_cpu_baseline:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(w, metrics[len(metrics) // 2], args.cpu_seconds)

    line = {
        "metric": "ES/sec (gene-set x perm) at 20k genes, 1 GPU",
        "value": value, "unit": "ES/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.double else "f32", "data": "synthetic",
        "config": {"workload": "%s: %d genes x %d samples, %d ragged gene sets, "
                               "%d permutations, %d metrics per step"
                               % (args.config, N, S, G, P, len(metrics)),
                   "genes": N, "samples": S, "gene_sets": G,
                   "permutations": P, "metrics": len(metrics),
                   "mean_ms_per_metric_run": stage_ms},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
