"""Autotuner of the engine's launch shapes (gsea_set_tuning).

The reference fixes these choices: 256-thread metric blocks
(correlate_genes.cuh:162-163), 128-permutation sort chunks
(batched_sort.cuh:85-90), 32/64-thread enrichment blocks
(enrichment_scores.cuh:240-245) and a free-memory batch size
(batch_scheduler.cuh:27-33).  Here they are measured on the workload itself:

  metric_gene_tile x metric_perms_per_block   per accumulator kind, timed by
                                              the metric stage of one metric
                                              of that kind
  es_lane_set_max x es_pair_set_max x es_waves
                                              timed by the enrichment stage
                                              (the pair class, sets of
                                              129..256 genes on two lanes,
                                              exists with lane split 128)

Each group is searched exhaustively with the others at their best (two
rounds of coordinate descent) on a permutation subset of the real inputs
(stage times are linear in the permutation count), timing the stages with the
engine's own HIP events, best of `reps` runs.  The accumulator kinds differ
in where their time goes (the naive sums are bound by the LDS reads of X and
prefer wide tiles streamed over many permutations, the Welford and
compensated kinds by arithmetic), hence one tile per kind.  The sort's radix
width is not a knob: k_rank's 8-bit digits keep the 16 wavefronts' digit
counters within LDS beside the keys.

Results are cached in tuning_cache.json next to this file, keyed by genes,
samples, sets, permutations, precision, metric kind mix and the device's CU
count; `python -m cudagsea_amd.autotune --config c3` fills the entry of a
benchmark configuration.  No setting changes any result: every shape computes
the same arithmetic on the same permutations.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import gsea as _g

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE_PATH = os.path.join(HERE, "tuning_cache.json")

METRIC_SPACE = {"metric_gene_tile": [16, 32, 64],
                "metric_perms_per_block": [128, 256, 512, 1024, 2048]}
ES_SPACE = {"es_lane_set_max": [-1, 32, 64, 128],
            "es_pair_set_max": [-1, 192, 256],
            "es_waves": [16, 14, 12]}


def es_candidates():
    """The enrichment-stage settings searched: every combination of
    ES_SPACE, except pair-class settings with a lane split other than 128
    (the pair class exists only with that split, so they would repeat the
    pair-less run)."""
    keys = list(ES_SPACE)
    cands = [dict(zip(keys, v)) for v in np.array(np.meshgrid(
        *[ES_SPACE[k] for k in keys])).T.reshape(-1, len(keys)).tolist()]
    return [c for c in cands
            if c["es_lane_set_max"] == 128 or c["es_pair_set_max"] == -1]


def cu_count():
    import torch
    return torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count


def _acc_kind(m):
    """Accumulator of a metric (functors.cuh): naive 0, Kahan 1, one-pass 2,
    two-pass 3, Welford 4, compensated two-pass 5."""
    i = _g.metric_id(m)
    return 0 if i <= 2 else 1 if i <= 5 else 2 + (i - 6) // 2


def cache_key(N, S, G, P, double, metrics, cus):
    kinds = sorted({_acc_kind(m) for m in metrics})
    return "N=%d,S=%d,G=%d,P=%d,dtype=%s,kinds=%s,cu=%d" % (
        N, S, G, P, "f64" if double else "f32", "-".join(map(str, kinds)), cus)


def load_cache(path=CACHE_PATH):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save_entry(key, entry, path=CACHE_PATH):
    cache = load_cache(path)
    cache[key] = entry
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def lookup(N, S, G, P, double, metrics, path=CACHE_PATH):
    """Cached knobs for this workload on this device, or None."""
    e = load_cache(path).get(cache_key(N, S, G, P, double, metrics, cu_count()))
    return e["knobs"] if e else None


def _stage_ms(X, labels, off, genes, P, metrics, knobs, stat, reps):
    """Best-of-reps sum over metrics of the timings' `stat` fields."""
    import torch
    G = off.numel() - 1
    st = torch.empty((5, G), dtype=X.dtype, device="cuda")
    _g.set_tuning(**knobs)
    try:
        total = 0.0
        for m in metrics:
            best = None
            for _ in range(reps + 1):            # first run warms up
                t = _g.run_device(X, labels, off, genes, P, m, st)
                v = sum(t[k] for k in stat)
                best = v if best is None else min(best, v)
            total += best
    finally:
        _g.set_tuning()
    return total


def _with_kind(knobs, kind, gt, ppb):
    """knobs with kind's metric tile set (per-kind lists of six values)."""
    out = dict(knobs)
    tiles = list(out.get("metric_gene_tile_kind", [0] * 6))
    ppbs = list(out.get("metric_perms_per_block_kind", [0] * 6))
    tiles[kind], ppbs[kind] = int(gt), int(ppb)
    out["metric_gene_tile_kind"], out["metric_perms_per_block_kind"] = tiles, ppbs
    return out


def tune(X, labels, off, genes, num_perms, metrics, sample_perms=4096, reps=2,
         log=None):
    """Search the knobs on the first min(num_perms, sample_perms)
    permutations; returns {"knobs", "default_ms", "tuned_ms", "trials",
    "sample_perms"} with run times scaled to num_perms (metric + rank + ES
    + statistics stages, summed over the metrics)."""
    P = int(min(num_perms, sample_perms))
    scale = num_perms / P
    say = log or (lambda *a: None)
    # one metric per accumulator kind times that kind's tile; the enrichment
    # stage hardly depends on the metric: tune it on one
    by_kind = {}
    for m in metrics:
        by_kind.setdefault(_acc_kind(m), m)
    es_metrics = metrics[:1]
    best = {}
    trials = 0
    t0 = time.perf_counter()
    for rnd in range(2):
        for kind, m in sorted(by_kind.items()):
            scored = []
            for gt in METRIC_SPACE["metric_gene_tile"]:
                for ppb in METRIC_SPACE["metric_perms_per_block"]:
                    knobs = _with_kind(best, kind, gt, ppb)
                    scored.append((_stage_ms(X, labels, off, genes, P, [m], knobs,
                                             ("ms_metric",), reps), knobs))
                    trials += 1
            scored.sort(key=lambda x: x[0])
            best = scored[0][1]
            say("round %d metric kind %d (%s): tile %d x %d, %.3f ms" % (
                rnd, kind, m, best["metric_gene_tile_kind"][kind],
                best["metric_perms_per_block_kind"][kind], scored[0][0]))
        keys = list(ES_SPACE)
        scored = []
        for c in es_candidates():
            knobs = dict(best)
            knobs.update({k: int(v) for k, v in c.items()})
            scored.append((_stage_ms(X, labels, off, genes, P, es_metrics, knobs, ("ms_es",),
                                     reps), knobs))
            trials += 1
        scored.sort(key=lambda x: x[0])
        best = scored[0][1]
        say("round %d ms_es: best %s %.3f ms" % (rnd, {k: best[k] for k in keys},
                                                  scored[0][0]))
    stages = ("ms_metric", "ms_rank", "ms_es", "ms_stats")
    default_ms = _stage_ms(X, labels, off, genes, P, metrics, {}, stages, reps) * scale
    tuned_ms = _stage_ms(X, labels, off, genes, P, metrics, best, stages, reps) * scale
    if tuned_ms > default_ms:
        best, tuned_ms = {}, default_ms      # never worse than the defaults
    return {"knobs": best, "default_ms": round(default_ms, 3),
            "tuned_ms": round(tuned_ms, 3), "trials": trials, "sample_perms": P,
            "search_s": round(time.perf_counter() - t0, 1)}


def tune_workload(w, double=False, sample_perms=4096, log=None, path=CACHE_PATH):
    """Tune a workloads.make() workload and store the cache entry."""
    import torch
    dt = np.float64 if double else np.float32
    X = torch.from_numpy(np.ascontiguousarray(w["X"], dtype=dt)).cuda()
    off = torch.from_numpy(w["set_offsets"]).cuda()
    genes = torch.from_numpy(w["set_genes"]).cuda()
    res = tune(X, w["labels"], off, genes, w["P"], w["metrics"], sample_perms, log=log)
    key = cache_key(w["N"], w["S"], w["G"], w["P"], double, w["metrics"], cu_count())
    res["device"] = torch.cuda.get_device_name()
    save_entry(key, res, path)
    return key, res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--config", default="c3")
    ap.add_argument("--double", action="store_true")
    ap.add_argument("--sample-perms", type=int, default=4096)
    ap.add_argument("--cache", default=CACHE_PATH, help="cache file to update")
    args = ap.parse_args()
    from . import workloads
    w = workloads.make(args.config, dtype=np.float64 if args.double else np.float32)
    key, res = tune_workload(w, args.double, args.sample_perms,
                             log=lambda m: print(m, file=sys.stderr, flush=True),
                             path=args.cache)
    print(json.dumps({"key": key, **res}))


if __name__ == "__main__":
    main()
