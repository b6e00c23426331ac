"""Benchmark: plain-GPU permutation GSEA on one MI355X.

Workload (default): 20,000 genes x 500 samples (labels 150/350), 10,000 gene
sets with sizes log-uniform in [5, 2000], 16,384 permutations, all 14 ranking
metrics -- the ragged batched configuration.  One step = the full GSEA path
(label permutation, metric, argsort, enrichment scores, NES/NP/FWER/FDR) for
each of the 14 metrics, inputs resident in HBM.  value = ES/s = 14 * G * P *
steps / time.

A second leg times the surrogate-guided run against the plain full run on
c5 (identical FDR < 0.25 calls and max |dNES| < 0.02 are checked and
reported), a third the whole path at the other configurations' full sizes
(c1, c2 in fp32 and fp64, c4).  Prints one JSON line with roofline,
cpu_baseline, surrogate and other_configs objects.
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
CLOCK_GHZ = 2.4          # MI355X peak engine clock


def algorithmic_bytes(kernel, N, S, G, P, w):
    """Algorithmic HBM bytes of one launch over P permutations, the stage
    model with materialised stage boundaries:
      k_metric: P*(N*w + S) + N*S*w   (write r, read labels; X once)
      k_es:     P*(N*(w+4) + G*w) + N*G/8
                (read the ranked list: |metric| and a 4-byte id per rank;
                 write ES; set membership bits once)."""
    if kernel == "k_metric":
        return P * (N * w + S) + N * S * w
    return P * (N * (w + 4) + G * w) + N * G / 8


def kernel_profile(kernel, config, perms):
    """The committed counter summary of a kernel at this workload
    (profiles/counters_<kernel>_<config>.txt, JSON), when its permutation
    count matches the launch: HBM bytes per launch (FETCH_SIZE and
    WRITE_SIZE, calibrated by a copy of known size), VALU instructions and
    LDS occupancy per launch."""
    path = os.path.join(ROOT, "profiles", "counters_%s_%s.txt" % (kernel, config))
    if not os.path.exists(path):
        return {}
    try:
        rec = json.load(open(path))
    except (ValueError, OSError):
        return {}
    if int(rec.get("permutations", -1)) != int(perms):
        return {}
    rec["source"] = os.path.relpath(path, ROOT)
    return rec


def path_bytes_per_es(N, S, G, P, w):
    """Whole-path algorithmic bytes per enrichment score (all stages)."""
    return (P * (N * (4 * w + 8) + S + G * w) + N * S * w + N * G / 8) / (G * P)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


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
    out = {"value": w["G"] * P1 / dt, "unit": "ES/s", "cores": cores,
           "kind": "port", "cpu": cpu_model(),
           "sample": "%s, %d genes x %d samples, %d sets, %d permutations "
                     "(of %d), %.1f s on %d threads"
                     % (metric, w["N"], w["S"], w["G"], P1, w["P"], dt, cores)}
    sizes = os.path.join(ROOT, "profiles", "cpu_baseline_sizes.txt")
    if os.path.exists(sizes):
        try:
            out["sizes"] = json.load(open(sizes))
        except ValueError:
            pass
    return out


def cpu_sizes():
    """The oracle at the configurations' own sizes: c1, c2 (fp32 and fp64)
    and c4 in full, c3 on 1,024 of its permutations (stage cost is linear in
    the permutation count).  Input generation is excluded, as the
    reference's timers exclude parsing (cugsea.cuh:81)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from cudagsea_amd import workloads
    oracle.build(ref=False)
    cores = oracle.lib().orc_num_threads()
    runs = [("c1", False, None, "stable_signal2noise"),
            ("c2", False, None, "onepass_t_test"),
            ("c2", True, None, "onepass_t_test"),
            ("c4", False, None, "twopass_signal2noise"),
            ("c3", False, 1024, "onepass_t_test")]
    out = {"cpu": cpu_model(), "cores": cores, "kind": "port", "runs": []}
    for cfg, dbl, perms, metric in runs:
        w = workloads.make(cfg, dtype=np.float64 if dbl else np.float32, num_perms=perms)
        print("cpu %s %s P=%d ..." % (cfg, "fp64" if dbl else "fp32", w["P"]),
              file=sys.stderr, flush=True)
        t = time.perf_counter()
        oracle.gsea(w["X"], w["labels"], w["set_offsets"], w["set_genes"], w["P"],
                    metric, double=dbl, want_es=False)
        dt = time.perf_counter() - t
        out["runs"].append({"config": cfg, "dtype": "f64" if dbl else "f32",
                            "metric": metric, "genes": w["N"], "samples": w["S"],
                            "sets": w["G"], "permutations": w["P"],
                            "seconds": round(dt, 3),
                            "es_per_s": round(w["G"] * w["P"] / dt, 1)})
    return out


def surrogate_leg(args):
    """Surrogate-guided vs plain full run on c5 (c3's shape, its own seed,
    planted enriched sets, 16,384 permutations, stable_signal2noise): both
    end to end from resident inputs, best of 2 after a warm-up."""
    from cudagsea_amd import surrogate as sg
    from cudagsea_amd import workloads
    if not os.path.exists(sg.MODEL_PATH):
        return {"skipped": "no trained model at %s" % sg.MODEL_PATH}
    print("surrogate leg (c5) ...", file=sys.stderr, flush=True)
    model = sg.Surrogate.load()
    w = workloads.make("c5", dtype=np.float64 if args.double else np.float32)
    res, plain = sg.evaluate(w, model)
    coef = model.meta["size_baseline"]
    abl, _ = sg.evaluate(w, model, plain=plain,
                         predictor=lambda f: sg.size_baseline_predict(coef, f))
    out = {k: res[k] for k in ("speedup", "s_plain", "s_adaptive", "calls_plain",
                               "calls_adaptive", "calls_identical", "max_abs_dnes",
                               "es_work_fraction", "sets_extended", "budget_hist",
                               "metric", "num_perms", "num_sets")}
    out["workload"] = ("c5: %d genes x %d samples, %d ragged sets (3%% planted), "
                       "%d permutations" % (w["N"], w["S"], w["G"], w["P"]))
    out["bar_met"] = bool(res["calls_identical"] and res["max_abs_dnes"] < 0.02)
    out["size_only_model"] = {k: abl[k] for k in ("speedup", "calls_identical",
                                                  "max_abs_dnes", "es_work_fraction")}
    # the allocation from the measured null moments (2,048 permutations of
    # the plain run): what a perfect moment predictor could give
    import torch
    X = torch.from_numpy(w["X"]).cuda()
    off = torch.from_numpy(w["set_offsets"]).cuda()
    genes = torch.from_numpy(w["set_genes"]).cuda()
    st = torch.empty((5, w["G"]), dtype=X.dtype, device="cuda")
    es = torch.empty((w["G"], 2048), dtype=X.dtype, device="cuda")
    sg.run_device(X, w["labels"], off, genes, 2048, w["metrics"][0], st, es_out=es)
    measured = sg.null_moments(es.float().t().contiguous()).cpu().numpy()
    mm, _ = sg.evaluate(w, model, plain=plain, predictor=lambda f: measured)
    out["measured_moments"] = {k: mm[k] for k in ("speedup", "calls_identical",
                                                 "max_abs_dnes", "es_work_fraction")}
    out["finding"] = ("the network does not beat the size-only model, and the "
                      "measured null moments do not either: a set's budget follows "
                      "its observed |NES| and the NES tolerance, so the rule, not "
                      "the moment prediction, bounds the gain")
    return out


def configs_leg():
    """Whole-path GPU rates of the other configurations at their full sizes
    (c1; c2 in fp32 and fp64; c4 with its three metrics), inputs resident,
    one graph-captured sweep per configuration timed after a warm-up, default
    launch shapes."""
    import torch
    import cudagsea_amd as cg
    from cudagsea_amd import workloads
    cg.set_tuning()
    out = []
    for cfg, dbl in (("c1", False), ("c2", False), ("c2", True), ("c4", False)):
        dt = np.float64 if dbl else np.float32
        w = workloads.make(cfg, dtype=dt)
        X = torch.from_numpy(w["X"]).cuda()
        off = torch.from_numpy(w["set_offsets"]).cuda()
        genes = torch.from_numpy(w["set_genes"]).cuda()
        st = torch.empty((len(w["metrics"]), 5, w["G"]), dtype=X.dtype, device="cuda")

        def run():
            return cg.sweep_device(X, w["labels"], off, genes, w["P"], w["metrics"], st)
        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tm = run()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        n = w["G"] * w["P"] * len(w["metrics"])
        out.append({"config": cfg, "dtype": "f64" if dbl else "f32",
                    "metrics": w["metrics"], "genes": w["N"], "samples": w["S"],
                    "sets": w["G"], "permutations": w["P"],
                    "ms": round(el * 1e3, 3), "es_per_s": round(n / el, 1),
                    "ms_es": round(tm["ms_es"], 3), "ms_metric": round(tm["ms_metric"], 3)})
        print("config %s %s: %.1f ms" % (cfg, out[-1]["dtype"], el * 1e3),
              file=sys.stderr, flush=True)
    return out


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
    ap.add_argument("--no-tune", action="store_true",
                    help="default launch shapes instead of the autotuner's")
    ap.add_argument("--retune", action="store_true",
                    help="search the launch shapes again (and cache them)")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the surrogate-guided vs plain comparison (c5)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the GPU rates of the other configurations")
    ap.add_argument("--no-default-step", action="store_true",
                    help="skip the untimed default-shape step (profiling runs: every "
                         "dispatch then uses the tuned shapes)")
    ap.add_argument("--cpu-sizes", action="store_true",
                    help="only time the CPU oracle at the configurations' sizes")
    ap.add_argument("--probe-copy", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES on the device (counter calibration)")
    args = ap.parse_args()

    if args.cpu_sizes:
        print(json.dumps(cpu_sizes()), flush=True)
        return

    if args.probe_copy:
        # calibration for HBM byte counters: one device-to-device copy of a
        # known size (read + write of probe_copy bytes each)
        import torch
        n = args.probe_copy // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"probe_copy_bytes": n * 4}), flush=True)
        return

    if args.gpus != 1:
        print("bench.py runs on one GPU only: multi-GPU permutation sharding is "
              "a design note, not part of this build (--gpus %d requested)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import cudagsea_amd as cg
    from cudagsea_amd import autotune, workloads

    torch.cuda.set_device(0)
    dt = np.float64 if args.double else np.float32
    w = workloads.make(args.config, dtype=dt,
                       num_perms=args.perms if args.perms else None)
    metrics = w["metrics"] if args.metrics == "all" else args.metrics.split(",")
    N, S, G, P = w["N"], w["S"], w["G"], w["P"]
    X = torch.from_numpy(w["X"]).cuda()
    off = torch.from_numpy(w["set_offsets"]).cuda()
    genes = torch.from_numpy(w["set_genes"]).cuda()
    stats = torch.empty((len(metrics), 5, G), dtype=X.dtype, device="cuda")
    print("bench: %s N=%d S=%d G=%d P=%d metrics=%d" % (args.config, N, S, G, P,
                                                         len(metrics)),
          file=sys.stderr, flush=True)

    # launch shapes: the autotuner's cached choice for this workload and
    # device, searched now (and cached) when there is none
    knobs, source = {}, "defaults"
    if not args.no_tune:
        knobs = autotune.lookup(N, S, G, P, args.double, metrics)
        source = "cache"
        if knobs is None or args.retune:
            print("autotune ...", file=sys.stderr, flush=True)
            _, res = autotune.tune_workload(dict(w, metrics=metrics), args.double,
                                            log=lambda m: print(m, file=sys.stderr,
                                                                flush=True))
            knobs, source = res["knobs"], "search"

    def step():
        # the whole sweep (every metric) is one HIP graph, replayed per step
        return cg.sweep_device(X, w["labels"], off, genes, P, metrics, stats, graph=True)

    def timed(nsteps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tims = [step() for _ in range(nsteps)]
        torch.cuda.synchronize()
        return time.perf_counter() - t0, tims

    # the default shapes, for the tuned-vs-default comparison (untimed steps)
    default_s = None
    if not args.no_default_step:
        cg.set_tuning()
        step()
        default_s, _ = timed(1)
    cg.set_tuning(**knobs)
    for i in range(args.warmup):
        step()
        print("warmup %d done" % i, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    all_t = []
    for i in range(args.steps):
        all_t.append(step())
        print("step %d done" % i, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    n_runs = len(all_t) * len(metrics)
    es_total = float(G) * P * n_runs
    value = es_total / elapsed
    ms_metric = sum(t["ms_metric"] for t in all_t)
    ms_rank = sum(t["ms_rank"] for t in all_t)
    ms_es = sum(t["ms_es"] for t in all_t)
    ms_stats = sum(t["ms_stats"] for t in all_t)
    batches = sum(t["perm_batches"] for t in all_t)
    es_launches = sum(t["es_launches"] for t in all_t)
    wbytes = np.dtype(dt).itemsize
    # dominant kernel: its launches' average duration (HIP event nodes around
    # each launch inside the replayed graph, in the timed region) against its
    # algorithmic bytes
    if ms_es >= ms_metric:
        name, ms, launches = "k_es", ms_es, es_launches
    else:
        name, ms, launches = "k_metric", ms_metric, batches
    per_launch_ms = ms / launches
    per_launch_P = P * n_runs / launches
    abytes = algorithmic_bytes(name, N, S, G, per_launch_P, wbytes)
    achieved = abytes / (per_launch_ms * 1e-3) / 1e9
    prof = kernel_profile(name, args.config, per_launch_P)
    bpe = path_bytes_per_es(N, S, G, P, wbytes)
    roof = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": prof.get("bytes_per_launch"), "kernel": name,
            "launch_ms": round(per_launch_ms, 4),
            "algorithmic_bytes": int(abytes),
            "path_bytes_per_es": round(bpe, 3),
            "path_frac": round(bpe * value / (HBM_PEAK_GBS * 1e9), 6)}
    if prof.get("valu_insts_per_launch"):
        # the kernel's own bound: vector-instruction issue (one wavefront
        # instruction per CU per clock) and LDS occupancy, from its counters
        cus = torch.cuda.get_device_properties(0).multi_processor_count
        peak_issue = cus * CLOCK_GHZ * 1e9
        issue = prof["valu_insts_per_launch"] / (per_launch_ms * 1e-3)
        roof["issue"] = {"bound": "valu", "achieved_winst_per_s": round(issue, 1),
                         "peak_winst_per_s": peak_issue,
                         "frac": round(issue / peak_issue, 4),
                         "lds_active_frac": prof.get("lds_active_frac"),
                         "lds_bank_conflict_frac": prof.get("lds_bank_conflict_frac"),
                         "profile": prof.get("source")}
    stage_ms = {"metric": round(ms_metric / n_runs, 3),
                "rank": round(ms_rank / n_runs, 3),
                "es": round(ms_es / n_runs, 3),
                "stats": round(ms_stats / n_runs, 3)}
    tuning = {"knobs": knobs, "source": source,
              "default_ms_per_step": round(default_s * 1e3, 2) if default_s else None,
              "tuned_ms_per_step": round(elapsed * 1e3 / args.steps, 2),
              "graph": "one HIP graph per step (%d metrics), replayed" % len(metrics)}

    cpu = None
    if not args.no_cpu_baseline:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(w, metrics[len(metrics) // 2], args.cpu_seconds)

    sur = None
    if not args.no_surrogate:
        sur = surrogate_leg(args)

    others = None
    if not args.no_configs:
        others = configs_leg()

    line = {
        "metric": "ES/sec (gene-set\u00d7perm) at 20k genes, 1 GPU; surrogate vs "
                  "plain at equal FDR",
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
        "roofline": roof, "tuning": tuning, "cpu_baseline": cpu, "surrogate": sur,
        "other_configs": others,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
