/*
 * gsea_hip.h -- C ABI of the MI355X (gfx950) permutation-GSEA engine,
 * libgsea_hip.so.  Plain pointers and sizes only.
 *
 * The reference binds its engine below Rcpp through the C++ proxy ABI of
 * cudaGSEA/src/cudaGSEA_wrapper.cuh:15-61 (implemented in
 * cudaGSEA/src/cudaGSEA_wrapper.cu).  That ABI is provided unchanged by
 * include/cudaGSEA_wrapper.cuh on top of the functions below; this header is
 * the thin C layer underneath it, callable from C, ctypes, cgo or JNI.
 *
 * Conventions
 *   - exprs is sample-major [S][N]: exprs[s*N + g], the layout the R path
 *     hands over (cudaGSEA_wrapper.cpp:231/:253, cudaGSEA_wrapper.cu:229-236).
 *   - labels: S bytes, 0 = class A, 1 = class B (cudaGSEA_wrapper.cu:239-244).
 *   - gene sets: CSR over gene indices, set_offsets[G+1], set_genes[...];
 *     within a set gene indices must be strictly increasing and < N (this is
 *     what gsea_map_sets produces from symbols).
 *   - es_out (optional) is [G][P] set-major, column 0 the observed labels:
 *     the reference's global_result / dump layout (copy_strided.cuh:9-36).
 *   - stats_out is [5][G]: ES, NES, NP, FWER, FDR (statistics.cuh:51-53; the
 *     reference leaves FDR NaN, statistics.cuh:112-114, this engine fills it).
 *   - Return value: 0 on success, otherwise an error code below; the message
 *     is available from gsea_last_error() (per thread).  Nothing calls exit().
 */
#ifndef GSEA_HIP_H
#define GSEA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes: 1-22 keep the numbering of cudaGSEA/src/include/error_codes.cuh */
#define GSEA_OK                         0
#define GSEA_ERR_UNKNOWN_METRIC         1   /* CUDA_GSEA_NO_KNOWN_METRIC_SPECIFIED_ERROR */
#define GSEA_ERR_CANNOT_OPEN_FILE       2   /* CUDA_GSEA_CANNOT_OPEN_FILE_ERROR */
#define GSEA_ERR_NO_PATHS               10  /* CUDA_GSEA_NO_PATHS_SPECIFIED_ERROR */
#define GSEA_ERR_EMPTY_DUMP_FILE_NAME   19  /* CUDA_GSEA_EMPTY_DUMP_FILE_NAME_ERROR */
#define GSEA_ERR_INVALID_ARGUMENT       30  /* bad sizes, labels or gene sets */
#define GSEA_ERR_UNSUPPORTED_SIZE       31  /* exceeds a documented engine limit */
#define GSEA_ERR_HIP                    32  /* HIP runtime error */
#define GSEA_ERR_OUT_OF_MEMORY          33

/* flags */
#define GSEA_FLAG_ASCENDING   1   /* reference sort_direction == false */
#define GSEA_FLAG_SWAP_LABELS 2   /* reference swap_labels == true */
/* Broad GSEA normalisation instead of statistics.cuh:59-70's single factor
 * P / sum_p |ES_p|: a score >= 0 is divided by the mean of its set's null
 * scores >= 0, a score < 0 by the mean magnitude of those < 0 (null =
 * permutations 1..P-1, NaN skipped; an empty side gives NaN).  NES, FWER and
 * FDR follow the normalised scores; ES and NP are unchanged.  No
 * counterpart in the reference, which has the single factor only. */
#define GSEA_FLAG_SIGNED_NES  4

/* engine limits */
#define GSEA_MAX_GENES   65535   /* 16-bit rank rows */
#define GSEA_MAX_SAMPLES 65535

/* per-stage device times of the last call, milliseconds (HIP events) */
typedef struct gsea_timings {
    double ms_lists;      /* permutation class lists */
    double ms_metric;     /* local ranking metric */
    double ms_rank_es;    /* argsort + enrichment scores (ms_rank + ms_es) */
    double ms_stats;      /* NES, NP, FWER, FDR */
    double ms_transpose;  /* [P][G] -> [G][P] when es_out is requested */
    double ms_total;      /* first launch to last completion */
    int64_t perm_batches; /* permutation batches used */
    double ms_rank;       /* argsort (rank rows) */
    double ms_es;         /* enrichment-score kernel launches */
    int64_t es_launches;  /* enrichment-score kernel launches timed */
} gsea_timings;

/* Launch-shape tuning (no counterpart in the reference, which fixes these
 * choices: batch_scheduler.cuh:27-33, correlate_genes.cuh:162-163,
 * batched_sort.cuh:85-90, enrichment_scores.cuh:240-245).  Zero fields mean
 * "choose automatically".  The setting is process-wide, applies to the calls
 * that follow, and never changes results: permutations are seeded by their
 * global index and every shape computes the same arithmetic.
 *   batch_perms        permutations per batch (default: the metric rows of
 *                      one batch stay below 2 GiB)
 *   metric_gene_tile   genes per k_metric workgroup: 16, 32 or 64 (the X
 *                      tile [S][tile] must fit LDS; default: the widest
 *                      within 64 KiB)
 *   metric_perms_per_block  permutations a k_metric workgroup streams over
 *                      its X tile (default: enough workgroups for 8 per CU)
 *   es_waves           wavefronts of each k_es workgroup that score sets,
 *                      1..16 (default: as many as the LDS holds)
 *   es_lane_set_max    largest set scored by one lane (8, 16, 32, 64 or 128;
 *                      -1: every set by a wavefront; default 128)
 *   es_pair_set_max    largest set scored by a pair of lanes (129..256; -1:
 *                      none; default 256); with es_lane_set_max 128 only
 *   metric_gene_tile_kind[a], metric_perms_per_block_kind[a]
 *                      the two metric knobs for the metrics of accumulator
 *                      kind a (0 naive, 1 Kahan, 2 one-pass, 3 two-pass,
 *                      4 Welford, 5 compensated two-pass); 0: the value above
 * The per-kind fields and es_pair_set_max occupy what earlier versions
 * reserved (zero-filled), so the structure's size and the other offsets are
 * unchanged. */
typedef struct gsea_tuning {
    int64_t batch_perms;
    int32_t metric_gene_tile;
    int32_t metric_perms_per_block;
    int32_t es_waves;
    int32_t es_lane_set_max;
    int32_t es_pair_set_max;
    int16_t metric_gene_tile_kind[6];
    int16_t metric_perms_per_block_kind[6];
    int64_t reserved[2];
} gsea_tuning;
int gsea_set_tuning(const gsea_tuning *tuning);   /* NULL: all automatic */
int gsea_get_tuning(gsea_tuning *tuning);

/* metric name <-> id; ids follow the dispatch order of
 * cudaGSEA/src/include/correlate_genes.cuh:166-295 */
int         gsea_metric_id(const char *name);          /* -1 if unknown */
const char *gsea_metric_name(int id);                  /* "" if unknown */
const char *gsea_last_error(void);
const char *gsea_version(void);

/* device management; replaces cudaGSEA_wrapper.cu:87-138
 * (listCudaDevices_proxy / getCudaDevice_proxy / setCudaDevice_proxy) */
int gsea_device_count(int *count);
int gsea_get_device(int *device);
int gsea_set_device(int device);
int gsea_device_name(int device, char *name, size_t name_len,
                     size_t *total_bytes);

/* Full path with host buffers: replaces gsea_proxy<...>'s device work,
 * cudaGSEA_wrapper.cu:173-401 (correllate_genes_gpu, pathway_sort_gpu,
 * compute_scores_gpu, final_statistics).  fp32: gsea_single_proxy's types
 * (cudaGSEA_wrapper.cu:403-419); fp64: gsea_double_proxy (:421-437). */
int gsea_run_f32(const float *exprs, int64_t num_genes, int64_t num_samples,
                 const uint8_t *labels, const int32_t *set_offsets,
                 const int32_t *set_genes, int64_t num_sets,
                 int64_t num_perms, int metric_id, int flags,
                 float *es_out, float *stats_out, gsea_timings *timings);
int gsea_run_f64(const double *exprs, int64_t num_genes, int64_t num_samples,
                 const uint8_t *labels, const int32_t *set_offsets,
                 const int32_t *set_genes, int64_t num_sets,
                 int64_t num_perms, int metric_id, int flags,
                 double *es_out, double *stats_out, gsea_timings *timings);

/* Same path with every array already in device memory (labels stay on the
 * host: S bytes, read synchronously).  d_es_out may be NULL.  Work is queued
 * on `stream` (a hipStream_t, NULL = default stream); the call returns after
 * the statistics are complete. */
int gsea_run_device_f32(const float *d_exprs, int64_t num_genes,
                        int64_t num_samples, const uint8_t *labels,
                        const int32_t *d_set_offsets,
                        const int32_t *d_set_genes, int64_t num_sets,
                        int64_t num_perms, int metric_id, int flags,
                        float *d_es_out, float *d_stats_out, void *stream,
                        gsea_timings *timings);
int gsea_run_device_f64(const double *d_exprs, int64_t num_genes,
                        int64_t num_samples, const uint8_t *labels,
                        const int32_t *d_set_offsets,
                        const int32_t *d_set_genes, int64_t num_sets,
                        int64_t num_perms, int metric_id, int flags,
                        double *d_es_out, double *d_stats_out, void *stream,
                        gsea_timings *timings);

/* Several metrics over the same device-resident inputs in one call (no
 * counterpart in the reference, whose R entry point runs one metric per
 * call): the class lists once, then per metric the whole path of
 * gsea_run_device_*, statistics into d_stats [num_metrics][5][G].  With
 * use_graph != 0 the sequence is captured into one HIP graph on the first
 * call and replayed by later calls with the same inputs, sizes, labels,
 * metrics, flags, output, stream, tuning and set list (the set CSR's contents
 * are checked); any other call captures again.  Results equal
 * gsea_run_device_*'s metric by metric.  timings: ms_total only. */
int gsea_sweep_device_f32(const float *d_exprs, int64_t num_genes, int64_t num_samples,
                          const uint8_t *labels, const int32_t *d_set_offsets,
                          const int32_t *d_set_genes, int64_t num_sets, int64_t num_perms,
                          const int *metric_ids, int num_metrics, int flags, int use_graph,
                          float *d_stats_out, void *stream, gsea_timings *timings);
int gsea_sweep_device_f64(const double *d_exprs, int64_t num_genes, int64_t num_samples,
                          const uint8_t *labels, const int32_t *d_set_offsets,
                          const int32_t *d_set_genes, int64_t num_sets, int64_t num_perms,
                          const int *metric_ids, int num_metrics, int flags, int use_graph,
                          double *d_stats_out, void *stream, gsea_timings *timings);

/* Budgeted run (no counterpart in the reference; used by the learned
 * permutation allocation).  Set s is scored for permutations
 * 0..budget[s]-1 only; every budget divides num_perms and is >= 2.  Its
 * NES, NP and FWER use those permutations, and the FDR counts each of its
 * null scores num_perms/budget[s] times.  With `ext`, the sets whose
 * significance call is not settled, or whose NES is imprecise, are then
 * scored for all permutations and the statistics recomputed; with se = a
 * set's NES standard error (coefficient of variation of |null ES| times
 * |NES| over sqrt(budget)):
 *   call rule (fdr > 0): the sets of the same NES sign with NES within
 *     z*se of the set's (itself included) have FDR values on both sides of
 *     fdr (or equal to it);
 *   precision rule (nes_tol > 0): z*se > nes_tol.
 * budget_out (host, may be NULL) receives the final budgets; d_es_pg (may
 * be NULL) receives es[P][G], whose entries beyond a set's budget are
 * unspecified.  With every budget equal to num_perms the results equal
 * gsea_run_device_*'s. */
typedef struct gsea_extension {
    double fdr;       /* significance threshold of the call rule */
    double z;         /* standard errors of NES uncertainty */
    double nes_tol;   /* NES precision target */
} gsea_extension;
int gsea_run_budget_device_f32(const float *d_exprs, int64_t num_genes,
                               int64_t num_samples, const uint8_t *labels,
                               const int32_t *d_set_offsets,
                               const int32_t *d_set_genes, int64_t num_sets,
                               int64_t num_perms, int metric_id, int flags,
                               const int32_t *budget, const gsea_extension *ext,
                               int32_t *budget_out, float *d_es_pg,
                               float *d_stats_out, void *stream,
                               gsea_timings *timings);
int gsea_run_budget_device_f64(const double *d_exprs, int64_t num_genes,
                               int64_t num_samples, const uint8_t *labels,
                               const int32_t *d_set_offsets,
                               const int32_t *d_set_genes, int64_t num_sets,
                               int64_t num_perms, int metric_id, int flags,
                               const int32_t *budget, const gsea_extension *ext,
                               int32_t *budget_out, double *d_es_pg,
                               double *d_stats_out, void *stream,
                               gsea_timings *timings);

/* Stage entry points (device pointers), for tests and callers that hold
 * intermediate results.  Replace correllate_genes_gpu
 * (correlate_genes.cuh:121-300) and pathway_sort_gpu + compute_scores_gpu
 * (batched_sort.cuh:85-269, enrichment_scores.cuh:208-256).
 *   gsea_metric_*: r[P][N] for global permutations shift..shift+P-1.
 *   gsea_rank_es_*: es[P][G] (permutation-major) from r[P][N]. */
int gsea_metric_f32(const float *d_exprs, int64_t num_genes,
                    int64_t num_samples, const uint8_t *labels,
                    int64_t num_perms, int64_t shift, int metric_id, int flags,
                    float *d_r, void *stream);
int gsea_metric_f64(const double *d_exprs, int64_t num_genes,
                    int64_t num_samples, const uint8_t *labels,
                    int64_t num_perms, int64_t shift, int metric_id, int flags,
                    double *d_r, void *stream);
/* gsea_metric_f32/f64 without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch (labels, each permutation's permuted labels and class lists, the
 * Welford reciprocal table) lives in a static device array of 32 MiB that
 * each GPU the library runs on holds until the process ends; a call needs
 * at most num_perms * (3 * num_samples + 28) + 9 * num_samples + 1024 bytes
 * of it and larger calls are
 * rejected with GSEA_ERR_UNSUPPORTED_SIZE (split the permutations with
 * `shift`).  Calls on one GPU share that array: calls that overlap on one GPU
 * race.  d_exprs and d_r are float (double_precision == 0) or double.
 * Results equal gsea_metric_f32/f64's. */
int gsea_metric_noalloc(int double_precision, const void *d_exprs, int64_t num_genes,
                        int64_t num_samples, const uint8_t *labels, int64_t num_perms,
                        int64_t shift, int metric_id, int flags, void *d_r,
                        void *stream);

int gsea_rank_es_f32(const float *d_r, int64_t num_genes, int64_t num_perms,
                     const int32_t *d_set_offsets, const int32_t *d_set_genes,
                     int64_t num_sets, int flags, float *d_es, void *stream);
int gsea_rank_es_f64(const double *d_r, int64_t num_genes, int64_t num_perms,
                     const int32_t *d_set_offsets, const int32_t *d_set_genes,
                     int64_t num_sets, int flags, double *d_es, void *stream);

/* Stage (b) alone: per permutation, the gene index at each rank under the
 * stated order (descending metric unless GSEA_FLAG_ASCENDING; ties by
 * ascending gene index; -0 below +0; NaN above +inf).  Replaces
 * pathway_sort_gpu's output Index (batched_sort.cuh:85-269). */
int gsea_argsort_f32(const float *d_keys, int64_t num_genes, int64_t num_perms,
                     int flags, uint32_t *d_index, void *stream);
int gsea_argsort_f64(const double *d_keys, int64_t num_genes,
                     int64_t num_perms, int flags, uint32_t *d_index,
                     void *stream);

/* Gene-set symbols -> CSR of gene indices with create_bitmaps semantics
 * (cudaGSEA/src/include/create_bitmaps.cuh:13-65): gene g belongs to set s
 * when gene_symbols[g] equals one of the set's symbols exactly; symbols
 * missing from the data are ignored; duplicates collapse.  set_symbols of set
 * s are set_symbols[set_symbol_offsets[s] .. set_symbol_offsets[s+1]).
 * set_offsets_out gets G+1 entries; set_genes_out up to `capacity` entries;
 * *total_out receives the number needed (call with capacity 0 to size). */
int gsea_map_sets(const char *const *gene_symbols, int64_t num_genes,
                  const char *const *set_symbols,
                  const int64_t *set_symbol_offsets, int64_t num_sets,
                  int32_t *set_offsets_out, int32_t *set_genes_out,
                  int64_t capacity, int64_t *total_out);

/* Raw ES dump, dump_binary.cuh:10-30 with the naming of
 * cudaGSEA_wrapper.cu:384-390: writes <prefix>_<G>_<P>_<bits>.es, native
 * endian, no header, row-major [G][P]. */
int gsea_dump_es(const char *prefix, const void *es, int64_t num_sets,
                 int64_t num_perms, int bits);

/* Whole R-level call with symbols (what gsea_single_proxy /
 * gsea_double_proxy do, cudaGSEA_wrapper.cu:173-437): map sets, run,
 * dump when dump_prefix is non-empty, return [5][G] statistics. */
int gsea_proxy_run(int double_precision, const void *exprs, int64_t num_genes,
                   int64_t num_samples, const char *const *gene_symbols,
                   const uint8_t *labels, const char *const *set_symbols,
                   const int64_t *set_symbol_offsets, int64_t num_sets,
                   const char *metric, int64_t num_perms, int sort_descending,
                   int swap_labels, const char *dump_prefix, void *stats_out,
                   gsea_timings *timings);

/* ---- learned permutation allocation (no counterpart in the reference) ----
 * A 16 -> 64 -> 64 -> 4 MLP with GELU predicts a set's null ES moments
 * (mean/std of the positive and of the negative null ES) from 16 set
 * features; trained in fp32 by these kernels, inferred in fp32 or bf16.
 * Parameters packed: W1[64][16] b1[64] W2[64][64] b2[64] W3[4][64] b3[4]. */
#define GSEA_MLP_IN 16
#define GSEA_MLP_HIDDEN 64
#define GSEA_MLP_OUT 4
#define GSEA_MLP_NUM_PARAMS 5508

int gsea_mlp_num_params(void);
/* y[B][4] = mlp(x[B][16]); bf16 != 0 rounds weights and activations to
 * bf16.  d_norm (may be NULL): 40 floats x_mean[16] x_inv_std[16]
 * y_scale[4] y_shift[4]; inputs become (x - x_mean) * x_inv_std (in fp32,
 * before any bf16 rounding) and outputs y * y_scale + y_shift. */
int gsea_mlp_forward(const float *d_params, const float *d_x, int64_t batch,
                     int bf16, const float *d_norm, float *d_y, void *stream);
/* gradient of mean((mlp(x) - t)^2) over batch x 4 outputs; *loss (host) */
int gsea_mlp_grad(const float *d_params, const float *d_x, const float *d_t,
                  int64_t batch, float *d_grad, float *loss, void *stream);
/* one Adam step (beta 0.9/0.999, eps 1e-8, decoupled weight decay added to
 * the gradient); step counts from 1 */
int gsea_mlp_adam(float *d_params, float *d_m, float *d_v, const float *d_grad,
                  int64_t step, float lr, float weight_decay, void *stream);
/* features[G][16] of every set from the observed metric r0[N] and the
 * observed ranking rank_genes[N] (gene at each rank, gsea_argsort_f32);
 * glob[4] receives summaries of r0 */
int gsea_set_features(const float *d_r0, const int32_t *d_rank_genes, int64_t num_genes,
                      const int32_t *d_set_offsets, const int32_t *d_set_genes,
                      int64_t num_sets, float *d_glob, float *d_feat,
                      void *stream);
/* out[G][5]: mean/std of positive null ES, mean/std of |negative null ES|,
 * fraction positive, over permutations 1..P-1 of es[P][G] */
int gsea_null_moments(const float *d_es, int64_t num_perms, int64_t num_sets,
                      float *d_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
