/*
 * gsea_oracle.h -- CPU restatement of the reference permutation-GSEA path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it; the product path (cudagsea_amd/) never does.
 *
 * Every function restates one piece of the reference CPU path
 * (cudaGSEA/src/include/cugsea.cuh:267-453, compute_gsea_cpu) and cites the
 * file:line it follows.  Deviations, each stated where it happens:
 *   - argsort uses a stated total order (descending metric, ties by ascending
 *     gene index, -0 below +0, every NaN canonicalised to +NaN which sorts
 *     above +inf) instead of the reference's unstable std::sort
 *     (batched_sort.cuh:309-317);
 *   - the two *_log2_ratio_of_classes metrics use gsea_log_f64/f32 below
 *     (a self-written logarithm, the same one the HIP kernels use) instead of
 *     glibc's logf/log (math_helpers.cuh:38-45);
 *   - FDR is defined by this project (the reference writes NaN,
 *     statistics.cuh:112-114).
 */
#ifndef GSEA_ORACLE_H
#define GSEA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG: rngpu/rngpu.hpp:27-38,145-178,288-303 ---- */
uint32_t orc_hash32(uint32_t x);
void     orc_kiss_init(uint32_t seed, uint32_t state[5]);
uint32_t orc_kiss_next(uint32_t state[5]);
/* Label vector of global permutation `perm` (0 = observed labels);
 * correlate_genes.cuh:94-103, seed = perm + 42 (configuration.cuh:29). */
void     orc_permute_labels(const uint8_t *labels, int64_t S, uint64_t perm,
                            uint8_t *out);

/* ---- logarithm used by the log2-ratio metrics on both CPU and GPU ---- */
double   orc_log_f64(double x);
float    orc_log_f32(float x);

/* ---- local ranking metrics: functors.cuh:10-433, correlate_genes.cuh ---- */
int      orc_metric_id(const char *name);       /* -1 if unknown */
const char *orc_metric_name(int id);
/* X is gene-major [N][S] if sample_major == 0 (reference CPU path,
 * correlate_genes.cuh:302-462 with transposed=false), else sample-major
 * [S][N] (R path, cudaGSEA_wrapper.cu:229-236).  Output correl[P][N]
 * for global permutations shift .. shift+P-1. */
void orc_correlate_f32(const float *X, const uint8_t *labels, int64_t N,
                       int64_t S, int64_t P, int64_t shift, int metric,
                       int sample_major, float *correl);
void orc_correlate_f64(const double *X, const uint8_t *labels, int64_t N,
                       int64_t S, int64_t P, int64_t shift, int metric,
                       int sample_major, double *correl);

/* ---- argsort: batched_sort.cuh:271-326 with the stated tie rule ---- */
uint32_t orc_sort_key_f32(float v, int ascending);
uint64_t orc_sort_key_f64(double v, int ascending);
/* keys[P][N] are permuted in place into rank order, index[P][N] receives the
 * gene id at each rank. */
void orc_pathway_sort_f32(float *keys, uint32_t *index, int64_t N, int64_t P,
                          int ascending);
void orc_pathway_sort_f64(double *keys, uint32_t *index, int64_t N, int64_t P,
                          int ascending);

/* ---- enrichment scores: enrichment_scores.cuh:259-336 (kahan=true) ----
 * Sets are given as CSR (set_offsets[G+1], set_genes = unique gene ids).
 * es[G][P] set-major, the reference's global_result layout. */
void orc_scores_f32(const float *sorted, const uint32_t *index,
                    const int32_t *set_offsets, const int32_t *set_genes,
                    int64_t G, int64_t N, int64_t P, float *es);
void orc_scores_f64(const double *sorted, const uint32_t *index,
                    const int32_t *set_offsets, const int32_t *set_genes,
                    int64_t G, int64_t N, int64_t P, double *es);

/* The same walks, also returning each walk's largest and smallest running
 * sum (0 when the walk never goes above / below 0): es is whichever of the two
 * has the larger magnitude, the first to occur on equal magnitudes.  Test
 * infrastructure: where two implementations round these two candidates
 * differently, a near-equal pair may flip the sign of es. */
void orc_scores_ext_f32(const float *sorted, const uint32_t *index,
                        const int32_t *set_offsets, const int32_t *set_genes,
                        int64_t G, int64_t N, int64_t P, float *es,
                        float *es_hi, float *es_lo);
void orc_scores_ext_f64(const double *sorted, const uint32_t *index,
                        const int32_t *set_offsets, const int32_t *set_genes,
                        int64_t G, int64_t N, int64_t P, double *es,
                        double *es_hi, double *es_lo);

/* ---- statistics: statistics.cuh:46-117 (+ FDR, defined here) ----
 * es[G][P] is normalised in place (as the reference does); out[5][G] =
 * ES, NES, NP, FWER, FDR. */
void orc_final_statistics_f32(float *es, int64_t G, int64_t P, float *out);
void orc_final_statistics_f64(double *es, int64_t G, int64_t P, double *out);
/* flags bit2: signed NES (Broad GSEA's per-sign normalisation by the mean of
 * the null scores of each sign, permutations 1..P-1; defined here, the
 * reference has the single factor only). */
void orc_final_statistics_ext_f32(float *es, int64_t G, int64_t P, int flags,
                                  float *out);
void orc_final_statistics_ext_f64(double *es, int64_t G, int64_t P, int flags,
                                  double *out);

/* ---- whole path: cugsea.cuh:267-453 ----
 * flags bit0: ascending sort (reference sort_direction=false),
 *       bit1: swap phenotype labels (cudaGSEA_wrapper.cu:246-251),
 *       bit2: signed NES (as orc_final_statistics_ext).
 * es_out[G][P] (may be NULL) receives the raw ES before normalisation.
 * Returns 0, or 1 for an unknown metric, or 21 for a bad argument. */
int orc_gsea_f32(const float *X, int64_t N, int64_t S, const uint8_t *labels,
                 const int32_t *set_offsets, const int32_t *set_genes,
                 int64_t G, int64_t P, int metric, int flags,
                 int sample_major, float *es_out, float *stats_out);
int orc_gsea_f64(const double *X, int64_t N, int64_t S, const uint8_t *labels,
                 const int32_t *set_offsets, const int32_t *set_genes,
                 int64_t G, int64_t P, int metric, int flags,
                 int sample_major, double *es_out, double *stats_out);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
