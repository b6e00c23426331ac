/*
 * gsea_oracle.c -- CPU restatement of the reference permutation-GSEA path.
 * TEST INFRASTRUCTURE ONLY (the checker); see gsea_oracle.h.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off, no fast-math),
 * so float and double arithmetic is plain IEEE, evaluated in source order
 * with no contraction, as in the reference's host build.
 *
 * Precision-generic code lives in oracle_impl.inc, included twice below.
 */
#include "gsea_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------
 * RNG.  Follows rngpu/rngpu.hpp:27-38 (hash used by the seeding,
 * `hash32` aliases it at :17), :145-160 (fast KISS step), :162-178 (seeding
 * with 8 rounds of 4 hashes) and :288-303 (Fisher-Yates).
 * ---------------------------------------------------------------------- */
uint32_t orc_hash32(uint32_t x) {
    x = (x + 0x7ed55d16u) + (x << 12);
    x = (x ^ 0xc761c23cu) ^ (x >> 19);
    x = (x + 0x165667b1u) + (x << 5);
    x = (x + 0xd3a2646cu) ^ (x << 9);
    x = (x + 0xfd7046c5u) + (x << 3);
    x = (x ^ 0xb55a4f09u) ^ (x >> 16);
    return x;
}

/* state = {x, y, z, w, c} */
void orc_kiss_init(uint32_t seed, uint32_t st[5]) {
    uint32_t w = seed ? seed : 0xffffffffu;
    uint32_t x = 0, y = 0, z = 0;
    for (int it = 0; it < 8; ++it) {
        x = orc_hash32(w);
        y = orc_hash32(x);
        z = orc_hash32(y);
        w = orc_hash32(z);
    }
    st[0] = x; st[1] = y; st[2] = z; st[3] = w; st[4] = 0;
}

uint32_t orc_kiss_next(uint32_t st[5]) {
    st[1] ^= st[1] << 5;
    st[1] ^= st[1] >> 7;
    st[1] ^= st[1] << 22;
    /* the reference adds in uint32 and tests the sign of the int32 result */
    uint32_t t = st[2] + st[3] + st[4];
    st[2] = st[3];
    st[4] = (t >> 31) & 1u;
    st[3] = t & 0x7fffffffu;
    st[0] += 1411392427u;
    return st[0] + st[1] + st[3];
}

void orc_permute_labels(const uint8_t *labels, int64_t S, uint64_t perm,
                        uint8_t *out) {
    memcpy(out, labels, (size_t)S);
    if (perm == 0) return;
    uint32_t st[5];
    /* correlate_genes.cuh:101: seed = perm + shift + 42, truncated to 32 bit */
    orc_kiss_init((uint32_t)(perm + 42u), st);
    for (int64_t i = S - 1; i >= 1; --i) {
        /* rngpu.hpp:298: (double)rng * (i+1) / 2^32, truncated */
        double d = (double)orc_kiss_next(st) * (double)(i + 1) / 4294967296.0;
        int64_t r = (int64_t)d;
        uint8_t tmp = out[i];
        out[i] = out[r];
        out[r] = tmp;
    }
}

/* ------------------------------------------------------------------------
 * Logarithm for the log2-ratio metrics.  Written for this project from the
 * series log(1+f) = 2 atanh(s), s = f/(2+f), after reducing x = m * 2^e with
 * m in [sqrt(1/2), sqrt(2)); ln 2 is split into a 32-bit head and a tail
 * (the usual Cody-Waite split).  Only IEEE +,-,*,/ in a fixed order, so the
 * HIP kernels evaluate exactly the same sequence and agree bit for bit.
 * ---------------------------------------------------------------------- */
static const double kLn2Hi = 6.93147180369123816490e-01;
static const double kLn2Lo = 1.90821492927058770002e-10;

double orc_log_f64(double x) {
    if (x != x) return (double)NAN;
    if (x < 0.0) return (double)NAN;
    if (x == 0.0) return -(double)INFINITY;
    if (x == (double)INFINITY) return x;
    uint64_t u;
    memcpy(&u, &x, 8);
    int e = 0;
    if ((u >> 52) == 0) {                   /* subnormal: scale by 2^54 */
        x = x * 18014398509481984.0;
        memcpy(&u, &x, 8);
        e = -54;
    }
    e += (int)((u >> 52) & 0x7ffu) - 1023;
    u = (u & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    memcpy(&m, &u, 8);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 2.0 / 23.0;
    R = 2.0 / 21.0 + z * R;
    R = 2.0 / 19.0 + z * R;
    R = 2.0 / 17.0 + z * R;
    R = 2.0 / 15.0 + z * R;
    R = 2.0 / 13.0 + z * R;
    R = 2.0 / 11.0 + z * R;
    R = 2.0 / 9.0 + z * R;
    R = 2.0 / 7.0 + z * R;
    R = 2.0 / 5.0 + z * R;
    R = 2.0 / 3.0 + z * R;
    R = z * R;
    double lm = f - s * (f - R);
    double de = (double)e;
    return de * kLn2Hi + (lm + de * kLn2Lo);
}

float orc_log_f32(float x) {
    if (x != x) return (float)NAN;
    return (float)orc_log_f64((double)x);
}

/* metric names in the reference's dispatch order (correlate_genes.cuh:342-457) */
static const char *kMetricNames[14] = {
    "naive_diff_of_classes",  "naive_ratio_of_classes",
    "naive_log2_ratio_of_classes", "stable_diff_of_classes",
    "stable_ratio_of_classes", "stable_log2_ratio_of_classes",
    "onepass_signal2noise",   "onepass_t_test",
    "twopass_signal2noise",   "twopass_t_test",
    "stable_signal2noise",    "stable_t_test",
    "overkill_signal2noise",  "overkill_t_test"};

int orc_metric_id(const char *name) {
    if (!name) return -1;
    for (int i = 0; i < 14; ++i)
        if (strcmp(name, kMetricNames[i]) == 0) return i;
    return -1;
}

const char *orc_metric_name(int id) {
    return (id >= 0 && id < 14) ? kMetricNames[id] : "";
}

/* sort keys: ascending unsigned order == the stated total order */
uint32_t orc_sort_key_f32(float v, int ascending) {
    uint32_t u;
    if (v != v) u = 0x7fc00000u; else memcpy(&u, &v, 4);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return ascending ? u : ~u;
}

uint64_t orc_sort_key_f64(double v, int ascending) {
    uint64_t u;
    if (v != v) u = 0x7ff8000000000000ull; else memcpy(&u, &v, 8);
    u = (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
    return ascending ? u : ~u;
}

/* ---- precision-generic parts ---- */
#define T float
#define SFX f32
#define ORC_LOG orc_log_f32
#define ORC_KEY_T uint32_t
#define ORC_KEY orc_sort_key_f32
#define ORC_SQRT sqrtf
#include "oracle_impl.inc"
#undef T
#undef SFX
#undef ORC_LOG
#undef ORC_KEY_T
#undef ORC_KEY
#undef ORC_SQRT

#define T double
#define SFX f64
#define ORC_LOG orc_log_f64
#define ORC_KEY_T uint64_t
#define ORC_KEY orc_sort_key_f64
#define ORC_SQRT sqrt
#include "oracle_impl.inc"
