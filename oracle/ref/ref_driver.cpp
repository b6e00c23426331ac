// ref_driver.cpp -- exposes the reference's own CPU arithmetic through a C
// ABI so the tests can pin the oracle against it.  TEST INFRASTRUCTURE ONLY.
//
// It #includes the reference headers where they lie under
// /root/reference/cudaGSEA/src/include (see oracle/ref/Makefile); nothing of
// the reference is copied here.  Only headers that compile on their own are
// used: rngpu/rngpu.hpp (RNG + shuffle), functors.cuh + math_helpers.cuh (the
// 14 metric functors, including glibc logf/log), statistics.cuh
// (final_statistics) and bitmap_types.cuh.  correlate_genes.cuh,
// batched_sort.cuh and enrichment_scores.cuh pull in the CUDA runtime or CUB
// and cannot be built here; the 10-line permutation loop of correlate_cpu
// (correlate_genes.cuh:91-114) is therefore restated below around the
// reference functors.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "rngpu/rngpu.hpp"
#include "functors.cuh"
#include "statistics.cuh"
#include "bitmap_types.cuh"

namespace {

template <class value_t, class funct_t>
void run_perms(funct_t accum, const value_t *X, const uint8_t *labels,
               int64_t N, int64_t S, int64_t P, int64_t shift,
               size_t nA, size_t nB, value_t *out) {
    std::vector<unsigned char> sigma(S);
    for (int64_t perm = 0; perm < P; ++perm) {
        for (int64_t i = 0; i < S; ++i) sigma[i] = labels[i];
        if (perm + shift > 0) {
            auto state = get_initial_fast_kiss_state32(perm + shift + 42);
            fisher_yates_shuffle(fast_kiss32, &state, sigma.data(), (size_t)S);
        }
        for (int64_t g = 0; g < N; ++g)
            out[perm * N + g] = accum(sigma.data(), const_cast<value_t *>(X),
                                      (size_t)S, (size_t)g, (size_t)N, nA, nB);
    }
}

// metric ids follow the dispatch order of correlate_genes.cuh:342-457
template <class value_t>
int correlate(const value_t *X, const uint8_t *labels, int64_t N, int64_t S,
              int64_t P, int64_t shift, int metric, value_t *out) {
    size_t nA = 0, nB = 0;
    for (int64_t i = 0; i < S; ++i) { nA += labels[i] == 0; nB += labels[i] == 1; }
    constexpr bool tr = false;   // gene-major table, as compute_gsea_cpu uses
    switch (metric) {
    case 0: run_perms(accumulate_naive_mean<combine_difference, tr>(combine_difference()), X, labels, N, S, P, shift, nA, nB, out); break;
    case 1: run_perms(accumulate_naive_mean<combine_quotient, tr>(combine_quotient()), X, labels, N, S, P, shift, nA, nB, out); break;
    case 2: run_perms(accumulate_naive_mean<combine_log2_quotient, tr>(combine_log2_quotient()), X, labels, N, S, P, shift, nA, nB, out); break;
    case 3: run_perms(accumulate_kahan_mean<combine_difference, tr>(combine_difference()), X, labels, N, S, P, shift, nA, nB, out); break;
    case 4: run_perms(accumulate_kahan_mean<combine_quotient, tr>(combine_quotient()), X, labels, N, S, P, shift, nA, nB, out); break;
    case 5: run_perms(accumulate_kahan_mean<combine_log2_quotient, tr>(combine_log2_quotient()), X, labels, N, S, P, shift, nA, nB, out); break;
    case 6: run_perms(accumulate_onepass_stats<combine_signal2noise<true>, tr>(combine_signal2noise<true>()), X, labels, N, S, P, shift, nA, nB, out); break;
    case 7: run_perms(accumulate_onepass_stats<combine_T_test<true>, tr>(combine_T_test<true>()), X, labels, N, S, P, shift, nA, nB, out); break;
    case 8: run_perms(accumulate_twopass_stats<combine_signal2noise<true>, tr>(combine_signal2noise<true>()), X, labels, N, S, P, shift, nA, nB, out); break;
    case 9: run_perms(accumulate_twopass_stats<combine_T_test<true>, tr>(combine_T_test<true>()), X, labels, N, S, P, shift, nA, nB, out); break;
    case 10: run_perms(accumulate_knuth_stats<combine_signal2noise<true>, tr>(combine_signal2noise<true>()), X, labels, N, S, P, shift, nA, nB, out); break;
    case 11: run_perms(accumulate_knuth_stats<combine_T_test<true>, tr>(combine_T_test<true>()), X, labels, N, S, P, shift, nA, nB, out); break;
    case 12: run_perms(accumulate_overkill_stats<combine_signal2noise<true>, tr>(combine_signal2noise<true>()), X, labels, N, S, P, shift, nA, nB, out); break;
    case 13: run_perms(accumulate_overkill_stats<combine_T_test<true>, tr>(combine_T_test<true>()), X, labels, N, S, P, shift, nA, nB, out); break;
    default: return 1;
    }
    return 0;
}

}  // namespace

extern "C" {

void ref_permute_labels(const uint8_t *labels, int64_t S, uint64_t perm,
                        uint8_t *out) {
    std::memcpy(out, labels, (size_t)S);
    if (perm == 0) return;
    auto state = get_initial_fast_kiss_state32((uint32_t)(perm + 42));
    fisher_yates_shuffle(fast_kiss32, &state, out, (size_t)S);
}

uint32_t ref_hash32(uint32_t x) { return nvidia_hash(x); }

void ref_kiss_stream(uint32_t seed, uint32_t *out, int64_t n) {
    auto state = get_initial_fast_kiss_state32(seed);
    for (int64_t i = 0; i < n; ++i) out[i] = fast_kiss32(&state);
}

int ref_correlate_f32(const float *X, const uint8_t *labels, int64_t N,
                      int64_t S, int64_t P, int64_t shift, int metric,
                      float *out) {
    return correlate<float>(X, labels, N, S, P, shift, metric, out);
}

int ref_correlate_f64(const double *X, const uint8_t *labels, int64_t N,
                      int64_t S, int64_t P, int64_t shift, int metric,
                      double *out) {
    return correlate<double>(X, labels, N, S, P, shift, metric, out);
}

// statistics.cuh:46-117; es[G][P] is normalised in place, out[5][G]
void ref_final_statistics_f32(float *es, int64_t G, int64_t P, float *out) {
    std::vector<float> v(es, es + G * P);
    auto r = final_statistics<float, size_t>(v, (size_t)G, (size_t)P);
    std::memcpy(out, r.data(), sizeof(float) * 5 * G);
    std::memcpy(es, v.data(), sizeof(float) * G * P);
}

void ref_final_statistics_f64(double *es, int64_t G, int64_t P, double *out) {
    std::vector<double> v(es, es + G * P);
    auto r = final_statistics<double, size_t>(v, (size_t)G, (size_t)P);
    std::memcpy(out, r.data(), sizeof(double) * 5 * G);
    std::memcpy(es, v.data(), sizeof(double) * G * P);
}

int ref_bitmap32_getbit(uint32_t word, int bit) {
    bitmap32_t b; b.x = word; return b.getbit(bit);
}

}  // extern "C"
