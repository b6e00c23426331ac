// gsea_kernels.hpp -- CDNA4 (gfx950) HIP kernels of the permutation-GSEA hot
// path:
//   k_class_lists  permutation p -> ordered sample lists of both classes
//   k_metric       14 local ranking metrics, many permutations per X tile
//   k_rank         per permutation: LDS radix argsort of all genes -> rank rows
//   k_es           per permutation: running-sum enrichment score of every set
//   k_set_stats / k_perm_extrema / k_fwer / k_fdr_*   final statistics
//   k_transpose    [P][G] -> [G][P] for the reference's ES dump layout
#pragma once

#include "gsea_device.hpp"

namespace gsea {

// ===========================================================================
// k_class_lists: one thread per permutation.  Labels of global permutation
// gp are the observed labels (gp == 0) or their Fisher-Yates shuffle seeded
// with gp + 42 (correlate_genes.cuh:44-55, rngpu.hpp:288-303).  The kernel
// writes, per permutation, the samples of class A (label 0, functors.cuh:110)
// in increasing sample order at [0, nA) and all other samples in increasing
// order at [offB, offB + S - nA): the metric kernels then visit each class in
// exactly the reference's summation order.
// ===========================================================================
__global__ void k_class_lists(const uint8_t *__restrict__ labels, int S,
                              int64_t perm_base, int Pb,
                              uint8_t *__restrict__ sig, int offB,
                              uint16_t *__restrict__ lists, int lstride) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Pb) return;
    uint8_t *s = sig + (size_t)p * S;
    for (int i = 0; i < S; ++i) s[i] = labels[i];
    const uint64_t gp = (uint64_t)(perm_base + p);
    if (gp) {
        KissState st = kiss_seed((uint32_t)(gp + 42u));
        for (int i = S - 1; i >= 1; --i) {
            // (double)rng * (i+1) / 2^32 truncated == (rng * (i+1)) >> 32
            const uint32_t r = (uint32_t)(((uint64_t)kiss_next(st) *
                                           (uint64_t)(i + 1)) >> 32);
            const uint8_t t = s[i];
            s[i] = s[r];
            s[r] = t;
        }
    }
    uint16_t *L = lists + (size_t)p * lstride;
    int a = 0, b = offB;
    for (int i = 0; i < S; ++i) {
        if (s[i] == 0) L[a++] = (uint16_t)i;
        else L[b++] = (uint16_t)i;
    }
}

// ===========================================================================
// k_metric: local ranking metrics (functors.cuh:10-433).
//
// Block = 4 wavefronts.  The block stages X[0:S][g0:g0+GT] (sample-major
// input) in LDS once and streams `ppb` permutations over it.  A wavefront is
// split into lane groups of GT/GPL lanes; lane `col` of a group owns genes
// g0+col*GPL .. +GPL-1 (GPL = 4 fp32 or 2 fp64 genes: one 16-byte LDS read
// per sample serves them all, so the class-list decoding and addressing is
// shared), and
// each lane carries PL permutations at once (independent accumulation
// chains).  Every class sum walks that class's samples in increasing sample
// order, one sample at a time, with the reference's exact arithmetic, so the
// metric is bit-identical to the reference CPU path.
// ===========================================================================
enum AccKind { kNaive = 0, kKahan = 1, kOnepass = 2, kTwopass = 3,
               kKnuth = 4, kOverkill = 5 };

__host__ __device__ constexpr int acc_kind(int metric) {
    return metric <= 2 ? kNaive
         : metric <= 5 ? kKahan
         : 2 + (metric - 6) / 2;
}

template <class T>
struct MetricArgs {
    const T *X;            // [S][N] sample-major
    int N, S;
    const uint16_t *lists; // [Pb][lstride]
    int lstride, nA, nB, offB;   // nA, nB: divisors (label 0 / 1 counts)
    int lenA;                    // samples with label 0 (class-A list)
    int Pb, ppb;
    int metric;            // 0..13, selects the combiner
    T *r;                  // [Pb][N]
    const double *rcp;     // [S]: 1/(i+1) in double (Welford running counts)
};

// visit list entries [off, off+cnt) of PL permutations in order; f(k, s, i)
template <int PL, class F>
__device__ __forceinline__ void walk_lists(const uint16_t *const *L, int off,
                                           int cnt, F &&f) {
    int i = 0;
    for (; i + 8 <= cnt; i += 8) {
        uint4 w[PL];
#pragma unroll
        for (int k = 0; k < PL; ++k)
            w[k] = *reinterpret_cast<const uint4 *>(L[k] + off + i);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
#pragma unroll
            for (int k = 0; k < PL; ++k) {
                const uint32_t word = j < 2 ? w[k].x : j < 4 ? w[k].y
                                    : j < 6 ? w[k].z : w[k].w;
                const int s = (j & 1) ? (int)(word >> 16) : (int)(word & 0xffffu);
                f(k, s, i + j);
            }
        }
    }
    for (; i < cnt; ++i) {
#pragma unroll
        for (int k = 0; k < PL; ++k) f(k, (int)L[k][off + i], i);
    }
}

template <class T>
__device__ __forceinline__ T combine_means(int metric, T mA, T mB) {
    const int c = metric % 3;
    if (c == 0) return mA - mB;                       // functors.cuh:10-15
    if (c == 1) return mA / mB;                       // functors.cuh:17-22
    return (T)((double)log_t(mA / mB) * 1.44269504089);  // functors.cuh:24-29
}

template <class T>
__device__ __forceinline__ T combine_stats(int metric, T mA, T vA, T nA,
                                           T mB, T vB, T nB) {
    if ((metric & 1) == 0) {                          // functors.cuh:44-51
        T sA = sqrt_t(vA), sB = sqrt_t(vB);
        const T minA = (T)(0.2 * (double)cuabs(mA));
        const T minB = (T)(0.2 * (double)cuabs(mB));
        sA = minA < sA ? sA : minA;
        sB = minB < sB ? sB : minB;
        return (mA - mB) / (sA + sB);
    }
    const T minA = (T)(0.04 * (double)cuabs(mA));     // functors.cuh:70-77
    const T minB = (T)(0.04 * (double)cuabs(mB));
    vA = minA < vA ? vA : minA;
    vB = minB < vB ? vB : minB;
    return (mA - mB) / sqrt_t(vA / nA + vB / nB);
}

// delta / (i+1), correctly rounded in T.  fp32: (double)delta * fl64(1/(i+1))
// rounded once to float is the correctly rounded quotient (the product is
// within 2^-52 of the exact quotient; a quotient of two floats is at least
// 2^-49 away from every float rounding midpoint), special values included.
__device__ __forceinline__ float div_count(float delta, int i, const double *rcp) {
    return (float)((double)delta * rcp[i]);
}
__device__ __forceinline__ double div_count(double delta, int i, const double *) {
    return delta / (double)(i + 1);
}

template <class T, int KIND, int GT, bool XLDS, int GPL>
__global__ __launch_bounds__(256) void k_metric(MetricArgs<T> a) {
    constexpr int PL = 4;
    constexpr int LPG = GT / GPL;     // lanes per group
    constexpr int PPW = kWave / LPG;  // lane groups per wavefront
    static_assert(GPL == 1 || (XLDS && (GPL == 2 || GPL == 4) && sizeof(T) * GPL <= 16),
                  "several genes per lane read one 8/16-byte group from the LDS tile");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *tile = reinterpret_cast<T *>(smem_raw);

    const int g0 = blockIdx.x * GT;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int grp = lane / LPG, col = lane % LPG;
    const int g = g0 + col * GPL;     // first gene of this lane
    const int gc = g < a.N ? g : a.N - 1;

    const T *xs;
    int xstride;
    if (XLDS) {
        for (int idx = threadIdx.x; idx < a.S * GT; idx += blockDim.x) {
            const int s = idx / GT, j = idx % GT;
            const int gg = g0 + j;
            tile[idx] = gg < a.N ? a.X[(size_t)s * a.N + gg] : T(0);
        }
        __syncthreads();
        xs = tile + col * GPL;
        xstride = GT;
    } else {
        xs = a.X + gc;
        xstride = a.N;
    }
    // the GPL values of sample s (one 8- or 16-byte LDS read for 2 or 4 fp32
    // genes)
    auto X_at = [&](int s, T (&v)[GPL]) {
        if constexpr (GPL == 4) {
            const float4 p = *reinterpret_cast<const float4 *>(xs + (size_t)s * xstride);
            v[0] = p.x;
            v[1] = p.y;
            v[2] = p.z;
            v[3] = p.w;
        } else if constexpr (GPL == 2 && sizeof(T) == 8) {
            const double2 p = *reinterpret_cast<const double2 *>(xs + (size_t)s * xstride);
            v[0] = p.x;
            v[1] = p.y;
        } else if constexpr (GPL == 2) {
            const float2 p = *reinterpret_cast<const float2 *>(xs + (size_t)s * xstride);
            v[0] = p.x;
            v[1] = p.y;
        } else {
            v[0] = xs[(size_t)s * xstride];
        }
    };

    // naive sums see 0*x of the other class (functors.cuh:115-116), which
    // turns into NaN for a non-finite x: track whether the gene has one
    bool nonfinite[GPL];
    bool any_nonfinite = false;
#pragma unroll
    for (int j = 0; j < GPL; ++j) nonfinite[j] = false;
    if (KIND == kNaive) {
        for (int s = 0; s < a.S; ++s) {
            T v[GPL];
            X_at(s, v);
#pragma unroll
            for (int j = 0; j < GPL; ++j) nonfinite[j] |= !(v[j] - v[j] == T(0));
        }
#pragma unroll
        for (int j = 0; j < GPL; ++j) any_nonfinite |= nonfinite[j];
    }

    const T nA = (T)a.nA, nB = (T)a.nB;
    const int nRest = a.S - a.lenA;   // samples whose label is not 0
    const int p_begin = blockIdx.y * a.ppb;
    const int p_end = min(a.Pb, p_begin + a.ppb);
    const int slot = wave * PPW + grp;
    const int step = 4 * PPW * PL;

    for (int base = p_begin + slot * PL; base < p_end; base += step) {
        const uint16_t *L[PL];
        int pk[PL];
#pragma unroll
        for (int k = 0; k < PL; ++k) {
            pk[k] = base + k;
            const int pe = pk[k] < p_end ? pk[k] : p_end - 1;
            L[k] = a.lists + (size_t)pe * a.lstride;
        }
        T out[PL][GPL];
        if (KIND == kNaive || KIND == kOnepass || KIND == kTwopass) {
            T sA[PL][GPL], sB[PL][GPL], qA[PL][GPL], qB[PL][GPL];
#pragma unroll
            for (int k = 0; k < PL; ++k)
#pragma unroll
                for (int j = 0; j < GPL; ++j) sA[k][j] = sB[k][j] = qA[k][j] = qB[k][j] = T(0);
            walk_lists<PL>(L, 0, a.lenA, [&](int k, int s, int) {
                T v[GPL];
                X_at(s, v);
#pragma unroll
                for (int j = 0; j < GPL; ++j) {
                    sA[k][j] = sA[k][j] + v[j];
                    if (KIND == kOnepass) qA[k][j] = qA[k][j] + v[j] * v[j];
                }
            });
            walk_lists<PL>(L, a.offB, nRest, [&](int k, int s, int) {
                T v[GPL];
                X_at(s, v);
#pragma unroll
                for (int j = 0; j < GPL; ++j) {
                    sB[k][j] = sB[k][j] + v[j];
                    if (KIND == kOnepass) qB[k][j] = qB[k][j] + v[j] * v[j];
                }
            });
            if (KIND == kNaive && any_nonfinite) {
                bool bA[PL][GPL], bB[PL][GPL];
#pragma unroll
                for (int k = 0; k < PL; ++k)
#pragma unroll
                    for (int j = 0; j < GPL; ++j) bA[k][j] = bB[k][j] = false;
                walk_lists<PL>(L, 0, a.lenA, [&](int k, int s, int) {
                    T v[GPL];
                    X_at(s, v);
#pragma unroll
                    for (int j = 0; j < GPL; ++j) bA[k][j] |= !(v[j] - v[j] == T(0));
                });
                walk_lists<PL>(L, a.offB, nRest, [&](int k, int s, int) {
                    T v[GPL];
                    X_at(s, v);
#pragma unroll
                    for (int j = 0; j < GPL; ++j) bB[k][j] |= !(v[j] - v[j] == T(0));
                });
#pragma unroll
                for (int k = 0; k < PL; ++k)
#pragma unroll
                    for (int j = 0; j < GPL; ++j) {
                        if (bB[k][j]) sA[k][j] = (T)__builtin_nan("");
                        if (bA[k][j]) sB[k][j] = (T)__builtin_nan("");
                    }
            }
            T mA[PL][GPL], mB[PL][GPL];
#pragma unroll
            for (int k = 0; k < PL; ++k)
#pragma unroll
                for (int j = 0; j < GPL; ++j) { mA[k][j] = sA[k][j] / nA; mB[k][j] = sB[k][j] / nB; }
            if (KIND == kNaive) {
#pragma unroll
                for (int k = 0; k < PL; ++k)
#pragma unroll
                    for (int j = 0; j < GPL; ++j)
                        out[k][j] = combine_means<T>(a.metric, mA[k][j], mB[k][j]);
            } else if (KIND == kOnepass) {
                // functors.cuh:212-218 (biased)
#pragma unroll
                for (int k = 0; k < PL; ++k)
#pragma unroll
                    for (int j = 0; j < GPL; ++j) {
                        const T vA = qA[k][j] / nA - mA[k][j] * mA[k][j];
                        const T vB = qB[k][j] / nB - mB[k][j] * mB[k][j];
                        out[k][j] = combine_stats<T>(a.metric, mA[k][j], vA, nA, mB[k][j],
                                                     vB, nB);
                    }
            } else {
                // functors.cuh:270-290
                walk_lists<PL>(L, 0, a.lenA, [&](int k, int s, int) {
                    T v[GPL];
                    X_at(s, v);
#pragma unroll
                    for (int j = 0; j < GPL; ++j) {
                        const T d = v[j] - mA[k][j];
                        qA[k][j] = qA[k][j] + d * d;
                    }
                });
                walk_lists<PL>(L, a.offB, nRest, [&](int k, int s, int) {
                    T v[GPL];
                    X_at(s, v);
#pragma unroll
                    for (int j = 0; j < GPL; ++j) {
                        const T d = v[j] - mB[k][j];
                        qB[k][j] = qB[k][j] + d * d;
                    }
                });
#pragma unroll
                for (int k = 0; k < PL; ++k)
#pragma unroll
                    for (int j = 0; j < GPL; ++j)
                        out[k][j] = combine_stats<T>(a.metric, mA[k][j], qA[k][j] / nA, nA,
                                                     mB[k][j], qB[k][j] / nB, nB);
            }
        } else if (KIND == kKahan || KIND == kOverkill) {
            // Kahan mean, functors.cuh:144-172 / 316-343
            T sA[PL][GPL], sB[PL][GPL], cA[PL][GPL], cB[PL][GPL];
#pragma unroll
            for (int k = 0; k < PL; ++k)
#pragma unroll
                for (int j = 0; j < GPL; ++j) sA[k][j] = sB[k][j] = cA[k][j] = cB[k][j] = T(0);
            walk_lists<PL>(L, 0, a.lenA, [&](int k, int s, int) {
                T v[GPL];
                X_at(s, v);
#pragma unroll
                for (int j = 0; j < GPL; ++j) {
                    const T alpha = v[j] - cA[k][j];
                    const T beta = sA[k][j] + alpha;
                    cA[k][j] = (beta - sA[k][j]) - alpha;
                    sA[k][j] = beta;
                }
            });
            walk_lists<PL>(L, a.offB, nRest, [&](int k, int s, int) {
                T v[GPL];
                X_at(s, v);
#pragma unroll
                for (int j = 0; j < GPL; ++j) {
                    const T alpha = v[j] - cB[k][j];
                    const T beta = sB[k][j] + alpha;
                    cB[k][j] = (beta - sB[k][j]) - alpha;
                    sB[k][j] = beta;
                }
            });
            T mA[PL][GPL], mB[PL][GPL];
#pragma unroll
            for (int k = 0; k < PL; ++k)
#pragma unroll
                for (int j = 0; j < GPL; ++j) { mA[k][j] = sA[k][j] / nA; mB[k][j] = sB[k][j] / nB; }
            if (KIND == kKahan) {
#pragma unroll
                for (int k = 0; k < PL; ++k)
#pragma unroll
                    for (int j = 0; j < GPL; ++j)
                        out[k][j] = combine_means<T>(a.metric, mA[k][j], mB[k][j]);
            } else {
                // compensated second pass, functors.cuh:346-370
                T uA[PL][GPL], uB[PL][GPL], vA[PL][GPL], vB[PL][GPL];
#pragma unroll
                for (int k = 0; k < PL; ++k)
#pragma unroll
                    for (int j = 0; j < GPL; ++j) uA[k][j] = uB[k][j] = vA[k][j] = vB[k][j] = T(0);
                walk_lists<PL>(L, 0, a.lenA, [&](int k, int s, int) {
                    T v[GPL];
                    X_at(s, v);
#pragma unroll
                    for (int j = 0; j < GPL; ++j) {
                        const T d = v[j] - mA[k][j];
                        uA[k][j] = uA[k][j] + d;
                        vA[k][j] = vA[k][j] + d * d;
                    }
                });
                walk_lists<PL>(L, a.offB, nRest, [&](int k, int s, int) {
                    T v[GPL];
                    X_at(s, v);
#pragma unroll
                    for (int j = 0; j < GPL; ++j) {
                        const T d = v[j] - mB[k][j];
                        uB[k][j] = uB[k][j] + d;
                        vB[k][j] = vB[k][j] + d * d;
                    }
                });
#pragma unroll
                for (int k = 0; k < PL; ++k)
#pragma unroll
                    for (int j = 0; j < GPL; ++j) {
                        const T va = vA[k][j] - uA[k][j] * uA[k][j] / nA;
                        const T vb = vB[k][j] - uB[k][j] * uB[k][j] / nB;
                        out[k][j] = combine_stats<T>(a.metric, mA[k][j], va / nA, nA,
                                                     mB[k][j], vb / nB, nB);
                    }
            }
        } else {
            // Welford/Knuth, functors.cuh:396-428: the i-th sample of a class
            // divides by its running count i+1
            T mA[PL][GPL], mB[PL][GPL], vA[PL][GPL], vB[PL][GPL];
#pragma unroll
            for (int k = 0; k < PL; ++k)
#pragma unroll
                for (int j = 0; j < GPL; ++j) mA[k][j] = mB[k][j] = vA[k][j] = vB[k][j] = T(0);
            walk_lists<PL>(L, 0, a.lenA, [&](int k, int s, int i) {
                T v[GPL];
                X_at(s, v);
#pragma unroll
                for (int j = 0; j < GPL; ++j) {
                    const T delta = v[j] - mA[k][j];
                    mA[k][j] = mA[k][j] + div_count(delta, i, a.rcp);
                    vA[k][j] = vA[k][j] + delta * (v[j] - mA[k][j]);
                }
            });
            walk_lists<PL>(L, a.offB, nRest, [&](int k, int s, int i) {
                T v[GPL];
                X_at(s, v);
#pragma unroll
                for (int j = 0; j < GPL; ++j) {
                    const T delta = v[j] - mB[k][j];
                    mB[k][j] = mB[k][j] + div_count(delta, i, a.rcp);
                    vB[k][j] = vB[k][j] + delta * (v[j] - mB[k][j]);
                }
            });
#pragma unroll
            for (int k = 0; k < PL; ++k)
#pragma unroll
                for (int j = 0; j < GPL; ++j)
                    out[k][j] = combine_stats<T>(a.metric, mA[k][j], vA[k][j] / nA, nA,
                                                 mB[k][j], vB[k][j] / nB, nB);
        }
#pragma unroll
        for (int k = 0; k < PL; ++k) {
            if (pk[k] >= p_end) continue;
            T *row = a.r + (size_t)pk[k] * a.N;
#pragma unroll
            for (int j = 0; j < GPL; ++j)
                if (g + j < a.N) row[g + j] = out[k][j];
        }
    }
}

// ===========================================================================
// k_rank_es: one workgroup (16 wavefronts) per permutation.
//
// Rank: stable LSD radix argsort in LDS, 8-bit digits.  Keys stay in place
// (indexed by gene); only 16-bit gene ids move.  Each wavefront ranks its
// contiguous segment chunk by chunk: lanes with equal digits are found with
// eight ballots, the lowest such lane bumps the wavefront's LDS counter for
// that digit, so ranks follow element order and each pass is stable.  A
// digit-major scan over (digit, wavefront) gives the scatter offsets.
//
// ES: ranks are inverted (pos[gene]) and |metric| is laid out by rank.  Each
// wavefront takes gene sets from an LDS work counter.  For a set of k genes
// it ORs the k rank bits into a private N-bit bitmap, scans the bitmap in
// lane-contiguous word ranges, and evaluates the running sum only at the
// hits: just before hit j at rank q the sum is S_{j-1} + (q-j)*decay, just
// after it S_j + (q-j)*decay, with S_j the prefix sum of |r|/N_R over hits in
// rank order.  The maximum deviation of the reference's walk
// (enrichment_scores.cuh:310-330) is attained at one of these points; the
// first strictly largest |value| in rank order wins, as in the walk.  Sums
// are carried in double and each candidate is rounded to T before it is
// compared, as the reference compares T values.
// ===========================================================================
constexpr int kRankWaves = 16;
constexpr int kRankThreads = kRankWaves * kWave;
// a wavefront ranks at most MAXCH chunks of 64 keys per pass, so one
// workgroup sorts up to 16 * MAXCH * 64 keys (MAXCH <= 32: N <= 32768)
constexpr int kMaxRankN = kRankWaves * 32 * kWave;

template <class T>
struct RankArgs {
    const T *r;              // [Pb][N]
    int N, Pb;
    int64_t perm_base;
    int G;
    const int32_t *set_off;  // [G+1]
    const int32_t *set_genes;
    const uint16_t *large_genes;  // k_es: large sets' gene ids + 1, at large_desc.y
    T *es;                   // [P][G]
    T *a_glob;               // [Pb][N] |metric| by rank when it is not in LDS
    uint32_t *idx_out;       // optional [Pb][N]: gene id at each rank
    uint16_t *pos_out;       // [Pb][N] rank of each gene (k_rank -> k_es)
    int p_lo;                // k_es: first rank row (and permutation) it scores
    int asc;
    int es_waves, nw;        // wavefronts that score sets, bitmap words (64*W)
    size_t off_ids, off_cnt, off_misc;
    // work items: n_large sets (> kSmallSet genes, one wavefront each), then
    // chunks of 64 small sets (one lane each) grouped by size class
    int n_items, n_large;
    const int4 *large_desc;            // [n_large] {set id, offset, size, 0},
                                       // in set order
    const T *large_decay;              // [n_large] decay of those sets
    const int32_t *small_ids;          // [n_chunks*64] set id or -1
    const int32_t *small_k;            // [n_chunks*64] set sizes
    const uint16_t *small_gm;          // [kSmallSet][small_stride] gene ids + 1
    int small_stride;
    int chunk_end8, chunk_end16, chunk_end32, chunk_end64;   // chunk class boundaries
    const T *decay;                    // [G] (T)(-1.0 / (double)(N - k_s))
};

constexpr int kSmallSet = 128;

// k_es reads gene ids shifted by one, as 16-bit values: id 0 is the
// sentinel gene, whose rank is N (past every real rank) and whose |metric| is
// 0.  A load past a set's entries reads 0 (out-of-range buffer loads return
// zero; the gene-id rows of small sets are zero-filled), so it marks the
// sentinel's bit, which is cleared before the walk, and adds 0 to N_R.

// large sets' ids, item i at desc[i].y (one block per item)
__global__ void k_large_genes(const int32_t *__restrict__ off,
                              const int32_t *__restrict__ genes,
                              const int4 *__restrict__ desc,
                              uint16_t *__restrict__ out) {
    const int4 d = desc[blockIdx.x];
    const int32_t *src = genes + off[d.x];
    for (int t = threadIdx.x; t < d.z; t += blockDim.x)
        out[d.y + t] = (uint16_t)(src[t] + 1);
}

// gene ids of the small sets, laid out [t][slot] so a wavefront's 64 lanes
// read row t of 64 consecutive slots; rows t >= k hold the sentinel
__global__ void k_small_gm(const int32_t *__restrict__ off,
                           const int32_t *__restrict__ genes,
                           const int32_t *__restrict__ small_ids, int nslots,
                           uint16_t *__restrict__ gm) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nslots) return;
    const int s = small_ids[slot];
    const int o0 = s >= 0 ? off[s] : 0;
    const int k = s >= 0 ? off[s + 1] - o0 : 0;
    for (int t = 0; t < kSmallSet; ++t)
        gm[(size_t)t * nslots + slot] = (uint16_t)(t < k ? genes[o0 + t] + 1 : 0);
}

template <class T>
struct KeyOf;
template <>
struct KeyOf<float> { typedef uint32_t type; };
template <>
struct KeyOf<double> { typedef uint64_t type; };

// ---- enrichment score of one set from its sorted hit ranks --------------
// The walk of enrichment_scores.cuh:310-330 evaluated at the hits.  With
// c = |r|/N_R formed in T as the reference does (N_R = the set's sum of |r|
// rounded to T), Sa_j = sum of c over hits 0..j (rank order) and m = misses
// before hit j (its rank minus j), the running sum just before hit j is
// Sa_{j-1} + m*decay and just after it Sa_j + m*decay.  Sa and the
// candidates are carried in double and each candidate is rounded to T before
// the first-strictly-largest |value| comparison.  A slot with valid == false
// leaves Sa and best unchanged, so unrolled loops need no branches.
template <class T>
__device__ __forceinline__ void es_candidates(double &Sa, double c, int q, int j,
                                              double decay, T &best,
                                              bool valid = true) {
    const int m = q - j;
    const double md = m ? (double)m * decay : 0.0;   // decay = -inf when k = N
    const T vb = (T)(Sa + md);
    // |x| > |y| compares like the reference's abs(x) > abs(y), -0 and NaN
    // included, and maps to one compare with |.| operand modifiers
    if (valid && fabs(vb) > fabs(best)) best = vb;
    Sa += valid ? c : 0.0;
    const T va = (T)(Sa + md);
    if (valid && fabs(va) > fabs(best)) best = va;
}

// c = |r| / N_R of one set, N_R rounded to T as in the reference.
// operator(): c carried in double as (double)|r| * rn with rn within one ulp
// of 1/N_R, i.e. within 1.5 * 2^-52 (relative) of the exact quotient; the reference rounds c to T, so an
// enrichment score differs from a walk over rounded values by at most 2^-24
// relative per step, i.e. by at most 6e-8 absolute (the c of a set sum to 1).
// rounded(): c rounded to T.  In fp32 that single rounding of the product
// is the correctly rounded quotient the reference's division yields: the
// product is within 1.5 * 2^-52 of the exact quotient, and a quotient of two
// floats is either a float or at least 2^-49 (relative) away from every
// float rounding midpoint (a midpoint has a 25-bit odd significand; times
// the divisor's significand it cannot fit the dividend's 24 bits).  Special
// values follow the division: N_R = 0 gives inf (|r| > 0) or NaN (|r| = 0),
// N_R = inf gives 0 (NaN for |r| = inf).
// fl64(1/n) to within one ulp for finite n > 0: v_rcp_f64 and two Newton
// steps (the error squares per step); 0, inf and NaN take the division
__device__ __forceinline__ double recip(double n) {
    if (n > 0.0 && n < __builtin_inf()) {
        double r = __builtin_amdgcn_rcp(n);
        double e = __builtin_fma(-n, r, 1.0);
        r = __builtin_fma(r, e, r);
        e = __builtin_fma(-n, r, 1.0);
        return __builtin_fma(r, e, r);
    }
    return 1.0 / n;
}

template <class T>
struct HitShare {
    double rn;
    __device__ __forceinline__ explicit HitShare(T n) : rn(recip((double)n)) {}
    __device__ __forceinline__ double operator()(T a) const { return (double)a * rn; }
    __device__ __forceinline__ double rounded(T a) const {
        return (double)(T)((double)a * rn);
    }
};

// leftmost lane value of largest |value| (first strict maximum in lane
// order), returned on every lane.  Lane values here are never NaN or -0.
template <class T>
__device__ __forceinline__ T wave_first_absmax(T x) {
    x = wave_incl_scan_op(x, [](T l, T r) { return fabs(r) > fabs(l) ? r : l; });
    return read_lane(x, kWave - 1);
}

// two 16-bit lanes of one register, compared lane by lane (v_pk_min_u16 /
// v_pk_max_u16)
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_min_u16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(
        __builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_max_u16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(
        __builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}
// low 16 bits of lo, high 16 bits of hi
__device__ __forceinline__ uint32_t halves(uint32_t lo, uint32_t hi) {
    return (lo & 0xffffu) | (hi & 0xffff0000u);
}

// In-register bitonic sort (compile-time network) of NC 16-bit keys, two per
// register: key i is in register i % R, low half for i < R, high half
// otherwise (R = NC/2).  A comparator (i, i ^ stride) with stride < R pairs
// the same halves of two registers, so one packed min and one packed max do
// two comparators; both halves share the direction ((i & size) == 0 means
// ascending) except in the merges of size R, where the low half ascends and
// the high half descends.  stride == R (the first step of the last merge)
// pairs the two halves of one register.
// one step (merge size SIZE, comparator distance STRIDE) of the network,
// then the rest by template recursion (nested loops this deep exceed the
// compiler's full-unroll limit at NC = 128)
template <int NC, int SIZE, int STRIDE>
__device__ __forceinline__ void bitonic_step_u16x2(uint32_t (&v)[NC / 2]) {
    constexpr int R = NC / 2;
    if constexpr (STRIDE == R) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t sw = __builtin_amdgcn_alignbit(v[r], v[r], 16);
            v[r] = halves(pk_min_u16(v[r], sw), pk_max_u16(v[r], sw));
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = r ^ STRIDE;
            if (j > r) {
                const uint32_t mn = pk_min_u16(v[r], v[j]);
                const uint32_t mx = pk_max_u16(v[r], v[j]);
                if (SIZE == R) {
                    v[r] = halves(mn, mx);
                    v[j] = halves(mx, mn);
                } else if ((r & SIZE) == 0) {
                    v[r] = mn;
                    v[j] = mx;
                } else {
                    v[r] = mx;
                    v[j] = mn;
                }
            }
        }
    }
    if constexpr (STRIDE > 1) bitonic_step_u16x2<NC, SIZE, STRIDE / 2>(v);
    else if constexpr (SIZE < NC) bitonic_step_u16x2<NC, SIZE * 2, SIZE>(v);
}

template <int NC>
__device__ __forceinline__ void bitonic_sort_u16x2(uint32_t (&v)[NC / 2]) {
    bitonic_step_u16x2<NC, 2, 1>(v);
}

// Sets of at most NC genes: one lane per set.  The lane gathers its hit
// ranks, sorts them in registers and walks them in rank order.  Rows t >= k
// of the gathered gene ids hold the sentinel: rank N (sorted last, as
// N <= 65535) and |metric| 0 (nothing added to N_R); the walk masks them.
template <class T, int NC, class AF>
__device__ __forceinline__ void es_small_sets(const RankArgs<T> &a,
                                              const uint16_t *pos, AF &A,
                                              int64_t gp, int chunk, int lane) {
    constexpr int R = NC / 2;
    const int slot = chunk * kWave + lane;
    const int sid = a.small_ids[slot];
    const int k = a.small_k[slot];
    uint32_t v[R];
    const uint16_t *gcol = a.small_gm + slot;
    // opaque per call: keeps the compiler from hoisting the row offsets out
    // of the work loop into (spilled) scalar registers
    int stride = a.small_stride;
    asm volatile("" : "+s"(stride));
    // gather in groups of 8 so that only 8 loads (and addresses) are live;
    // N_R is summed here, in set order (a double sum of at most 128 values
    // of T, rounded to T below)
    double nr = 0.0;
#pragma unroll
    for (int t0 = 0; t0 < NC; t0 += 8) {
        int gg[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) gg[u] = gcol[(size_t)(t0 + u) * stride];
        uint32_t g[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) g[u] = pos[gg[u]];
        T x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = A((int)g[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int t = t0 + u;
            if (t < R) v[t] = g[u];
            else v[t - R] |= g[u] << 16;
            nr += (double)x[u];
        }
        // the group's results are final here: its loads, masks and values
        // die before the next group starts
        asm volatile("" : "+v"(nr));
#pragma unroll
        for (int u = 0; u < 8; ++u) asm volatile("" : "+v"(v[(t0 + u) % R]));
    }
    bitonic_sort_u16x2<NC>(v);
    auto rank_at = [&](int t) -> int {
        return t < R ? (int)(v[t] & 0xffffu) : (int)(v[t - R] >> 16);
    };
    const HitShare<T> share((T)nr);
    // opaque keys: the walk recomputes the addresses per group instead of
    // keeping all NC of them live from the pass above (spills)
#pragma unroll
    for (int r = 0; r < R; ++r) asm volatile("" : "+v"(v[r]));

    const double decay = sid >= 0 ? (double)a.decay[sid] : 0.0;
    T best = T(0);
    double Sa = 0.0;
#pragma unroll
    for (int t0 = 0; t0 < NC; t0 += 8) {
        T x[8];
        // an opaque copy of k per group: the compiler would otherwise compute
        // all NC masks t < k at once and keep them live (spills at NC = 128)
        int kk = k;
        asm volatile("" : "+v"(kk));
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = A(rank_at(t0 + u));
#pragma unroll
        for (int u = 0; u < 8; ++u)
            es_candidates<T>(Sa, share.rounded(x[u]), rank_at(t0 + u), t0 + u, decay,
                             best, t0 + u < kk);
        __builtin_amdgcn_sched_barrier(0);
    }
    if (k == 0) best = (T)((double)a.N * decay);
    if (sid >= 0) a.es[(size_t)gp * a.G + sid] = best;
}

// Phase A of a large set: B gene-id slots per lane (set entries j0 + 64u +
// lane, u < B), the bits of their ranks set in the wavefront's bitmap and
// their |r| summed.  rs covers the set's k entries: slots past them read 0,
// the sentinel (rank N, |r| 0).
template <class T, int B, class AF>
__device__ __forceinline__ void es_mark_hits(__amdgpu_buffer_rsrc_t rs, const uint16_t *pos,
                                             AF &A, uint32_t *bm, int j0, int lane,
                                             double &nr) {
    const int vo = 2 * (j0 + lane);   // byte offset; slot u adds 128u
    uint32_t q[B];
#pragma unroll
    for (int u = 0; u < B; ++u)
        q[u] = __builtin_amdgcn_raw_buffer_load_b16(rs, vo + 2 * kWave * u, 0, 0);
#pragma unroll
    for (int u = 0; u < B; ++u) q[u] = pos[q[u]];
    T x[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
        x[u] = A((int)q[u]);
        // word index kept opaque: (q >> 5) then one shift-add for the address
        // (the compiler's (q >> 3) & ~3 form takes one instruction more)
        uint32_t wi = q[u] >> 5;
        asm volatile("" : "+v"(wi));
        atomicOr(&bm[wi], 1u << (q[u] & 31u));
    }
#pragma unroll
    for (int u = 0; u < B; ++u) nr += (double)x[u];
}

// All of a set's entries through es_mark_hits, 8 slots per lane per round;
// the last round takes the smallest batch that covers the set's end.  The
// sentinel's bit (rank N) is cleared at the end.  sg: the set's ids + 1.
template <class T, class AF>
__device__ __forceinline__ void es_mark_set(const uint16_t *sg, const uint16_t *pos,
                                            AF &A, uint32_t *bm, int k, int N,
                                            int lane, double &nr) {
    // the set's entries as one buffer resource: scalar base, k entries in
    // range, a 32-bit offset per lane and the slot offsets as immediates
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)sg, (short)0, 2 * k, kBufferRsrcWord3);
    for (int j0 = 0; j0 < k; j0 += 8 * kWave) {
        const int ns = (k - j0 + kWave - 1) / kWave;   // slots with set entries
        if (ns > 4) es_mark_hits<T, 8>(rs, pos, A, bm, j0, lane, nr);
        else if (ns > 2) es_mark_hits<T, 4>(rs, pos, A, bm, j0, lane, nr);
        else if (ns > 1) es_mark_hits<T, 2>(rs, pos, A, bm, j0, lane, nr);
        else es_mark_hits<T, 1>(rs, pos, A, bm, j0, lane, nr);
    }
    // LDS operations of one wavefront complete in order: this follows the ORs
    if (lane == 0) atomicAnd(&bm[N >> 5], ~(1u << (N & 31)));
}

// One lane's walk over its hits in rank order.  The lane's W bitmap words
// hold ranks [lane*W*32, (lane+1)*W*32); nz flags its nonempty words and j is
// the set index of its first hit.  Every word read is cleared.  EXACT == false: L accumulates the lane's
// sum of c from 0, vmin / vmax keep the smallest value just before a hit and
// the largest just after one, of L + misses*decay (NaN means none yet:
// fmin / fmax skip it, as the reference's comparison skips a NaN running
// sum).  EXACT == true: L starts at the lane's Sa and every candidate goes
// through es_candidates (first strict maximum, rounded to T).
// min / max of two doubles with NaN operands skipped (IEEE-mode v_min_f64 /
// v_max_f64).  The walk's accumulators only ever hold quiet NaN or results
// of these instructions, so the canonicalisation the compiler adds around
// fmin / fmax (for signalling NaN inputs) is not needed.
__device__ __forceinline__ double min_skip_nan(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double max_skip_nan(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// the same for float (operands are quiet NaN or arithmetic results)
__device__ __forceinline__ float min_skip_nan(float a, float b) {
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float max_skip_nan(float a, float b) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

template <class T, int W, bool EXACT, class AF>
__device__ __forceinline__ void es_walk(uint32_t *bm, AF &A,
                                        const HitShare<T> &share, int lane,
                                        uint32_t nz, int hits, int j, double dec,
                                        double &L, double &vmin, double &vmax,
                                        T &best) {
    uint32_t cur = 0;
    int wq = 0;        // first rank of the current word
    int jt = j;        // set index of the current hit
    for (int t = 0; t < hits; ++t, ++jt) {
        if (cur == 0u) {
            const int w = lane * W + __builtin_ctz(nz);
            nz &= nz - 1u;
            cur = bm[w];
            bm[w] = 0u;   // leaves the bitmap clear for the next set
            wq = w * 32;
        }
        const int q = wq + __builtin_ctz(cur);
        cur &= cur - 1u;
        const double c = share(A(q));
        if (EXACT) {
            es_candidates<T>(L, c, q, jt, dec, best);
        } else {
            // running sum just before the hit, then just after it
            const double before = __builtin_fma((double)(q - jt), dec, L);
            vmin = min_skip_nan(vmin, before);
            L += c;
            vmax = max_skip_nan(vmax, before + c);
        }
    }
}

// Larger sets: one wavefront per set.  Phase A sets the k rank bits in the
// wavefront's N-bit bitmap (lane l owns words l*W .. l*W+W-1, so lanes own
// consecutive rank ranges) and sums N_R.  Each lane then walks its own hits
// once, keeping its sum of c and the extremes of its local running sum; one
// wavefront scan of the sums places every lane's extremes on the set's
// running sum.  Rounding to T is monotone, so the largest (smallest) rounded
// candidate is the rounded largest (smallest) one, and the enrichment score
// is whichever has the larger magnitude.  Only when both magnitudes are
// equal does the order of occurrence decide (the reference keeps the first):
// that case walks the hits again with the first-strict-maximum rule.
template <class T, bool ALDS, int W, class AF>
__device__ __forceinline__ void es_large_set(const RankArgs<T> &a,
                                             const uint16_t *pos, AF &A,
                                             uint32_t *bm, int64_t gp, int s,
                                             int o0, int k, double decay, int lane) {
    const uint16_t *sg = a.large_genes + o0;
    double nr = 0.0;
    es_mark_set<T>(sg, pos, A, bm, k, a.N, lane, nr);
    nr = wave_sum(nr);
    const HitShare<T> share((T)nr);
    // k == N: no misses; the reference's decay (-inf) is never applied
    const double dec = k == a.N ? 0.0 : decay;

    wave_sync_lds();
    int hits = 0;
    uint32_t nz = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const uint32_t w = bm[lane * W + i];
        hits += __popc(w);
        nz |= (w != 0u ? 1u : 0u) << i;
    }
    const int j = wave_excl_scan(hits);
    double L = 0.0, vmin = __builtin_nan(""), vmax = __builtin_nan("");
    T best = T(0);
    es_walk<T, W, false>(bm, A, share, lane, nz, hits, j, dec, L, vmin, vmax, best);
    const double Sa = wave_excl_scan(L);
    const T hi = wave_reduce_idem((T)(Sa + vmax), [](T x, T y) { return max_skip_nan(x, y); });
    const T lo = wave_reduce_idem((T)(Sa + vmin), [](T x, T y) { return min_skip_nan(x, y); });
    const T ah = hi == hi ? fabs(hi) : T(0);
    const T al = lo == lo ? fabs(lo) : T(0);
    if (ah == al && ah != T(0)) {
        // the walk cleared the bitmap: mark the hits again
        wave_sync_lds();
        double unused = 0.0;
        es_mark_set<T>(sg, pos, A, bm, k, a.N, lane, unused);
        wave_sync_lds();
        double S2 = Sa;
        es_walk<T, W, true>(bm, A, share, lane, nz, hits, j, dec, S2, vmin, vmax, best);
        best = wave_first_absmax(best);
    } else {
        best = ah > al ? hi : (al > ah ? lo : T(0));
    }
    if (lane == 0) a.es[(size_t)gp * a.G + s] = best;
    wave_sync_lds();
}

template <class T, bool ALDS, int MAXCH>
__global__ __launch_bounds__(kRankThreads) void k_rank(RankArgs<T> a) {
    typedef typename KeyOf<T>::type KeyT;
    constexpr int NPASS = (int)sizeof(KeyT);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // LDS: [keys | a_rank] (ALDS only) | ids -> pos | counters -> bitmaps | misc
    uint32_t *keys = reinterpret_cast<uint32_t *>(smem_raw);
    uint16_t *ids = reinterpret_cast<uint16_t *>(smem_raw + a.off_ids);
    uint32_t *cnt = reinterpret_cast<uint32_t *>(smem_raw + a.off_cnt);
    int *misc = reinterpret_cast<int *>(smem_raw + a.off_misc);

    const int N = a.N;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int w = tid / kWave;
    const bool asc = a.asc != 0;

    // keys are padded to Npad = 16 wavefronts x CH chunks x 64 lanes with a
    // sentinel above every real key (NaN is canonicalised, so no metric maps
    // to all ones); padding sorts after the N genes and is never read back
    constexpr int CH = MAXCH;           // the host picks MAXCH >= N / 1024
    const int seg = CH * kWave;
    const int Npad = kRankWaves * seg;
    const KeyT kSentinel = ~(KeyT)0;

    for (int p = blockIdx.x; p < a.Pb; p += gridDim.x) {
        const T *rp = a.r + (size_t)p * N;
        for (int g = tid; g < Npad; g += kRankThreads) {
            if (ALDS) keys[g] = g < N ? (uint32_t)sort_key(rp[g], asc) : 0xffffffffu;
            ids[g] = (uint16_t)g;
        }
        if (tid == 0) misc[0] = 0;
        __syncthreads();

        auto key_of = [&](int g) -> KeyT {
            if (ALDS) return (KeyT)keys[g];
            return g < N ? (KeyT)sort_key(rp[g], asc) : kSentinel;
        };

        const int start = w * seg;
        uint32_t *wcnt = cnt + w * 256;

        for (int pass = 0; pass < NPASS; ++pass) {
            const int shift = pass * 8;
            for (int i = lane; i < 256; i += kWave) wcnt[i] = 0;
            // per chunk: gene id (bits 0-15) | digit (bits 16-23); local rank
            uint32_t pk[MAXCH], lr[MAXCH];
#pragma unroll
            for (int c = 0; c < MAXCH; ++c) {
                if (c < CH) {
                    const uint32_t id = ids[start + c * kWave + lane];
                    const uint32_t d = (uint32_t)((key_of((int)id) >> shift) & 255u);
                    pk[c] = id | (d << 16);
                    // keep one packed register per chunk (no hoisted parts)
                    asm volatile("" : "+v"(pk[c]));
                }
            }
            wave_sync_lds();
#pragma unroll
            for (int c = 0; c < MAXCH; ++c) {
                if (c < CH) {
                    const uint32_t d = pk[c] >> 16;
                    uint64_t peers = ~0ull;
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        const bool bit = (d >> b) & 1u;
                        const uint64_t bal = ballot64(bit);
                        peers &= bit ? bal : ~bal;
                    }
                    const int before = mbcnt64(peers);
                    const int tot = __popcll(peers);
                    const uint32_t base = wcnt[d];
                    lr[c] = base + (uint32_t)before;
                    asm volatile("" : "+v"(lr[c]));
                    wave_sync_lds();
                    if (before == 0) wcnt[d] = base + (uint32_t)tot;
                    wave_sync_lds();
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            __syncthreads();
            // exclusive scan over (digit, wavefront), digit-major
            uint32_t tot_d = 0;
            if (tid < 256) {
                uint32_t run = 0;
                for (int w2 = 0; w2 < kRankWaves; ++w2) {
                    const uint32_t v = cnt[w2 * 256 + tid];
                    cnt[w2 * 256 + tid] = run;
                    run += v;
                }
                tot_d = run;
                uint32_t inc = tot_d;
#pragma unroll
                for (int off = 1; off < kWave; off <<= 1) {
                    const uint32_t o = __shfl_up(inc, off, kWave);
                    if (lane >= off) inc += o;
                }
                if (lane == kWave - 1) misc[4 + w] = (int)inc;
                tot_d = inc - tot_d;   // exclusive within the wavefront
            }
            __syncthreads();
            if (tid < 256) {
                uint32_t base = tot_d;
                for (int w2 = 0; w2 < w; ++w2) base += (uint32_t)misc[4 + w2];
                for (int w2 = 0; w2 < kRankWaves; ++w2) cnt[w2 * 256 + tid] += base;
            }
            __syncthreads();
#pragma unroll
            for (int c = 0; c < MAXCH; ++c) {
                if (c < CH)
                    ids[wcnt[pk[c] >> 16] + lr[c]] = (uint16_t)(pk[c] & 0xffffu);
            }
            __syncthreads();
        }

        if (a.idx_out) {
            for (int i = tid; i < N; i += kRankThreads)
                a.idx_out[(size_t)p * N + i] = ids[i];
        }
        if (a.G == 0) {
            __syncthreads();
            continue;
        }

        // |metric| by rank, then invert ids into pos[gene] = rank; both rows go
        // to HBM for k_es
        if (a.pos_out) {
            T *arow = a.a_glob + (size_t)p * N;
            uint32_t gi[MAXCH];
#pragma unroll
            for (int j = 0; j < MAXCH; ++j) {
                const int i = tid + j * kRankThreads;
                const uint32_t g = ids[i];
                gi[j] = g;
                if (i < N) {
                    if (ALDS) arow[i] = (T)key_abs(keys[g], asc);
                    else arow[i] = (T)__builtin_fabs((double)rp[g]);
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < MAXCH; ++j)
                ids[gi[j]] = (uint16_t)(tid + j * kRankThreads);
            __syncthreads();
            for (int g = tid; g < N; g += kRankThreads)
                a.pos_out[(size_t)p * N + g] = ids[g];
        }
        __syncthreads();
    }
}

// ===========================================================================
// k_es: enrichment scores of every gene set for one permutation per
// workgroup, from the rank rows k_rank wrote (pos[gene], |metric| by rank).
// Both rows are staged in LDS (|metric| only in fp32; fp64 reads it from L2).
// Wavefronts take work items from an LDS counter: large sets (one wavefront
// each, bitmap scan) first, then chunks of 64 small sets (one lane each).
// ===========================================================================
template <class T, bool ALDS, int W>
__global__ __launch_bounds__(kRankThreads) void k_es(RankArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *a_lds = reinterpret_cast<float *>(smem_raw);
    uint16_t *pos = reinterpret_cast<uint16_t *>(smem_raw + a.off_ids);
    uint32_t *bms = reinterpret_cast<uint32_t *>(smem_raw + a.off_cnt);
    int *misc = reinterpret_cast<int *>(smem_raw + a.off_misc);
    const int N = a.N;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int w = tid / kWave;

    for (int p = blockIdx.x; p < a.Pb; p += gridDim.x) {
        const T *arow = a.a_glob + (size_t)(a.p_lo + p) * N;
        const uint16_t *prow = a.pos_out + (size_t)(a.p_lo + p) * N;
        // pos is indexed by gene id + 1: entry 0 is the sentinel gene, rank N,
        // whose |metric| (a_lds[N], or 0 below) is 0
        for (int g = tid; g < N; g += kRankThreads) {
            pos[g + 1] = prow[g];
            if (ALDS) a_lds[g] = (float)arow[g];
        }
        if (tid == 0) {
            misc[0] = 0;
            pos[0] = (uint16_t)N;
            if (ALDS) a_lds[N] = 0.0f;
        }
        if (w < a.es_waves)
            for (int i = lane; i < a.nw; i += kWave) bms[(size_t)w * a.nw + i] = 0u;
        __syncthreads();
        auto A = [&](int q) -> T {
            if (ALDS) return (T)a_lds[q];
            return q < N ? arow[q] : T(0);
        };
        if (w < a.es_waves) {
            uint32_t *bm = bms + (size_t)w * a.nw;
            const int64_t gp = a.perm_base + a.p_lo + p;
            // work-item indices and descriptors are wavefront-uniform: they
            // live in scalar registers (scalar loads, scalar address math)
            int item = 0;
            if (lane == 0) item = atomicAdd(&misc[0], 1);
            item = __builtin_amdgcn_readfirstlane(item);
            while (item < a.n_items) {
                // fetch the next work item now; its latency hides behind this one
                int next = 0;
                if (lane == 0) next = atomicAdd(&misc[0], 1);
                if (item < a.n_large) {
                    // read-only tables through the constant address space:
                    // scalar loads of the uniform descriptor and decay
                    const __attribute__((address_space(4))) int *d =
                        (const __attribute__((address_space(4))) int *)(a.large_desc + item);
                    const T dec = *((const __attribute__((address_space(4))) T *)
                                        a.large_decay + item);
                    es_large_set<T, ALDS, W>(a, pos, A, bm, gp, d[0], d[1], d[2], (double)dec,
                                             lane);
                } else {
                    const int chunk = item - a.n_large;
                    if (chunk < a.chunk_end8)
                        es_small_sets<T, 8>(a, pos, A, gp, chunk, lane);
                    else if (chunk < a.chunk_end16)
                        es_small_sets<T, 16>(a, pos, A, gp, chunk, lane);
                    else if (chunk < a.chunk_end32)
                        es_small_sets<T, 32>(a, pos, A, gp, chunk, lane);
                    else if (chunk < a.chunk_end64)
                        es_small_sets<T, 64>(a, pos, A, gp, chunk, lane);
                    else
                        es_small_sets<T, 128>(a, pos, A, gp, chunk, lane);
                }
                item = __builtin_amdgcn_readfirstlane(next);
            }
        }
        __syncthreads();
    }
}

// ===========================================================================
// statistics, statistics.cuh:46-117 (+ FDR defined by this project)
// es is [P][G] (permutation-major).
// ===========================================================================

// per set: ES, NES (naive sequential sum of |ES| in permutation order,
// statistics.cuh:60-70), NP (statistics.cuh:73-77).  budget (optional):
// set s was scored for permutations 0..budget[s]-1 only and its statistics
// use those; cv (optional, with budget) receives std/mean of |null ES|.
template <class T>
__global__ void k_set_stats(const T *__restrict__ es, int64_t P, int G,
                            T *__restrict__ out, T *__restrict__ factor,
                            const int32_t *__restrict__ budget,
                            float *__restrict__ cv) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= G) return;
    const int64_t B = budget ? (int64_t)budget[s] : P;
    const T e0 = es[s];
    const T ref = __builtin_fabs(e0);
    T sum = T(0);
    double s1 = 0.0, s2 = 0.0;
    int64_t c = 0;
    int64_t p = 0;
    // 32 loads in flight per lane; the sum stays sequential in p
    constexpr int U = 32;
    for (; p + U <= B; p += U) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = __builtin_fabs(es[(p + u) * G + s]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            sum += v[u];
            c += v[u] >= ref;
            if (cv) { s1 += (double)v[u]; s2 += (double)v[u] * v[u]; }
        }
    }
    for (; p < B; ++p) {
        const T v = __builtin_fabs(es[p * G + s]);
        sum += v;
        c += v >= ref;
        if (cv) { s1 += (double)v; s2 += (double)v * v; }
    }
    const T f = (T)B / sum;
    out[s] = e0;
    out[G + s] = e0 * f;
    out[2 * G + s] = (T)c / (T)B;
    factor[s] = f;
    if (cv) {
        const double m = s1 / (double)B;
        const double var = s2 / (double)B - m * m;
        cv[s] = (float)(m > 0.0 && var > 0.0 ? sqrt(var) / m : 0.0);
    }
}

// per permutation: extrema of normalised ES over sets (statistics.cuh:85-97),
// over the sets scored for that permutation when budgets are given
template <class T>
__global__ void k_perm_extrema(const T *__restrict__ es, int64_t P, int G,
                               const T *__restrict__ factor,
                               const int32_t *__restrict__ budget,
                               T *__restrict__ maxv, T *__restrict__ minv) {
    const int64_t p = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (p >= P) return;
    const int lane = lane_id();
    T mx = -__builtin_inf(), mn = __builtin_inf();
    for (int s = lane; s < G; s += kWave) {
        if (budget && p >= budget[s]) continue;
        const T v = es[p * G + s] * factor[s];
        mx = mx < v ? v : mx;
        mn = mn > v ? v : mn;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const T omx = __shfl_xor(mx, off, kWave);
        const T omn = __shfl_xor(mn, off, kWave);
        mx = mx < omx ? omx : mx;
        mn = mn > omn ? omn : mn;
    }
    if (lane == 0) { maxv[p] = mx; minv[p] = mn; }
}

// FWER, statistics.cuh:100-110 (over the set's own permutations when
// budgets are given)
template <class T>
__global__ void k_fwer(T *__restrict__ out, int64_t P, int G,
                       const T *__restrict__ maxv, const T *__restrict__ minv,
                       const int32_t *__restrict__ budget) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= G) return;
    const int64_t B = budget ? (int64_t)budget[s] : P;
    const T score = out[G + s];
    int64_t c = 0;
    if (score > T(0)) {
        for (int64_t p = 0; p < B; ++p) c += maxv[p] >= score;
    } else {
        for (int64_t p = 0; p < B; ++p) c += minv[p] <= score;
    }
    out[3 * G + s] = (T)c / (T)B;
}

// FWER from the per-permutation extrema sorted ascending (same counts as
// k_fwer's loop: #max >= NES for NES > 0, #min <= NES otherwise; NaN NES
// counts nothing)
template <class T>
__device__ __forceinline__ int64_t count_ge(const T *v, int64_t n, T x) {
    int64_t lo = 0, hi = n;   // first index with v >= x
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (v[mid] < x) lo = mid + 1; else hi = mid;
    }
    return n - lo;
}

template <class T>
__device__ __forceinline__ int64_t count_le(const T *v, int64_t n, T x) {
    int64_t lo = 0, hi = n;   // first index with v > x
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (v[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

template <class T>
__global__ void k_fwer_sorted(T *__restrict__ out, int64_t P, int G,
                              const T *__restrict__ max_sorted,
                              const T *__restrict__ min_sorted) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= G) return;
    const T score = out[G + s];
    int64_t c = 0;
    if (score > T(0)) c = count_ge(max_sorted, P, score);
    else if (score == score) c = count_le(min_sorted, P, score);
    out[3 * G + s] = (T)c / (T)P;
}

template <class T>
__device__ __forceinline__ int lower_bound_t(const T *v, int n, T x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

template <class T>
__device__ __forceinline__ int upper_bound_t(const T *v, int n, T x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// FDR histograms.  thr[0..n) are the sorted non-NaN NES.  For every null
// score z (permutations 1..P-1, normalised as statistics.cuh:90): hub[#thr<=z]
// and hlb[#thr<z] are counted, plus #z>=0 and #z<=0.
// LDS=true keeps thresholds and per-workgroup histograms in LDS; LDS=false
// searches and counts in global memory (large numbers of sets).
template <class T, bool LDS>
__global__ void k_fdr_hist(const T *__restrict__ es, int64_t P, int G,
                           const T *__restrict__ factor,
                           const int32_t *__restrict__ budget,
                           const T *__restrict__ thr, int n,
                           const int32_t *__restrict__ table, int nb, T lo, T inv_w,
                           unsigned long long *__restrict__ hub,
                           unsigned long long *__restrict__ hlb,
                           unsigned long long *__restrict__ signs) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const T *t = thr;
    uint32_t *h1 = nullptr, *h2 = nullptr;
    int32_t *tb = nullptr;
    if (LDS) {
        T *tl = reinterpret_cast<T *>(smem_raw);
        tb = reinterpret_cast<int32_t *>(smem_raw + ((sizeof(T) * n + 15) / 16) * 16);
        h1 = reinterpret_cast<uint32_t *>(reinterpret_cast<unsigned char *>(tb) +
                                          ((sizeof(int32_t) * (nb + 1) + 15) / 16) * 16);
        h2 = h1 + (n + 1);
        for (int i = threadIdx.x; i < n; i += blockDim.x) tl[i] = thr[i];
        for (int i = threadIdx.x; i <= nb; i += blockDim.x) tb[i] = table[i];
        for (int i = threadIdx.x; i <= n; i += blockDim.x) { h1[i] = 0; h2[i] = 0; }
        __syncthreads();
        t = tl;
    }
    unsigned long long npos = 0, nneg = 0;
    // whole permutation rows per block (no index division per score)
    for (int64_t p = 1 + blockIdx.x; p < P; p += gridDim.x) {
        for (int s = threadIdx.x; s < G; s += blockDim.x) {
            // a set scored for B of the P permutations stands for P/B null
            // samples per score (budgets divide P)
            uint32_t w = 1;
            if (budget) {
                if (p >= budget[s]) continue;
                w = (uint32_t)(P / budget[s]);
            }
            const T z = es[p * G + s] * factor[s];
            if (z != z) continue;
            int lb;
            if (LDS && nb > 0) {
                // uniform buckets over [t[0], t[n-1]]: the bucket's table
                // entries (one bucket of slack each side for the rounding of
                // the bucket index) bound the binary search
                if (!(z > t[0])) {
                    lb = 0;
                } else if (z > t[n - 1]) {
                    lb = n;
                } else {
                    int b = (int)((z - lo) * inv_w);
                    b = b < 0 ? 0 : (b >= nb ? nb - 1 : b);
                    int a = tb[b > 0 ? b - 1 : 0];
                    int c = tb[b + 2 <= nb ? b + 2 : nb];
                    while (a < c) {
                        const int mid = (a + c) >> 1;
                        if (t[mid] < z) a = mid + 1; else c = mid;
                    }
                    lb = a;
                }
            } else {
                lb = lower_bound_t(t, n, z);
            }
            int ub = lb;                       // thresholds equal to z are rare
            while (ub < n && !(z < t[ub])) ++ub;
            if (LDS) {
                atomicAdd(&h1[ub], w);
                atomicAdd(&h2[lb], w);
            } else {
                atomicAdd(&hub[ub], (unsigned long long)w);
                atomicAdd(&hlb[lb], (unsigned long long)w);
            }
            npos += z >= T(0) ? w : 0u;
            nneg += z <= T(0) ? w : 0u;
        }
    }
    npos = wave_sum(npos);
    nneg = wave_sum(nneg);
    if (lane_id() == 0) {
        atomicAdd(&signs[0], npos);
        atomicAdd(&signs[1], nneg);
    }
    if (LDS) {
        __syncthreads();
        for (int i = threadIdx.x; i <= n; i += blockDim.x) {
            if (h1[i]) atomicAdd(&hub[i], (unsigned long long)h1[i]);
            if (h2[i]) atomicAdd(&hlb[i], (unsigned long long)h2[i]);
        }
    }
}

// inclusive prefix sums of both histograms (single workgroup, in place)
__global__ void k_prefix_u64(unsigned long long *__restrict__ h1,
                             unsigned long long *__restrict__ h2, int len) {
    __shared__ unsigned long long carry[2];
    if (threadIdx.x == 0) { carry[0] = 0; carry[1] = 0; }
    __syncthreads();
    for (int base = 0; base < len; base += kWave) {
        if (threadIdx.x < kWave) {
            const int i = base + (int)threadIdx.x;
            unsigned long long a = i < len ? h1[i] : 0ull;
            unsigned long long b = i < len ? h2[i] : 0ull;
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                const unsigned long long oa = __shfl_up(a, off, kWave);
                const unsigned long long ob = __shfl_up(b, off, kWave);
                if ((int)threadIdx.x >= off) { a += oa; b += ob; }
            }
            a += carry[0];
            b += carry[1];
            if (i < len) { h1[i] = a; h2[i] = b; }
            const unsigned long long la = __shfl(a, kWave - 1, kWave);
            const unsigned long long lb = __shfl(b, kWave - 1, kWave);
            if (threadIdx.x == 0) { carry[0] = la; carry[1] = lb; }
        }
        __syncthreads();
    }
}

// FDR per set from the prefix-summed histograms:
// #z >= t = total - cum_hub[lb(t)], #z <= t = cum_hlb[lb(t)]
template <class T>
__global__ void k_fdr_final(T *__restrict__ out, int G,
                            const T *__restrict__ thr, int n,
                            const unsigned long long *__restrict__ cub,
                            const unsigned long long *__restrict__ clb,
                            const unsigned long long *__restrict__ signs) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= G) return;
    const T nes = out[G + s];
    if (nes != nes) { out[4 * G + s] = (T)__builtin_nan(""); return; }
    const unsigned long long total = cub[n];
    const int lbt = lower_bound_t(thr, n, nes);
    unsigned long long a, b, dn, dobs;
    if (nes > T(0)) {
        a = total - cub[lbt];
        b = (unsigned long long)(n - lbt);
        dn = signs[0];
        dobs = (unsigned long long)(n - lower_bound_t(thr, n, T(0)));
    } else {
        a = clb[lbt];
        b = (unsigned long long)upper_bound_t(thr, n, nes);
        dn = signs[1];
        dobs = (unsigned long long)upper_bound_t(thr, n, T(0));
    }
    const double num = dn ? (double)a / (double)dn : 0.0;
    const double den = (double)b / (double)dobs;
    const double q = num / den;
    out[4 * G + s] = (T)(q < 1.0 ? q : 1.0);
}

// [P][G] -> [G][P]
template <class T>
__global__ void k_transpose(const T *__restrict__ in, int64_t P, int G,
                            T *__restrict__ out) {
    __shared__ T tile[32][33];
    const int64_t p0 = (int64_t)blockIdx.y * 32;
    const int s0 = blockIdx.x * 32;
    for (int i = threadIdx.y; i < 32; i += blockDim.y) {
        const int64_t p = p0 + i;
        const int s = s0 + threadIdx.x;
        if (p < P && s < G) tile[i][threadIdx.x] = in[p * G + s];
    }
    __syncthreads();
    for (int i = threadIdx.y; i < 32; i += blockDim.y) {
        const int s = s0 + i;
        const int64_t p = p0 + threadIdx.x;
        if (p < P && s < G) out[(int64_t)s * P + p] = tile[threadIdx.x][i];
    }
}

}  // namespace gsea
