// gsea_metric.hpp -- label permutations' local ranking metrics (k_metric):
// the 14 metrics of the reference, many permutations per LDS tile of X.
#pragma once

#include "gsea_device.hpp"

namespace gsea {

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

}  // namespace gsea
