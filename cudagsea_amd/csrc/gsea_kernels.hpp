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
#include "gsea_metric.hpp"
#include "gsea_rank.hpp"

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
    // k_es gathers pos[id + 1] for 32 consecutive entries per half-wavefront
    // access; pos is 16-bit, so entry id + 1 = g1 sits in LDS bank
    // (g1 >> 1) mod 32 up to a rotation that is the same for every entry.
    // The entries are stored sorted by (r, bank), r being an entry's index
    // among its set's entries of the same bank in set order: the r-th round
    // holds one entry of each bank with more than r entries, so an aligned
    // window of 32 entries spans at most the end of one round and the start
    // of the next, and its banks are distinct but where the two overlap.
    // The order of a set's entries changes nothing but the order of its N_R
    // summands.  An entry's position is the number of (r, bank) keys below
    // its own: sum over banks c of min(cnt_c, r) + #{c < bank : cnt_c > r}.
    // 256 threads: thread (b, s) = (tid % 32, tid / 32) handles the ids of
    // bank b in segment s, the s-th eighth of the set; a prefix over the
    // segments gives each thread the round index of its first id
    constexpr int kSeg = 8;
    __shared__ int cs[kSeg][32];
    __shared__ int cnt[32];
    const int4 d = desc[blockIdx.x];
    const int32_t *src = genes + off[d.x];
    const int b = threadIdx.x & 31, s = threadIdx.x >> 5;
    const int len = (d.z + kSeg - 1) / kSeg;
    const int t0 = min(s * len, d.z), t1 = min(t0 + len, d.z);
    int n = 0;
    for (int t = t0; t < t1; ++t) n += (((src[t] + 1) >> 1) & 31) == b ? 1 : 0;
    cs[s][b] = n;
    __syncthreads();
    int r = 0;
    for (int s2 = 0; s2 < s; ++s2) r += cs[s2][b];
    if (s == kSeg - 1) cnt[b] = r + n;
    __syncthreads();
    for (int t = t0; t < t1; ++t) {
        const int g1 = src[t] + 1;
        if (((g1 >> 1) & 31) != b) continue;
        int p = 0;
        for (int c = 0; c < 32; ++c) {
            const int m = cnt[c];
            p += min(m, r) + (c < b && m > r ? 1 : 0);
        }
        out[d.y + p] = (uint16_t)g1;
        ++r;
    }
}

// gene ids of the small sets, laid out [t][slot] so a wavefront's 64 lanes
// read row t of 64 consecutive slots; rows t >= k hold the sentinel.  Slots
// from pair0 on hold pair sets: an odd slot has the set's entries from
// kSmallSet on.
__global__ void k_small_gm(const int32_t *__restrict__ off,
                           const int32_t *__restrict__ genes,
                           const int32_t *__restrict__ small_ids, int nslots, int pair0,
                           uint16_t *__restrict__ gm) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nslots) return;
    const int s = small_ids[slot];
    const int skip = slot >= pair0 && (slot & 1) ? kSmallSet : 0;
    const int o0 = s >= 0 ? off[s] + skip : 0;
    const int k = s >= 0 ? off[s + 1] - o0 : 0;
    for (int t = 0; t < kSmallSet; ++t)
        gm[(size_t)t * nslots + slot] = (uint16_t)(t < k ? genes[o0 + t] + 1 : 0);
}

// ===========================================================================
// statistics, statistics.cuh:46-117 (+ FDR defined by this project)
// es is [P][G] (permutation-major).
// ===========================================================================

// per set: ES, NES (naive sequential sum of |ES| in permutation order,
// statistics.cuh:60-70), NP (statistics.cuh:73-77).  budget (optional):
// set s was scored for permutations 0..budget[s]-1 only and its statistics
// use those; cv (optional, with budget) receives std/mean of |null ES|.
// order-sensitive 64-bit checksum of a set CSR (offsets[G+1], then the
// offsets[G] gene ids), one workgroup: identifies a set list whose derived
// work tables can be reused
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

__global__ __launch_bounds__(1024) void k_csr_checksum(const int32_t *__restrict__ off,
                                                       const int32_t *__restrict__ genes,
                                                       int64_t G,
                                                       unsigned long long *__restrict__ out) {
    __shared__ unsigned long long part[1024 / kWave];
    const int64_t total = off[G];
    unsigned long long h = 0;
    for (int64_t i = threadIdx.x; i <= G; i += blockDim.x)
        h += mix64(((unsigned long long)(uint32_t)off[i] << 32) ^ (unsigned long long)i);
    for (int64_t i = threadIdx.x; i < total; i += blockDim.x)
        h += mix64(((unsigned long long)(uint32_t)genes[i] << 32) ^
                   (0x8000000000000000ull | (unsigned long long)i));
    h = wave_sum(h);
    if (lane_id() == 0) part[threadIdx.x / kWave] = h;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += part[w];
        out[0] = t;
        out[1] = (unsigned long long)total;
    }
}

// Per set: ES, NES, NP (statistics.cuh:56-78) and the normalisation
// factors, factor[s] for scores >= 0 and factor[G + s] for scores < 0.
// SIGNED=false: the reference's one factor B / sum_p |ES_p| over permutations
// 0..B-1 (both entries equal).  SIGNED=true (GSEA_FLAG_SIGNED_NES): the
// Broad GSEA normalisation, each sign divided by the mean magnitude of the
// null scores (permutations 1..B-1, NaN skipped) of that sign.  NP is the
// reference's two-sided count either way.
template <class T, bool SIGNED>
__global__ void k_set_stats(const T *__restrict__ es, int64_t P, int G,
                            T *__restrict__ out, T *__restrict__ factor,
                            const int32_t *__restrict__ budget,
                            float *__restrict__ cv) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= G) return;
    const int64_t B = budget ? (int64_t)budget[s] : P;
    const T e0 = es[s];
    const T ref = __builtin_fabs(e0);
    T sum = T(0), sneg = T(0);
    int64_t npos = 0, nneg = 0;
    double s1 = 0.0, s2 = 0.0;
    int64_t c = 0;
    int64_t p = 0;
    // 32 loads in flight per lane; the sums stay sequential in p
    constexpr int U = 32;
    for (; p + U <= B; p += U) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = es[(p + u) * G + s];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const T a = __builtin_fabs(v[u]);
            if (!SIGNED) {
                sum += a;
            } else if (p + u > 0) {
                if (v[u] < T(0)) { sneg += a; ++nneg; }
                else if (v[u] >= T(0)) { sum += a; ++npos; }
            }
            c += a >= ref;
            if (cv) { s1 += (double)a; s2 += (double)a * a; }
        }
    }
    for (; p < B; ++p) {
        const T raw = es[p * G + s];
        const T a = __builtin_fabs(raw);
        if (!SIGNED) {
            sum += a;
        } else if (p > 0) {
            if (raw < T(0)) { sneg += a; ++nneg; }
            else if (raw >= T(0)) { sum += a; ++npos; }
        }
        c += a >= ref;
        if (cv) { s1 += (double)a; s2 += (double)a * a; }
    }
    const T fp = SIGNED ? (T)npos / sum : (T)B / sum;
    const T fn = SIGNED ? (T)nneg / sneg : fp;
    out[s] = e0;
    out[G + s] = e0 * (e0 < T(0) ? fn : fp);
    out[2 * G + s] = (T)c / (T)B;
    factor[s] = fp;
    factor[G + s] = fn;
    if (cv) {
        const double m = s1 / (double)B;
        const double var = s2 / (double)B - m * m;
        cv[s] = (float)(m > 0.0 && var > 0.0 ? sqrt(var) / m : 0.0);
    }
}

// a score normalised by its set's factor of its sign (statistics.cuh:90)
template <class T>
__device__ __forceinline__ T normalised(T v, const T *__restrict__ factor, int G, int s) {
    return v * factor[v < T(0) ? G + s : s];
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
        const T v = normalised(es[p * G + s], factor, G, s);
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

// Sorted observed NES and the bucket table that bounds searches in them,
// produced on the device by k_sort_thresholds (or uploaded by the host when
// there are more sets than one workgroup sorts)
template <class T>
struct FdrInfo {
    int n;      // non-NaN NES, sorted ascending in thr[0..n)
    int nb;     // buckets of the table (0: no table)
    T lo;       // thr[0]
    T inv_w;    // buckets per unit of NES
};

constexpr int kFdrBuckets = 4096;

// zero fill (stream-ordered, graph-safe: no memset nodes)
__global__ void k_fill_u64(unsigned long long *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

template <class K>
__device__ __forceinline__ K key_max();
template <>
__device__ __forceinline__ uint32_t key_max<uint32_t>() { return 0xffffffffu; }
template <>
__device__ __forceinline__ uint64_t key_max<uint64_t>() { return ~0ull; }

// value of an ascending sort key (sort_key(v, true)); NaN keys are not
// converted back
__device__ __forceinline__ float key_value(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
__device__ __forceinline__ double key_value(uint64_t u) {
    return __longlong_as_double((long long)((u & 0x8000000000000000ull)
                                                ? (u & 0x7fffffffffffffffull) : ~u));
}

// One workgroup: the NES row of stats [5][G] sorted ascending (bitonic
// network over npow2 >= G keys in LDS, NaN and padding last), the non-NaN
// values written to thr[0..n), and the bucket table of the FDR search:
// nb = min(4096, n) uniform buckets over [thr[0], thr[n-1]] (width w in
// double), table[b] = #thresholds below thr[0] + b w, table[0] = 0,
// table[nb] = n; no table when n < 2 or all thresholds are equal.
template <class T>
__global__ __launch_bounds__(1024) void k_sort_thresholds(const T *__restrict__ stats, int G,
                                                          int npow2, T *__restrict__ thr,
                                                          FdrInfo<T> *__restrict__ info,
                                                          int32_t *__restrict__ table) {
    typedef typename KeyOf<T>::type K;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    K *key = reinterpret_cast<K *>(smem_raw);
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int mine = 0;
    for (int i = threadIdx.x; i < npow2; i += blockDim.x) {
        K k = key_max<K>();
        if (i < G) {
            const T v = stats[G + i];
            k = (K)sort_key(v, true);   // NaN canonicalised above +inf
            mine += v == v;
        }
        key[i] = k;
    }
    atomicAdd(&cnt, mine);
    for (int size = 2; size <= npow2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int t = threadIdx.x; t < (npow2 >> 1); t += blockDim.x) {
                const int i = 2 * stride * (t / stride) + (t % stride);
                const int j = i + stride;
                const K x = key[i], y = key[j];
                const bool up = (i & size) == 0;
                if ((x > y) == up) { key[i] = y; key[j] = x; }
            }
        }
    }
    __syncthreads();
    const int n = cnt;
    for (int i = threadIdx.x; i < n; i += blockDim.x) thr[i] = (T)key_value(key[i]);
    int nb = 0;
    double w = 0.0;
    if (n > 1) {
        const T t0 = (T)key_value(key[0]), t1 = (T)key_value(key[n - 1]);
        if (t1 > t0) {
            nb = n < kFdrBuckets ? n : kFdrBuckets;
            w = ((double)t1 - (double)t0) / nb;
        }
    }
    if (nb) {
        const double t0 = (double)(T)key_value(key[0]);
        for (int b = threadIdx.x; b <= nb; b += blockDim.x) {
            int v;
            if (b == 0) {
                v = 0;
            } else if (b == nb) {
                v = n;
            } else {
                const double edge = t0 + b * w;
                int lo = 0, hi = n;   // first threshold with (double)x >= edge
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if ((double)(T)key_value(key[mid]) < edge) lo = mid + 1; else hi = mid;
                }
                v = lo;
            }
            table[b] = v;
        }
    }
    if (threadIdx.x == 0) {
        FdrInfo<T> f;
        f.n = n;
        f.nb = nb;
        f.lo = n ? (T)key_value(key[0]) : T(0);
        f.inv_w = nb ? (T)(1.0 / w) : T(0);
        *info = f;
    }
}

// FWER counts through the sorted NES: for every permutation p (0 included),
// hmax[#thresholds <= max_p] and hmin[#thresholds < min_p] are counted.  A set
// with NES > 0 then has #{p : max_p >= NES} = P - (prefix of hmax at
// lower_bound(NES)), one with NES <= 0 #{p : min_p <= NES} = prefix of hmin
// at lower_bound(NES) -- the counts of statistics.cuh:100-110.
template <class T>
__global__ void k_fwer_hist(const T *__restrict__ maxv, const T *__restrict__ minv, int64_t P,
                            const T *__restrict__ thr, const FdrInfo<T> *__restrict__ info,
                            unsigned long long *__restrict__ hmax,
                            unsigned long long *__restrict__ hmin) {
    const int n = info->n;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P;
         p += (int64_t)gridDim.x * blockDim.x) {
        atomicAdd(&hmax[upper_bound_t(thr, n, maxv[p])], 1ull);
        atomicAdd(&hmin[lower_bound_t(thr, n, minv[p])], 1ull);
    }
}

template <class T>
__global__ void k_fwer_final(T *__restrict__ out, int64_t P, int G, const T *__restrict__ thr,
                             const FdrInfo<T> *__restrict__ info,
                             const unsigned long long *__restrict__ cmax,
                             const unsigned long long *__restrict__ cmin) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= G) return;
    const int n = info->n;
    const T score = out[G + s];
    int64_t c = 0;
    if (score > T(0)) c = P - (int64_t)cmax[lower_bound_t(thr, n, score)];
    else if (score == score) c = (int64_t)cmin[lower_bound_t(thr, n, score)];
    out[3 * G + s] = (T)c / (T)P;
}

// FDR histograms.  thr[0..n) are the sorted non-NaN NES.  For every null
// score z (permutations 1..P-1, normalised as statistics.cuh:90): hub[#thr<=z]
// and hlb[#thr<z] are counted, plus #z>=0 and #z<=0.
// LDS=true keeps thresholds and per-workgroup histograms in LDS; LDS=false
// searches and counts in global memory (large numbers of sets).
template <class T, bool LDS>
__global__ __launch_bounds__(1024) void k_fdr_hist(const T *__restrict__ es, int64_t P, int G,
                           const T *__restrict__ factor,
                           const int32_t *__restrict__ budget,
                           const T *__restrict__ thr, const FdrInfo<T> *__restrict__ info,
                           const int32_t *__restrict__ table,
                           unsigned long long *__restrict__ hub,
                           unsigned long long *__restrict__ hlb,
                           unsigned long long *__restrict__ signs) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n = info->n, nb = LDS ? info->nb : 0;
    const T lo = info->lo, inv_w = info->inv_w;
    const T *t = thr;
    uint32_t *h1 = nullptr, *h2 = nullptr;
    int32_t *tb = nullptr;
    if (LDS) {
        // LDS layout sized by the host for G thresholds and the full table
        T *tl = reinterpret_cast<T *>(smem_raw);
        tb = reinterpret_cast<int32_t *>(smem_raw + ((sizeof(T) * G + 15) / 16) * 16);
        h1 = reinterpret_cast<uint32_t *>(reinterpret_cast<unsigned char *>(tb) +
                                          ((sizeof(int32_t) * (kFdrBuckets + 1) + 15) / 16) * 16);
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
            const T z = normalised(es[p * G + s], factor, G, s);
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

// inclusive prefix sums of both histograms (single workgroup, in place) over
// *nptr + 1 entries
__global__ void k_prefix_u64(unsigned long long *__restrict__ h1,
                             unsigned long long *__restrict__ h2, const int *__restrict__ nptr) {
    const int len = *nptr + 1;
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
                            const T *__restrict__ thr, const FdrInfo<T> *__restrict__ info,
                            const unsigned long long *__restrict__ cub,
                            const unsigned long long *__restrict__ clb,
                            const unsigned long long *__restrict__ signs) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= G) return;
    const int n = info->n;
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