// gsea_rank.hpp -- per-permutation argsort (k_rank) and enrichment scores
// of every gene set (k_es).
#pragma once

#include "gsea_device.hpp"

namespace gsea {

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
    int W;                   // k_es: bitmap words per lane, (N + 1) bits in all
    size_t off_ids, off_cnt, off_misc;
    // work items: the chunks (64 lanes each: sets of one size class, one lane
    // per set, or 32 pair sets), then the n_large sets above the pair class
    // (one wavefront each)
    int n_items, n_large;
    const int4 *large_desc;            // [n_large] {set id, offset, size, 0},
                                       // in set order
    const T *large_decay;              // [n_large] decay of those sets
    const int32_t *small_ids;          // [n_chunks*64] set id or -1
    const int32_t *small_k;            // [n_chunks*64] set sizes
    const uint16_t *small_gm;          // [kSmallSet][small_stride] gene ids + 1
    const int32_t *small_kmax;         // [n_chunks] largest set of each chunk
    int small_stride;
    int chunk_end8, chunk_end16, chunk_end32, chunk_end64;   // chunk class boundaries
    int chunk_end128;                  // chunks from here on: pair sets (129..256 genes)
    const T *decay;                    // [G] (T)(-1.0 / (double)(N - k_s))
};

constexpr int kSmallSet = 128;

// k_es reads gene ids shifted by one, as 16-bit values: id 0 is the
// sentinel gene, whose rank is N (past every real rank) and whose |metric| is
// 0.  A load past a set's entries reads 0 (out-of-range buffer loads return
// zero; the gene-id rows of small sets are zero-filled), so it marks the
// sentinel's bit, which is cleared before the walk, and adds 0 to N_R.

// large sets' ids, item i at desc[i].y (one block per item)
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
// N <= 65535) and |metric| 0 (nothing added to N_R).  The sets of a chunk are
// of similar size (sorted by size within their class) and kmax, the chunk's
// largest, bounds the gather and the walk: groups of 8 entries at or past it
// are sentinels only, and their keys are the sentinel from the start.
// The walk keeps the extremes of the running sum just before and just after
// each hit (in double, c unrounded, as the wavefront path); rounding to T is
// monotone, so the enrichment score is the rounded extreme of larger
// magnitude, and only on equal magnitudes (the reference keeps the first)
// does the lane walk again with the first-strict-maximum rule.
template <class T, int NC, class AF>
__device__ __forceinline__ void es_small_sets(const RankArgs<T> &a,
                                              const uint16_t *pos, AF &A,
                                              int64_t gp, int chunk, int lane) {
    constexpr int R = NC / 2;
    const int slot = chunk * kWave + lane;
    const int sid = a.small_ids[slot];
    const int k = a.small_k[slot];
    const int kmax = __builtin_amdgcn_readfirstlane(a.small_kmax[chunk]);
    uint32_t v[R];
    const uint32_t sent = (uint32_t)a.N * 0x00010001u;
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = sent;
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
        if (t0 >= kmax) continue;   // uniform: keeps the loop fully unrolled
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
            if (t < R) v[t] = (v[t] & 0xffff0000u) | g[u];
            else v[t - R] = (v[t - R] & 0xffffu) | (g[u] << 16);
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
    // k == N: no misses; the reference's decay (-inf) is never applied
    const double dec = k == a.N ? 0.0 : decay;
    const double qnan = __builtin_nan("");
    double L = 0.0, vmin = qnan, vmax = qnan;
#pragma unroll
    for (int t0 = 0; t0 < NC; t0 += 8) {
        if (t0 >= kmax) continue;   // uniform: keeps the loop fully unrolled
        T x[8];
        // an opaque copy of k per group: the compiler would otherwise compute
        // all NC masks t < k at once and keep them live (spills at NC = 128)
        int kk = k;
        asm volatile("" : "+v"(kk));
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = A(rank_at(t0 + u));
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int t = t0 + u;
            const bool valid = t < kk;
            const double before = __builtin_fma((double)(rank_at(t) - t), dec, L);
            vmin = min_skip_nan(vmin, valid ? before : qnan);
            const double c = share(x[u]);   // 0 for a sentinel
            L += c;
            vmax = max_skip_nan(vmax, valid ? before + c : qnan);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    const T hi = (T)vmax, lo = (T)vmin;
    const T ah = hi == hi ? fabs(hi) : T(0);
    const T al = lo == lo ? fabs(lo) : T(0);
    T best = ah > al ? hi : (al > ah ? lo : T(0));
    if (ah == al && ah != T(0)) {
        // equal magnitudes: the first strictly largest candidate, in order
        best = T(0);
        double Sa = 0.0;
#pragma unroll
        for (int t0 = 0; t0 < NC; t0 += 8) {
            if (t0 >= kmax) continue;   // uniform: keeps the loop fully unrolled
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int t = t0 + u;
                es_candidates<T>(Sa, share(A(rank_at(t))), rank_at(t), t, decay, best, t < k);
            }
        }
    }
    if (k == 0) best = (T)((double)a.N * decay);
    if (sid >= 0) a.es[(size_t)gp * a.G + sid] = best;
}

// Sets of 129..256 genes: a pair of lanes each (lanes 2i and 2i+1 of a
// chunk of 32 sets).  Lane 2i gathers the set's entries 0..127, lane 2i+1
// the rest (rows past the set's end hold the sentinel, rank N), and each
// sorts its 128 ranks in registers; lane 2i+1 sorts descending (ascending on
// complemented keys).  The two sequences then form one bitonic sequence of
// 256 ranks, and one compare-exchange of key i of the even lane with key i of
// the odd lane (a DPP move, a packed min and max) leaves the 128 smallest
// ranks on the even lane and the 128 largest on the odd lane, each as a
// bitonic sequence that a size-128 bitonic merge sorts ascending.  Each lane
// then walks its 128 ranks as the one-lane path does, with global hit
// indices (the odd lane's start at 128) and its own running sum from 0; the
// odd lane's extremes are offset by the even lane's final sum, and the pair
// combines its extremes through one more DPP move each.  N_R is the sum of
// the two lanes' partial sums (the same addition on both lanes).
template <class T, class AF>
__device__ __forceinline__ void es_pair_sets(const RankArgs<T> &a, const uint16_t *pos,
                                             AF &A, int64_t gp, int chunk, int lane) {
    constexpr int NC = kSmallSet, R = NC / 2;
    const int slot = chunk * kWave + lane;
    const int sid = a.small_ids[slot];
    const int k = a.small_k[slot];   // the whole set's size, on both lanes
    const bool odd = (lane & 1) != 0;
    const int t_base = odd ? NC : 0;  // set index of the lane's first entry
    uint32_t v[R];
    const uint32_t sent = (uint32_t)a.N * 0x00010001u;
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = sent;
    const uint16_t *gcol = a.small_gm + slot;
    int stride = a.small_stride;
    asm volatile("" : "+s"(stride));
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
            if (t < R) v[t] = (v[t] & 0xffff0000u) | g[u];
            else v[t - R] = (v[t - R] & 0xffffu) | (g[u] << 16);
            nr += (double)x[u];
        }
        asm volatile("" : "+v"(nr));
#pragma unroll
        for (int u = 0; u < 8; ++u) asm volatile("" : "+v"(v[(t0 + u) % R]));
    }
    // the odd lane sorts its complemented keys, i.e. its ranks descending
    const uint32_t flip = odd ? 0xffffffffu : 0u;
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] ^= flip;
    bitonic_sort_u16x2<NC>(v);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        v[r] ^= flip;
        const uint32_t p = dpp_move_any<kDppQuadXor1, 0xf>(v[r]);
        v[r] = odd ? pk_max_u16(v[r], p) : pk_min_u16(v[r], p);
        asm volatile("" : "+v"(v[r]));
    }
    bitonic_step_u16x2<NC, NC, R>(v);
    auto rank_at = [&](int t) -> int {
        return t < R ? (int)(v[t] & 0xffffu) : (int)(v[t - R] >> 16);
    };
    nr = nr + dpp_move_any<kDppQuadXor1, 0xf>(nr);
    const HitShare<T> share((T)nr);
#pragma unroll
    for (int r = 0; r < R; ++r) asm volatile("" : "+v"(v[r]));

    const double decay = sid >= 0 ? (double)a.decay[sid] : 0.0;
    const double dec = k == a.N ? 0.0 : decay;
    const double qnan = __builtin_nan("");
    double L = 0.0, vmin = qnan, vmax = qnan;
#pragma unroll
    for (int t0 = 0; t0 < NC; t0 += 8) {
        T x[8];
        int kk = k - t_base;   // the lane's entries
        asm volatile("" : "+v"(kk));
        int tb = t_base;
        asm volatile("" : "+v"(tb));
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = A(rank_at(t0 + u));
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int t = t0 + u;
            const bool valid = t < kk;
            const double before = __builtin_fma((double)(rank_at(t) - t - tb), dec, L);
            vmin = min_skip_nan(vmin, valid ? before : qnan);
            const double c = share(x[u]);   // 0 for a sentinel
            L += c;
            vmax = max_skip_nan(vmax, valid ? before + c : qnan);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    // the even lane's sum of c offsets the odd lane's running sum
    const double La = dpp_move_any<kDppQuadXor1, 0xf>(L);
    const double off = odd ? La : 0.0;
    vmin += off;
    vmax += off;
    vmin = min_skip_nan(vmin, dpp_move_any<kDppQuadXor1, 0xf>(vmin));
    vmax = max_skip_nan(vmax, dpp_move_any<kDppQuadXor1, 0xf>(vmax));
    const T hi = (T)vmax, lo = (T)vmin;
    const T ah = hi == hi ? fabs(hi) : T(0);
    const T al = lo == lo ? fabs(lo) : T(0);
    T best = ah > al ? hi : (al > ah ? lo : T(0));
    if (ah == al && ah != T(0)) {
        // equal magnitudes: the first strictly largest candidate, in order,
        // on each lane, then the even lane's unless the odd lane's is larger
        T b = T(0);
        double Sa = off;
        // reload keys and |metric| here: reusing the walk's values would keep
        // all 128 of each live across it
        asm volatile("" ::: "memory");
#pragma unroll
        for (int r = 0; r < R; ++r) asm volatile("" : "+v"(v[r]));
#pragma unroll
        for (int t0 = 0; t0 < NC; t0 += 8) {
            int kk = k - t_base;
            asm volatile("" : "+v"(kk));
            int tb = t_base;
            asm volatile("" : "+v"(tb));
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int t = t0 + u;
                es_candidates<T>(Sa, share(A(rank_at(t))), rank_at(t), t + tb, decay, b,
                                 t < kk);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        const T bo = dpp_move_any<kDppQuadXor1, 0xf>(b);   // the partner's
        const T be = odd ? bo : b, bd = odd ? b : bo;        // even's, odd's
        best = fabs(bd) > fabs(be) ? bd : be;
    }
    if (!odd && sid >= 0) a.es[(size_t)gp * a.G + sid] = best;
}

// Phase A of a large set: B gene-id slots per lane (set entries j0 + 64u +
// lane, u < B), the bits of their ranks set in the wavefront's bitmap and
// their |r| summed.  rs covers the set's k entries: slots past them read 0,
// the sentinel (rank N, |r| 0).
template <class T, int B, class AF>
__device__ __forceinline__ void es_mark_hits(__amdgpu_buffer_rsrc_t rs, const uint16_t *pos,
                                             AF &A, uint32_t *bm, int j0, int lane,
                                             double &nr) {
    // byte offset; slot u adds 128u, folded into the loads' immediate offset
    // (an opaque base keeps the compiler from rebasing the slots on the last
    // one, which needs a vector add per slot for the negative offsets).  The
    // range check of the buffer covers voffset + immediate offset.
    int vo = 2 * (j0 + lane);
    asm volatile("" : "+v"(vo));
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
// the last round takes exactly the slots that hold set entries.  The
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
        switch (ns < 8 ? ns : 8) {
        case 8: es_mark_hits<T, 8>(rs, pos, A, bm, j0, lane, nr); break;
        case 7: es_mark_hits<T, 7>(rs, pos, A, bm, j0, lane, nr); break;
        case 6: es_mark_hits<T, 6>(rs, pos, A, bm, j0, lane, nr); break;
        case 5: es_mark_hits<T, 5>(rs, pos, A, bm, j0, lane, nr); break;
        case 4: es_mark_hits<T, 4>(rs, pos, A, bm, j0, lane, nr); break;
        case 3: es_mark_hits<T, 3>(rs, pos, A, bm, j0, lane, nr); break;
        case 2: es_mark_hits<T, 2>(rs, pos, A, bm, j0, lane, nr); break;
        default: es_mark_hits<T, 1>(rs, pos, A, bm, j0, lane, nr); break;
        }
    }
    // LDS operations of one wavefront complete in order: this follows the ORs
    if (lane == 0) atomicAnd(&bm[N >> 5], ~(1u << (N & 31)));
}

// One lane's walk over its hits in rank order, first-strict-maximum rule
// (the tie case of es_large_set).  The lane's W bitmap words hold ranks
// [lane*W*32, (lane+1)*W*32); nz flags its nonempty words and j is the set
// index of its first hit.  S starts at the lane's Sa (sum of c over the
// hits of earlier lanes); every candidate goes through es_candidates.
template <class T, class AF>
__device__ __forceinline__ void es_walk_exact(uint32_t *bm, int W, AF &A,
                                              const HitShare<T> &share, int lane,
                                              uint32_t nz, int hits, int j, double dec,
                                              double &S, T &best) {
    uint32_t cur = 0;
    int wq = 0;        // first rank of the current word
    int jt = j;        // set index of the current hit
    for (int t = 0; t < hits; ++t, ++jt) {
        if (cur == 0u) {
            const int w = lane * W + __builtin_ctz(nz);
            nz &= nz - 1u;
            // one LDS swap reads the word and leaves it clear for the next set
            cur = atomicExch(&bm[w], 0u);
            wq = w * 32;
        }
        const int q = wq + __builtin_ctz(cur);
        cur &= cur - 1u;
        es_candidates<T>(S, share(A(q)), q, jt, dec, best);
    }
}

// The same walk for the lane's extremes, relative to the lane's start.
// With L the lane's running sum of c and t the number of its earlier hits,
// the set's running sum just before the lane's hit at rank q is
// Sa + L + (q - j - t)*decay (Sa: sum of c over earlier lanes, j: their hit
// count) and just after it that plus c.  The walk carries Y = L - t*decay,
// so each hit costs one product with its rank: before = q*decay + Y,
// after = c + before, Y += c - decay; the lane's offset Sa - j*decay is added
// afterwards.  vmin keeps the smallest value before a hit and vmax the
// largest after one (NaN means none yet: v_min / v_max skip it, as the
// reference's comparison skips a NaN running sum).  Every word read is
// cleared.  Returns the final Y, L - hits*decay (hits: the lane's hit count):
// the exclusive wavefront scan of it is each lane's offset Sa - j*decay.
template <class T, class AF>
__device__ __forceinline__ double es_walk_extremes(uint32_t *bm, int W, AF &A, double rn,
                                                   int lane, uint32_t nz,
                                                   double dec, double &vmin, double &vmax) {
    double Y = 0.0;
    uint32_t cur = 0;
    int wq = 0;
    const int base = lane * W;
    if (nz == 0u) return Y;
    // one hit per step; the lane is done when its current word and nz are
    // both empty (one 64-bit compare of the pair)
    do {
        if (cur == 0u) {
            const int w = base + __builtin_ctz(nz);
            nz &= nz - 1u;
            // one LDS swap reads the word and leaves it clear for the next set
            cur = atomicExch(&bm[w], 0u);
            wq = w * 32;
        }
        const int q = wq + __builtin_ctz(cur);
        cur &= cur - 1u;
        const double x = (double)A(q);
        const double before = __builtin_fma((double)q, dec, Y);
        vmin = min_skip_nan(vmin, before);
        vmax = max_skip_nan(vmax, __builtin_fma(x, rn, before));
        Y = __builtin_fma(x, rn, Y) - dec;
    } while ((((uint64_t)nz << 32) | cur) != 0u);
    return Y;
}

// The same walk for |metric| staged in LDS (fp32 runs), with the rank taken
// from the LDS byte address of its |metric|, a_base + 4q: the address serves
// both the gather and the product, as fma((double)(a_base + 4q), decay/4, Y).
// decay/4 is exact, so for a_base = 0 each candidate has the bits of
// fma((double)q, decay, Y); for a_base != 0 the candidates are offset by
// a_base*decay/4, which the caller subtracts.  Each word's address base is
// formed once per refill from the lane's (opaque) base.
__device__ __forceinline__ double es_walk_extremes_lds(uint32_t *bm, int W, uint32_t a_base,
                                                       double rn, int lane, uint32_t nz,
                                                       double dec, double &vmin,
                                                       double &vmax) {
    double Y = 0.0;
    uint32_t cur = 0;
    uint32_t wq4 = 0;   // LDS byte address of |metric| at the word's first rank
    uint32_t *bml = bm + lane * W;                        // the lane's words
    uint32_t al = a_base + (uint32_t)(lane * W) * 128u;   // |metric| of its ranks
    // opaque: one shift-add per refill instead of the compiler's re-derived
    // (w + lane*W)*128 + a_base
    asm volatile("" : "+v"(al));
    const double dec4 = dec * 0.25;   // exact: a power-of-two scaling
    if (nz == 0u) return Y;
    do {
        if (cur == 0u) {
            const uint32_t w = (uint32_t)__builtin_ctz(nz);
            nz &= nz - 1u;
            // one LDS swap reads the word and leaves it clear for the next set
            cur = atomicExch(&bml[w], 0u);
            wq4 = al + w * 128u;
        }
        const uint32_t addr = wq4 + 4u * (uint32_t)__builtin_ctz(cur);
        cur &= cur - 1u;
        const double x =
            (double)*(const __attribute__((address_space(3))) float *)(uintptr_t)addr;
        const double before = __builtin_fma((double)addr, dec4, Y);
        vmin = min_skip_nan(vmin, before);
        vmax = max_skip_nan(vmax, __builtin_fma(x, rn, before));
        Y = __builtin_fma(x, rn, Y) - dec;
    } while ((((uint64_t)nz << 32) | cur) != 0u);
    return Y;
}

// max and min over the 64 lanes at once, returned on every lane: the
// butterfly of wave_sum with the combination folded into DPP-modified
// v_max_f32 / v_min_f32 (IEEE mode: a NaN operand is skipped).  The two
// chains alternate so that each DPP read of a register comes two
// instructions after its write (the s_nop fills the second wait state).
__device__ __forceinline__ void wave_maxmin_f32(float &hi, float &lo) {
    asm volatile(
        "s_nop 1\n\t"
        "v_max_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
        "v_min_f32_dpp %1, %1, %1 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_max_f32_dpp %0, %0, %0 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
        "v_min_f32_dpp %1, %1, %1 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_max_f32_dpp %0, %0, %0 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_min_f32_dpp %1, %1, %1 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_max_f32_dpp %0, %0, %0 row_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_min_f32_dpp %1, %1, %1 row_mirror row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_max_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "v_min_f32_dpp %1, %1, %1 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_max_f32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
        "v_min_f32_dpp %1, %1, %1 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
        "s_nop 1"
        : "+v"(hi), "+v"(lo));
    hi = read_lane(hi, kWave - 1);
    lo = read_lane(lo, kWave - 1);
}

// Larger sets: one wavefront per set.  Phase A sets the k rank bits in the
// wavefront's (N+1)-bit bitmap (lane l owns words l*W .. l*W+W-1, so lanes
// own consecutive rank ranges) and sums N_R.  Each lane then walks its own hits
// once, keeping its sum of c and the extremes of its local running sum; one
// wavefront scan of the sums places every lane's extremes on the set's
// running sum.  Rounding to T is monotone, so the largest (smallest) rounded
// candidate is the rounded largest (smallest) one, and the enrichment score
// is whichever has the larger magnitude.  Only when both magnitudes are
// equal does the order of occurrence decide (the reference keeps the first):
// that case walks the hits again with the first-strict-maximum rule.
template <class T, bool ALDS, class AF>
__device__ __forceinline__ void es_large_set(const RankArgs<T> &a, const float *a_lds,
                                             const uint16_t *pos, AF &A,
                                             uint32_t *bm, int64_t gp, int s,
                                             int o0, int k, double decay, int lane) {
    if (k == 0) {
        // no hits (an empty set reaches this path when every set is scored
        // by a wavefront): the walk only falls, to N*decay at its end, as on
        // the one-lane path
        if (lane == 0) a.es[(size_t)gp * a.G + s] = (T)((double)a.N * decay);
        return;
    }
    const uint16_t *sg = a.large_genes + o0;
    double nr = 0.0;
    es_mark_set<T>(sg, pos, A, bm, k, a.N, lane, nr);
    nr = wave_sum(nr);
    const HitShare<T> share((T)nr);
    // k == N: no misses; the reference's decay (-inf) is never applied
    const double dec = k == a.N ? 0.0 : decay;

    wave_sync_lds();
    // the lane's nonempty words (W <= 32)
    uint32_t nz = 0;
    const int W = a.W;
    // an opaque 1: min(word, 1) stays one v_min_u32 (with a literal 1 the
    // compiler rewrites it as a compare and a select)
    uint32_t one = 1u;
    asm volatile("" : "+v"(one));
#pragma unroll 4
    for (int i = 0; i < W; ++i) nz |= min(bm[lane * W + i], one) << i;
    double vmin = __builtin_nan(""), vmax = __builtin_nan("");
    T best = T(0);
    double off;   // Sa - j*decay of each lane
    if constexpr (ALDS) {
        const uint32_t a_base =
            (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) float *)a_lds;
        off = wave_excl_scan(es_walk_extremes_lds(bm, W, a_base, share.rn, lane, nz, dec,
                                                  vmin, vmax));
        // the walk's frame is offset by a_base * decay / 4 (0 when the rows
        // start at LDS byte 0)
        const double o2 = a_base ? off - (double)a_base * 0.25 * dec : off;
        vmin += o2;
        vmax += o2;
    } else {
        off = wave_excl_scan(es_walk_extremes<T>(bm, W, A, share.rn, lane, nz, dec, vmin,
                                                 vmax));
        vmin += off;
        vmax += off;
    }
    T hi, lo;
    if constexpr (sizeof(T) == 4) {
        float h = (float)vmax, l = (float)vmin;
        wave_maxmin_f32(h, l);
        hi = h;
        lo = l;
    } else {
        hi = wave_reduce_idem((T)vmax, [](T x, T y) { return max_skip_nan(x, y); });
        lo = wave_reduce_idem((T)vmin, [](T x, T y) { return min_skip_nan(x, y); });
    }
    const T ah = hi == hi ? fabs(hi) : T(0);
    const T al = lo == lo ? fabs(lo) : T(0);
    if (ah == al && ah != T(0)) {
        // the walk cleared the bitmap: mark the hits again
        wave_sync_lds();
        double unused = 0.0;
        es_mark_set<T>(sg, pos, A, bm, k, a.N, lane, unused);
        wave_sync_lds();
        int hits = 0;
        for (int i = 0; i < W; ++i) hits += __popc(bm[lane * W + i]);
        const int j = wave_excl_scan(hits);
        double S2 = off + (double)j * dec;   // Sa
        es_walk_exact<T>(bm, W, A, share, lane, nz, hits, j, dec, S2, best);
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
// k_rank_big: the rank rows of k_rank for gene counts beyond one workgroup's
// LDS sort (N > 32,768; 16-bit ids, N <= 65,535).  The same stable LSD radix
// passes with 8-bit digits and ballot-ranked local ranks, but the ids
// ping-pong through a global scratch of 2 x Npad entries per workgroup, and
// each pass sweeps the wavefront's chunks twice -- counting digits, then,
// after the scan over (digit, wavefront), scattering -- instead of holding
// every chunk in registers.  Keys are read from the metric row (L2).
// ===========================================================================
template <class T>
__global__ __launch_bounds__(kRankThreads) void k_rank_big(RankArgs<T> a,
                                                           uint16_t *__restrict__ scratch,
                                                           int Npad) {
    typedef typename KeyOf<T>::type KeyT;
    constexpr int NPASS = (int)sizeof(KeyT);
    __shared__ uint32_t cnt[kRankWaves * 256];
    __shared__ int misc[kRankWaves];
    const int N = a.N;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int w = tid / kWave;
    const bool asc = a.asc != 0;
    const int seg = Npad / kRankWaves;          // a multiple of 64
    const int CH = seg / kWave;
    const KeyT kSentinel = ~(KeyT)0;
    uint16_t *bufA = scratch + (size_t)blockIdx.x * 2 * Npad;
    uint16_t *bufB = bufA + Npad;

    for (int p = blockIdx.x; p < a.Pb; p += gridDim.x) {
        const T *rp = a.r + (size_t)p * N;
        auto key_of = [&](int g) -> KeyT {
            return g < N ? (KeyT)sort_key(rp[g], asc) : kSentinel;
        };
        for (int g = tid; g < Npad; g += kRankThreads) bufA[g] = (uint16_t)g;
        __syncthreads();
        uint16_t *src = bufA, *dst = bufB;
        const int start = w * seg;
        uint32_t *wcnt = cnt + w * 256;
        for (int pass = 0; pass < NPASS; ++pass) {
            const int shift = pass * 8;
            for (int i = lane; i < 256; i += kWave) wcnt[i] = 0;
            wave_sync_lds();
            // sweep 1: digit counts of the wavefront's segment
            for (int c = 0; c < CH; ++c) {
                const uint32_t id = src[start + c * kWave + lane];
                const uint32_t d = (uint32_t)((key_of((int)id) >> shift) & 255u);
                uint64_t peers = ~0ull;
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const bool bit = (d >> b) & 1u;
                    const uint64_t bal = ballot64(bit);
                    peers &= bit ? bal : ~bal;
                }
                if (mbcnt64(peers) == 0) wcnt[d] += (uint32_t)__popcll(peers);
                wave_sync_lds();
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
                if (lane == kWave - 1) misc[w] = (int)inc;
                tot_d = inc - tot_d;
            }
            __syncthreads();
            if (tid < 256) {
                uint32_t base = tot_d;
                for (int w2 = 0; w2 < w; ++w2) base += (uint32_t)misc[w2];
                for (int w2 = 0; w2 < kRankWaves; ++w2) cnt[w2 * 256 + tid] += base;
            }
            __syncthreads();
            // sweep 2: the same chunks in the same order, scattered
            for (int c = 0; c < CH; ++c) {
                const uint32_t id = src[start + c * kWave + lane];
                const uint32_t d = (uint32_t)((key_of((int)id) >> shift) & 255u);
                uint64_t peers = ~0ull;
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const bool bit = (d >> b) & 1u;
                    const uint64_t bal = ballot64(bit);
                    peers &= bit ? bal : ~bal;
                }
                const int before = mbcnt64(peers);
                const uint32_t base = wcnt[d];
                dst[base + (uint32_t)before] = (uint16_t)id;
                wave_sync_lds();
                if (before == 0) wcnt[d] = base + (uint32_t)__popcll(peers);
                wave_sync_lds();
            }
            __syncthreads();
            uint16_t *t = src;
            src = dst;
            dst = t;
        }
        // src holds the gene at each rank (padding ranks >= N last)
        if (a.idx_out)
            for (int i = tid; i < N; i += kRankThreads)
                a.idx_out[(size_t)p * N + i] = src[i];
        if (a.pos_out) {
            T *arow = a.a_glob + (size_t)p * N;
            for (int i = tid; i < N; i += kRankThreads) {
                const int g = src[i];
                arow[i] = (T)__builtin_fabs((double)rp[g]);
                a.pos_out[(size_t)p * N + g] = (uint16_t)i;
            }
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
template <class T, bool ALDS>
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
            // items run longest first, so that the wavefronts' last items
            // (which the workgroup waits for before the next permutation) are
            // short: the chunks from the pair class down to the 8-gene class,
            // then the large sets in set order
            const int n_chunks = a.n_items - a.n_large;
            while (item < a.n_items) {
                // fetch the next work item now; its latency hides behind this one
                int next = 0;
                if (lane == 0) next = atomicAdd(&misc[0], 1);
                if (item >= n_chunks) {
                    const int li = item - n_chunks;
                    // read-only tables through the constant address space:
                    // scalar loads of the uniform descriptor and decay
                    const __attribute__((address_space(4))) int *d =
                        (const __attribute__((address_space(4))) int *)(a.large_desc + li);
                    const T dec = *((const __attribute__((address_space(4))) T *)
                                        a.large_decay + li);
                    es_large_set<T, ALDS>(a, a_lds, pos, A, bm, gp, d[0], d[1], d[2], (double)dec,
                                          lane);
                } else {
                    const int chunk = n_chunks - 1 - item;
                    if (chunk < a.chunk_end8)
                        es_small_sets<T, 8>(a, pos, A, gp, chunk, lane);
                    else if (chunk < a.chunk_end16)
                        es_small_sets<T, 16>(a, pos, A, gp, chunk, lane);
                    else if (chunk < a.chunk_end32)
                        es_small_sets<T, 32>(a, pos, A, gp, chunk, lane);
                    else if (chunk < a.chunk_end64)
                        es_small_sets<T, 64>(a, pos, A, gp, chunk, lane);
                    else if (chunk < a.chunk_end128)
                        es_small_sets<T, 128>(a, pos, A, gp, chunk, lane);
                    else
                        es_pair_sets<T>(a, pos, A, gp, chunk, lane);
                }
                item = __builtin_amdgcn_readfirstlane(next);
            }
        }
        __syncthreads();
    }
}

}  // namespace gsea
