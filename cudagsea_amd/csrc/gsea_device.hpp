// gsea_device.hpp -- device-side building blocks shared by the GSEA kernels:
// the permutation RNG, the logarithm of the log2-ratio metrics, sort-key
// transforms and wavefront (64-lane) primitives for CDNA4 (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gsea {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Permutation RNG: the reference seeds a fast KISS generator per permutation
// with hash rounds and shuffles the labels Fisher-Yates
// (rngpu/rngpu.hpp:27-38,145-178,288-303; seed = perm + 42,
// correlate_genes.cuh:50-54, configuration.cuh:29).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t seed_hash(uint32_t x) {
    x = (x + 0x7ed55d16u) + (x << 12);
    x = (x ^ 0xc761c23cu) ^ (x >> 19);
    x = (x + 0x165667b1u) + (x << 5);
    x = (x + 0xd3a2646cu) ^ (x << 9);
    x = (x + 0xfd7046c5u) + (x << 3);
    x = (x ^ 0xb55a4f09u) ^ (x >> 16);
    return x;
}

struct KissState {
    uint32_t x, y, z, w, c;
};

__device__ __forceinline__ KissState kiss_seed(uint32_t seed) {
    KissState s;
    s.w = seed ? seed : 0xffffffffu;
    s.x = s.y = s.z = 0;
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        s.x = seed_hash(s.w);
        s.y = seed_hash(s.x);
        s.z = seed_hash(s.y);
        s.w = seed_hash(s.z);
    }
    s.c = 0;
    return s;
}

__device__ __forceinline__ uint32_t kiss_next(KissState &s) {
    s.y ^= s.y << 5;
    s.y ^= s.y >> 7;
    s.y ^= s.y << 22;
    const uint32_t t = s.z + s.w + s.c;
    s.z = s.w;
    s.c = t >> 31;
    s.w = t & 0x7fffffffu;
    s.x += 1411392427u;
    return s.x + s.y + s.w;
}

// ---------------------------------------------------------------------------
// Logarithm of the log2-ratio metrics: log(1+f) = 2 atanh(s), s = f/(2+f),
// x = m 2^e, m in [sqrt(1/2), sqrt(2)), ln 2 split into head and tail.  Only
// IEEE +,-,*,/ in a fixed order (the build disables contraction), so host and
// device produce identical bits.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double log_f64(double x) {
    if (x != x) return __builtin_nan("");
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    uint64_t u = __double_as_longlong(x);
    int e = 0;
    if ((u >> 52) == 0) {
        x = x * 18014398509481984.0;
        u = __double_as_longlong(x);
        e = -54;
    }
    e += (int)((u >> 52) & 0x7ffu) - 1023;
    u = (u & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m = __longlong_as_double(u);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
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
    const double lm = f - s * (f - R);
    const double de = (double)e;
    return de * 6.93147180369123816490e-01 +
           (lm + de * 1.90821492927058770002e-10);
}

__device__ __forceinline__ float log_f32(float x) {
    if (x != x) return __builtin_nanf("");
    return (float)log_f64((double)x);
}

__device__ __forceinline__ float log_t(float x) { return log_f32(x); }
__device__ __forceinline__ double log_t(double x) { return log_f64(x); }
__device__ __forceinline__ float sqrt_t(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double sqrt_t(double x) { return __builtin_sqrt(x); }

// math_helpers.cuh:10-13
template <class T>
__device__ __forceinline__ T cuabs(T x) { return x < T(0) ? -x : x; }

// ---------------------------------------------------------------------------
// Sort keys: ascending unsigned order of the key is the stated rank order
// (descending metric unless `asc`; -0 below +0; NaN canonicalised above +inf;
// ties broken by ascending gene id by the stable radix passes).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t sort_key(float v, bool asc) {
    uint32_t u = (v != v) ? 0x7fc00000u : __float_as_uint(v);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return asc ? u : ~u;
}

__device__ __forceinline__ uint64_t sort_key(double v, bool asc) {
    uint64_t u = (v != v) ? 0x7ff8000000000000ull
                          : (uint64_t)__double_as_longlong(v);
    u = (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
    return asc ? u : ~u;
}

// |value| recovered from a key
__device__ __forceinline__ float key_abs(uint32_t k, bool asc) {
    uint32_t t = asc ? k : ~k;
    uint32_t u = (t & 0x80000000u) ? (t & 0x7fffffffu) : ~t;
    return __uint_as_float(u & 0x7fffffffu);
}

// dword 3 of a raw buffer resource on gfx9-family devices (32-bit data
// format, no swizzle, bounds checked against num_records bytes)
constexpr int kBufferRsrcWord3 = 0x00020000;

// ---------------------------------------------------------------------------
// wavefront primitives
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ uint64_t ballot64(bool p) {
    return (uint64_t)__ballot(p);
}

// popcount(mask & lanes below me)
__device__ __forceinline__ int mbcnt64(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- cross-lane moves on the DPP path (VALU, no LDS traffic) -------------
// dpp_move<CTRL, ROWS>(v): every lane of the rows selected by ROWS takes v
// from the lane CTRL names; lanes with no such source, and lanes of the other
// rows, get 0 (all-zero bits).  64-bit values move as two 32-bit halves.
// KEEP: those lanes keep their own v instead (for idempotent combinations).
template <int CTRL, int ROWS, bool KEEP>
__device__ __forceinline__ uint32_t dpp32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(KEEP ? (int)v : 0, (int)v, CTRL, ROWS,
                                                 0xf, false);
}

template <int CTRL, int ROWS, class V, bool KEEP = false>
__device__ __forceinline__ V dpp_move(V v) {
    static_assert(sizeof(V) == 4 || sizeof(V) == 8, "32- or 64-bit lane values");
    if constexpr (sizeof(V) == 4) {
        return __builtin_bit_cast(
            V, dpp32<CTRL, ROWS, KEEP>(__builtin_bit_cast(uint32_t, v)));
    } else {
        const uint64_t b = __builtin_bit_cast(uint64_t, v);
        const uint64_t lo = dpp32<CTRL, ROWS, KEEP>((uint32_t)b);
        const uint64_t hi = dpp32<CTRL, ROWS, KEEP>((uint32_t)(b >> 32));
        return __builtin_bit_cast(V, lo | (hi << 32));
    }
}

// dpp_move_bc<CTRL>(v): every lane takes v from the lane CTRL names, lanes
// with no such source get 0 (bound_ctrl), all rows are written: one
// v_mov_b32_dpp per 32 bits, no initialisation of the destination.
// dpp_move_any<CTRL, ROWS>(v): the same for the rows selected by ROWS; the
// lanes of the other rows hold unspecified values (for reductions that
// never read them).
template <int CTRL, int ROWS, bool BC, class V>
__device__ __forceinline__ V dpp_mov_raw(V v) {
    static_assert(sizeof(V) == 4 || sizeof(V) == 8, "32- or 64-bit lane values");
    if constexpr (sizeof(V) == 4) {
        return __builtin_bit_cast(
            V, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, ROWS, 0xf, BC));
    } else {
        const uint64_t b = __builtin_bit_cast(uint64_t, v);
        const uint64_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)b, CTRL, ROWS,
                                                               0xf, BC);
        const uint64_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), CTRL,
                                                               ROWS, 0xf, BC);
        return __builtin_bit_cast(V, lo | (hi << 32));
    }
}
template <int CTRL, class V>
__device__ __forceinline__ V dpp_move_bc(V v) {
    return dpp_mov_raw<CTRL, 0xf, true>(v);
}
template <int CTRL, int ROWS, class V>
__device__ __forceinline__ V dpp_move_any(V v) {
    return dpp_mov_raw<CTRL, ROWS, false>(v);
}

constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114,
              kDppRowShr8 = 0x118, kDppWaveShr1 = 0x138, kDppRowBcast15 = 0x142,
              kDppRowBcast31 = 0x143, kDppQuadXor1 = 0xb1, kDppQuadXor2 = 0x4e,
              kDppRowHalfMirror = 0x141, kDppRowMirror = 0x140;

// value of lane l, as a wave-uniform value
template <class V>
__device__ __forceinline__ V read_lane(V v, int l) {
    static_assert(sizeof(V) == 4 || sizeof(V) == 8, "32- or 64-bit lane values");
    if constexpr (sizeof(V) == 4) {
        return __builtin_bit_cast(
            V, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
    } else {
        const uint64_t b = __builtin_bit_cast(uint64_t, v);
        const uint64_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
        const uint64_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
        return __builtin_bit_cast(V, lo | (hi << 32));
    }
}

// inclusive prefix combination in lane order, op(left, right) associative:
// Hillis-Steele within each row of 16 lanes, then the row totals carried by
// the row broadcasts.  Lanes without a left neighbour see a 0 operand, so
// op(0, x) must be x for the values that occur.
template <class V, class Op>
__device__ __forceinline__ V wave_incl_scan_op(V x, Op op) {
    x = op(dpp_move_bc<kDppRowShr1>(x), x);
    x = op(dpp_move_bc<kDppRowShr2>(x), x);
    x = op(dpp_move_bc<kDppRowShr4>(x), x);
    x = op(dpp_move_bc<kDppRowShr8>(x), x);
    x = op(dpp_move<kDppRowBcast15, 0xa>(x), x);
    x = op(dpp_move<kDppRowBcast31, 0xc>(x), x);
    return x;
}

template <class V>
__device__ __forceinline__ V wave_incl_scan(V x) {
    return wave_incl_scan_op(x, [](V l, V r) { return l + r; });
}

// sum over the 64 lanes in a fixed order, the same value on every lane: a
// butterfly within each row (lanes xor 1, xor 2, the mirrored half-row, the
// mirrored row), then the row sums carried into lane 63 by the broadcasts
template <class V>
__device__ __forceinline__ V wave_sum(V v) {
    v = dpp_move_any<kDppQuadXor1, 0xf>(v) + v;
    v = dpp_move_any<kDppQuadXor2, 0xf>(v) + v;
    v = dpp_move_any<kDppRowHalfMirror, 0xf>(v) + v;
    v = dpp_move_any<kDppRowMirror, 0xf>(v) + v;
    v = dpp_move_any<kDppRowBcast15, 0xa>(v) + v;
    v = dpp_move_any<kDppRowBcast31, 0xc>(v) + v;
    return read_lane(v, kWave - 1);
}

// op over all 64 lanes for an associative, commutative op (min, max),
// returned on every lane: the butterfly of wave_sum
template <class V, class Op>
__device__ __forceinline__ V wave_reduce_idem(V x, Op op) {
    x = op(dpp_move_any<kDppQuadXor1, 0xf>(x), x);
    x = op(dpp_move_any<kDppQuadXor2, 0xf>(x), x);
    x = op(dpp_move_any<kDppRowHalfMirror, 0xf>(x), x);
    x = op(dpp_move_any<kDppRowMirror, 0xf>(x), x);
    x = op(dpp_move_any<kDppRowBcast15, 0xa>(x), x);
    x = op(dpp_move_any<kDppRowBcast31, 0xc>(x), x);
    return read_lane(x, kWave - 1);
}

// exclusive prefix sum over lanes in lane order (lane 0 gets 0)
template <class V>
__device__ __forceinline__ V wave_excl_scan(V v) {
    return dpp_move_bc<kDppWaveShr1>(wave_incl_scan(v));
}

}  // namespace gsea
