// lcf_device.hpp -- device-side building blocks shared by the gfx950 kernels:
// precision-dispatched math, Philox4x32-10, wave64 reductions, Kahan sums.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lcf {

constexpr int WAVE = 64;

// Largest argument for which exp() is finite in IEEE binary64.  The reference
// evaluates every model in FP64, so a term whose FP64 exponential overflows is
// non-finite there; the kernels test arguments against this bound in both
// precisions to reproduce exactly those cases.
constexpr double EXP_OVF64 = 709.782712893384;

// ---------------------------------------------------------------------------
// Math per precision.  FP32 uses the hardware transcendental unit directly
// (v_exp_f32 / v_log_f32 / v_rcp_f32 / v_sqrt_f32: base-2, ~1 ulp);
// FP64 uses the device math library.
// ---------------------------------------------------------------------------
template <typename R> struct Math;

template <> struct Math<float> {
    static __device__ __forceinline__ float exp(float x) {
        return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
    }
    static __device__ __forceinline__ float log(float x) {
        return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
    }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
    static __device__ __forceinline__ bool finite(float x) { return __builtin_isfinite(x); }
    // Sentinel standing in for the reference's 1e99 "invalid" cost.
    static __device__ __forceinline__ float big() { return 1e30f; }
};

template <> struct Math<double> {
    static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
    static __device__ __forceinline__ double log(double x) { return ::log(x); }
    static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
    static __device__ __forceinline__ double rsqrt(double x) { return 1.0 / ::sqrt(x); }
    static __device__ __forceinline__ bool finite(double x) { return __builtin_isfinite(x); }
    static __device__ __forceinline__ double big() { return 1e99; }
};

// log(1 + e) for e in [0, 1], accurate also for tiny e.
template <typename R>
__device__ __forceinline__ R log1p_unit(R e) {
    using M = Math<R>;
    // Below 2^-7 use the series e - e^2/2 + e^3/3 - e^4/4 (rel. err < 2^-30).
    R s = e * (R(1) - e * (R(0.5) - e * (R(1) / R(3) - e * R(0.25))));
    R l = M::log(R(1) + e);
    return e < R(0.0078125) ? s : l;
}

template <>
__device__ __forceinline__ double log1p_unit<double>(double e) { return ::log1p(e); }

// softplus(x) = log(1 + exp(x)) evaluated without overflow.
template <typename R>
__device__ __forceinline__ R softplus(R x) {
    using M = Math<R>;
    R ax = x < R(0) ? -x : x;
    R l = log1p_unit<R>(M::exp(-ax));
    return (x > R(0) ? x : R(0)) + l;
}

// 1 - exp(-y) for y >= 0, accurate for small y.
template <typename R>
__device__ __forceinline__ R one_minus_exp_neg(R y) {
    using M = Math<R>;
    R s = y * (R(1) - y * (R(0.5) - y * (R(1) / R(6) - y * (R(1) / R(24)))));
    R d = R(1) - M::exp(-y);
    return y < R(0.03125) ? s : d;
}

template <>
__device__ __forceinline__ double one_minus_exp_neg<double>(double y) { return -::expm1(-y); }

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Counter (c0..c3), key (k0, k1).
// ---------------------------------------------------------------------------
struct U4 { uint32_t x, y, z, w; };

__host__ __device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                     uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        // one 32 x 32 -> 64-bit product per word (v_mad_u64_u32 on the
        // device: one instruction for both halves)
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{c0, c1, c2, c3};
}

// xoshiro128** (Blackman and Vigna, "Scrambled linear pseudorandom number
// generators", ACM TOMS 2021): a particle's velocity-update stream, seeded
// from one Philox block (an all-zero block becomes {1, 0, 0, 0}).  Ten
// full-rate integer operations per 32-bit output (the multiplications by 5
// and 9 are shift-adds), against ten rounds of 32-bit multiplies per
// Philox block.
struct Xoshiro128 {
    uint32_t s0, s1, s2, s3;
    __host__ __device__ __forceinline__ void seed(const U4& r) {
        s0 = r.x; s1 = r.y; s2 = r.z; s3 = r.w;
        if ((s0 | s1 | s2 | s3) == 0u) s0 = 1u;
    }
    __host__ __device__ __forceinline__ uint32_t next() {
        const uint32_t m = s1 * 5u;
        const uint32_t out = ((m << 7) | (m >> 25)) * 9u;
        const uint32_t t = s1 << 9;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = (s3 << 11) | (s3 >> 21);
        return out;
    }
};

// 24-bit uniform in [0,1): exactly representable in both precisions, so the
// FP32 and FP64 swarms (and the oracle) consume identical random numbers.
template <typename R>
__device__ __forceinline__ R u24(uint32_t x) {
    return R(x >> 8) * R(1.0 / 16777216.0);
}

// ---------------------------------------------------------------------------
// Wave64 helpers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int bcast_int(int v, int src_lane) {
    return __builtin_amdgcn_readlane(v, src_lane);
}
__device__ __forceinline__ float bcast(float v, int src_lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src_lane));
}
__device__ __forceinline__ double bcast(double v, int src_lane) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src_lane);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), src_lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Lexicographic (value, index) minimum across the wave.  Callers map NaN to
// +inf beforehand.  Returns the winning index, wave-uniform.
template <typename R>
__device__ __forceinline__ int wave_argmin(R v, int idx) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        R ov = __shfl_xor(v, off, WAVE);
        int oi = __shfl_xor(idx, off, WAVE);
        bool take = (ov < v) || (ov == v && oi < idx);
        v = take ? ov : v;
        idx = take ? oi : idx;
    }
    return __builtin_amdgcn_readfirstlane(idx);
}

template <typename R>
__device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Cross-lane moves without LDS.  swap32: lanes 32..63 of a trade places with
// lanes 0..31 of b (v_permlane32_swap_b32); swap16: the odd 16-lane rows of
// a trade places with the even rows of b (v_permlane16_swap_b32); dpp<CTRL>:
// a DPP move within 16-lane rows.  All lanes must be active.
__device__ __forceinline__ void swap32(float& a, float& b) {
    auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap16(float& a, float& b) {
    auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap32(double& a, double& b) {
    uint64_t x = (uint64_t)__double_as_longlong(a), y = (uint64_t)__double_as_longlong(b);
    auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)x, (uint32_t)y, false, false);
    auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(x >> 32), (uint32_t)(y >> 32), false, false);
    a = __longlong_as_double((long long)(((uint64_t)hi[0] << 32) | lo[0]));
    b = __longlong_as_double((long long)(((uint64_t)hi[1] << 32) | lo[1]));
}
__device__ __forceinline__ void swap16(double& a, double& b) {
    uint64_t x = (uint64_t)__double_as_longlong(a), y = (uint64_t)__double_as_longlong(b);
    auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)x, (uint32_t)y, false, false);
    auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(x >> 32), (uint32_t)(y >> 32), false, false);
    a = __longlong_as_double((long long)(((uint64_t)hi[0] << 32) | lo[0]));
    b = __longlong_as_double((long long)(((uint64_t)hi[1] << 32) | lo[1]));
}
constexpr int DPP_QP_1032 = 0xB1;        // quad_perm [1,0,3,2]: lane ^ 1
constexpr int DPP_QP_2301 = 0x4E;        // quad_perm [2,3,0,1]: lane ^ 2
constexpr int DPP_ROW_HALF_MIRROR = 0x141;  // 7 - lane within 8-lane halves
constexpr int DPP_ROW_ROR8 = 0x128;      // rotate by 8 within 16-lane rows: lane ^ 8
template <int CTRL> __device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <int CTRL> __device__ __forceinline__ double dpp(double v) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xf, 0xf, true);
    int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Lowest lane holding the wave minimum of v (no NaN; all lanes active):
// a min reduction by permlane swaps and DPP moves, then a ballot of the
// lanes equal to it.  Wave-uniform.
__device__ __forceinline__ int wave_argmin_lane(float v) {
    float a = v, b = v;
    swap32(a, b);
    float m = __builtin_fminf(a, b);  // lanes l and l ^ 32
    a = m;
    b = m;
    swap16(a, b);
    m = __builtin_fminf(a, b);  // and l ^ 16
    m = __builtin_fminf(m, dpp<DPP_ROW_ROR8>(m));         // and l ^ 8
    m = __builtin_fminf(m, dpp<DPP_ROW_HALF_MIRROR>(m));  // 7 - l within 8-lane halves
    m = __builtin_fminf(m, dpp<DPP_QP_2301>(m));          // l ^ 2
    m = __builtin_fminf(m, dpp<DPP_QP_1032>(m));          // l ^ 1: all 64 lanes
    const unsigned long long eq = __ballot(v == m);
    return eq ? (int)__builtin_ctzll(eq) : 0;
}

// Sixteen wave sums at once (a reduce-scatter): returns in lane l the sum
// over all 64 lanes of v[l >> 2].  Each stage halves the values a lane
// carries, pairing lanes that differ in one bit (bit 5, 4, 3, then a mirror
// partner that differs in bit 2) and ends with a full sum over the quad:
// 8 + 4 + 2 + 1 + 2 cross-lane moves where sixteen butterflies take 96.
template <typename R>
__device__ __forceinline__ R wave_sum16(const R (&v)[16]) {
    const int lane = lane_id();
    R w8[8], w4[4], w2[2];
#pragma unroll
    for (int i = 0; i < 8; i++) {  // lanes 0..31 keep values 0..7, lanes 32..63 values 8..15
        R a = v[i], b = v[8 + i];
        swap32(a, b);
        w8[i] = a + b;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {  // even rows keep the lower four, odd rows the upper
        R a = w8[i], b = w8[4 + i];
        swap16(a, b);
        w4[i] = a + b;
    }
    const bool b3 = (lane & 8) != 0, b2 = (lane & 4) != 0;
#pragma unroll
    for (int i = 0; i < 2; i++) {
        R keep = b3 ? w4[2 + i] : w4[i], send = b3 ? w4[i] : w4[2 + i];
        w2[i] = keep + dpp<DPP_ROW_ROR8>(send);
    }
    R keep = b2 ? w2[1] : w2[0], send = b2 ? w2[0] : w2[1];
    R w = keep + dpp<DPP_ROW_HALF_MIRROR>(send);
    w += dpp<DPP_QP_2301>(w);
    w += dpp<DPP_QP_1032>(w);
    return w;
}

// Next task of a persistent single-wave workgroup from a device work queue.
// Every lane takes part in the atomic (lane 0 adds 1, the others 0) and the
// value lane 0 received is read back with readlane, so the fetch has no
// branch on the lane index.  The obvious form `if (lane == 0) t =
// atomicAdd(q, 1); t = shfl(t, 0)` is unsafe at -O3: jump threading merges
// that branch with a lane-0 branch at the end of the loop body, lanes 1..63
// then re-enter the loop without lane 0 fetching a ticket and spin forever
// on a stale index (the loop nest of the compiled kernel shows it).
__device__ __forceinline__ int wave_next_task(int* queue) {
    int mine = atomicAdd(queue, lane_id() == 0 ? 1 : 0);
    return __builtin_amdgcn_readlane(mine, 0);
}

// Convergence point at the end of a loop body whose lanes must stay together
// (cross-lane operations inside, lane-dependent branches around them): a
// block holding a convergent operation is never duplicated, so no lane
// condition known at the end of one iteration can be threaded into the next.
__device__ __forceinline__ void wave_reconverge() { __builtin_amdgcn_wave_barrier(); }

// Kahan-compensated accumulator.
template <typename R>
struct Kahan {
    R s, c;
    __device__ __forceinline__ void init() { s = R(0); c = R(0); }
    __device__ __forceinline__ void add(R x) {
        R y = x - c;
        R t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

}  // namespace lcf
