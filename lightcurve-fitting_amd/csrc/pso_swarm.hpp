// pso_swarm.hpp -- batched particle-swarm model fitting for gfx950: the
// device code, instantiated per precision by pso_f32.hip and pso_f64.hip.
//
// One wavefront owns one band ("BatchSource" of the reference) at a time and
// runs the swarms of every requested model on it, one after another:
//   * particles live in lanes (particle p = lane + 64*k, k < PPL);
//   * the band's photometry is staged once into the wave's LDS slice as packed
//     {t, y, var, flag} records and read back as wave-uniform broadcasts;
//   * the per-particle NLL is a lane-private Kahan sum over observations, so
//     no cross-lane traffic is needed inside the observation loop;
//   * the swarm state (position, velocity, personal best) stays in registers
//     for all iterations; the global best is a wave argmin per iteration.
// Semantics follow GpuContext::batch_pso (src/gpu_cuda.rs:368-567) with the
// cost of batch_pso_cost (cuda/models.cu:516-584) plus the host-side prior
// term (src/gpu_cuda.rs:461-480), and the model-selection loop of
// batch_model_select / batch_all_models (src/gpu_cuda.rs:573-637).  Random
// numbers come from Philox4x32-10 keyed by the seed with the band index in
// the counter (initial swarm, and the seed of each particle's xoshiro128**
// velocity-update stream) instead of the reference's sequential SmallRng,
// which makes every band's result independent of batch composition and
// launch order.
#pragma once
#include "pso_kernels.hpp"
#include "model_tables.hpp"
#include "models_fit.hpp"
#include "pso_fast32.hpp"

namespace lcf {

template <typename R> struct ObsT;
template <> struct ObsT<float> { using T = float4; };
template <> struct ObsT<double> { using T = double4; };

// Copy `cnt` packed observations into the wave's LDS slice (coalesced: lane i
// moves records i, i+64, ...).  Single-wave workgroups: __syncthreads() is a
// wave-local ordering point.
template <typename R>
__device__ __forceinline__ void stage_obs(typename ObsT<R>::T* lds,
                                          const typename ObsT<R>::T* __restrict__ g, int cnt) {
    __syncthreads();
    for (int i = lane_id(); i < cnt; i += WAVE) lds[i] = g[i];
    __syncthreads();
}

template <int MODEL, typename R, int PPL>
struct Swarm {
    static constexpr int NP = ModelTraits<MODEL>::NP;
    // FP32 swarms of models 0-6 and MultiBazin use the detections-first LDS
    // tile and the instruction-lean model forms of pso_fast32.hpp.
    static constexpr bool FAST = sizeof(R) == 4 && (MODEL <= 6 || MODEL >= MB_MODEL0);
#ifndef LCF_PSO_F32_UNIT
    static_assert(!FAST, "FP32 swarms are instantiated in pso_f32.hip only (it builds their table)");
#endif
    using M = Math<R>;
    using O = typename ObsT<R>::T;

    R pos[PPL][NP], vel[PPL][NP], pb[PPL][NP], pcost[PPL];
    R gb[NP], gcost;
    R ipw[NP];      // FAST: sqrt(1/2) / population-prior width (0 where the width is 0)
    float inv_nf;   // FAST: 1 / max(n, 1), set once per band

    // FAST cost: the tile holds {t log2 e, y, var, flag} with the `nd`
    // detections first (staged once by the caller when the band fits one
    // tile, else re-staged here per tile).  Four detections per Kahan update,
    // then the upper limits, four per update.  When no particle of the wave
    // can reach a phase test of its model on the tile's time range, the
    // tests are compiled out (CHECK = false).  A non-finite prediction
    // reaches the sum as +inf or NaN and maps to the invalid sentinel below,
    // as the per-observation flag of the general path does.
    template <bool CHECK>
    __device__ __forceinline__ void tile_fast(Kahan<float> (&acc)[PPL], const f32::Fast<MODEL> (&mdl)[PPL],
                                              const float (&se2)[PPL], const O* lds, int nd, int cnt) {
        int i = 0;
        for (; i + 4 <= nd; i += 4) {
            const O r0 = lds[i], r1 = lds[i + 1], r2 = lds[i + 2], r3 = lds[i + 3];
#pragma unroll
            for (int k = 0; k < PPL; k++) acc[k].add(f32::cost_quad<CHECK>(mdl[k], se2[k], r0, r1, r2, r3));
        }
        if (i + 2 <= nd) {
            const O r0 = lds[i], r1 = lds[i + 1];
#pragma unroll
            for (int k = 0; k < PPL; k++) acc[k].add(f32::cost_pair<CHECK>(mdl[k], se2[k], r0, r1));
            i += 2;
        }
        if (i < nd) {
            const O r0 = lds[i];
#pragma unroll
            for (int k = 0; k < PPL; k++) acc[k].add(f32::cost_one<CHECK>(mdl[k], se2[k], r0));
            i++;
        }
        float se2u[PPL];  // sigma_x^2 / 64 against the upper limits' var / 64
#pragma unroll
        for (int k = 0; k < PPL; k++) se2u[k] = se2[k] * (1.0f / 64.0f);
        for (; i + 4 <= cnt; i += 4) {
            const O r0 = lds[i], r1 = lds[i + 1], r2 = lds[i + 2], r3 = lds[i + 3];
#pragma unroll
            for (int k = 0; k < PPL; k++)
                acc[k].add(f32::cost_upper_quad<CHECK>(mdl[k], se2u[k], r0, r1, r2, r3));
        }
        for (; i < cnt; i++) {
            const O r0 = lds[i];
#pragma unroll
            for (int k = 0; k < PPL; k++) acc[k].add(f32::cost_upper<CHECK>(mdl[k], se2u[k], r0));
        }
    }

    __device__ __forceinline__ void costs_fast(R (&cost)[PPL], const O* lds, const O* __restrict__ g,
                                               int n, int tile, int nd_staged, float tmin_staged,
                                               float tmax_staged) {
        f32::Fast<MODEL> mdl[PPL];
        float se2[PPL];
        Kahan<float> acc[PPL];
#pragma unroll
        for (int k = 0; k < PPL; k++) {
            mdl[k].setup(pos[k]);
            se2[k] = f32::ex2((2.0f * f32::L2E) * pos[k][NP - 1]);
            acc[k].init();
        }
        for (int base = 0; base < n; base += tile) {
            const int cnt = n - base < tile ? n - base : tile;
            int nd = nd_staged;
            float tmin = tmin_staged, tmax = tmax_staged;
            if (n > tile) nd = f32::stage_split(const_cast<O*>(lds), g + base, cnt, tmin, tmax);
            bool safe = true;
#pragma unroll
            for (int k = 0; k < PPL; k++) safe = safe && mdl[k].safe(tmin, tmax);
            if (__all(safe)) tile_fast<false>(acc, mdl, se2, lds, nd, cnt);
            else tile_fast<true>(acc, mdl, se2, lds, nd, cnt);
        }
        const float inv_n = inv_nf;
#pragma unroll
        for (int k = 0; k < PPL; k++) {
            float pri = 0.0f;  // sum of z^2 / 2, z = (p - centre) / width
            if constexpr (MODEL < MB_MODEL0) {  // MultiBazin has no population prior
#pragma unroll
                for (int j = 0; j < NP; j++) {
                    float z = (pos[k][j] - (float)c_pc[MODEL][j]) * ipw[j];
                    pri = __builtin_fmaf(z, z, pri);
                }
            }
            float c = n > 0 ? acc[k].s * inv_n : M::big();
            c = __builtin_fmaf(pri * inv_n, inv_n, c);
            if (!(c < M::big())) c = M::big();
            cost[k] = c;
        }
    }

    // NLL/n + prior/n^2 for this lane's particles (cuda/models.cu:540-583 and
    // src/gpu_cuda.rs:461-480).  FP64 follows the reference's arithmetic
    // step for step (plain sum, divisions, exp(ln sigma_x) squared, 1e99 for a
    // non-finite prediction, +inf kept when the sum overflows: an infinite
    // cost never becomes a personal best, 1e99 does).  FP32 uses a Kahan sum
    // and reciprocals, and folds every overflow into its 1e30 sentinel.
    __device__ __forceinline__ void costs(R (&cost)[PPL], const O* lds, const O* __restrict__ g,
                                          int n, int tile, int P, int nd_staged, float tmin_staged,
                                          float tmax_staged) {
        if constexpr (FAST) {
            costs_fast(cost, lds, g, n, tile, nd_staged, tmin_staged, tmax_staged);
            return;
        }
        constexpr bool F64 = sizeof(R) == 8;
        Model<MODEL, R> mdl[PPL];
        R se2[PPL];
        Kahan<R> acc[PPL];
        bool ok[PPL];
#pragma unroll
        for (int k = 0; k < PPL; k++) {
            mdl[k].setup(pos[k]);
            if (F64) {
                R se = M::exp(pos[k][NP - 1]);
                se2[k] = se * se;
            } else {
                se2[k] = M::exp(R(2) * pos[k][NP - 1]);
            }
            acc[k].init();
            ok[k] = n > 0;
        }
        for (int base = 0; base < n; base += tile) {
            int cnt = n - base < tile ? n - base : tile;
            if (n > tile) stage_obs<R>(const_cast<O*>(lds), g + base, cnt);
            for (int i = 0; i < cnt; i++) {
                O o = lds[i];  // wave-uniform address: LDS broadcast
#pragma unroll
                for (int k = 0; k < PPL; k++) {
                    R f;
                    bool good = mdl[k].pred(o.x, f) && M::finite(f);
                    ok[k] = ok[k] && good;
                    R tv = o.z + se2[k];
                    if (F64) {
                        if (o.w != R(0)) {
                            acc[k].s -= log_normal_cdf<R>((o.y - f) / M::sqrt(tv));
                        } else {
                            R d = f - o.y;
                            acc[k].s += d * d / tv + M::log(tv);
                        }
                    } else {
                        R term;
                        if (o.w != R(0)) {
                            term = -log_normal_cdf<R>((o.y - f) * M::rsqrt(tv));
                        } else {
                            R d = f - o.y;
                            term = d * d * M::rcp(tv) + M::log(tv);
                        }
                        acc[k].add(term);
                    }
                }
            }
        }
        R nn = R(n > 0 ? n : 1);
        R inv_n = R(1) / nn;
#pragma unroll
        for (int k = 0; k < PPL; k++) {
            R pri = R(0);
            if constexpr (MODEL < MB_MODEL0) {  // MultiBazin has no population prior
#pragma unroll
                for (int j = 0; j < NP; j++) {
                    R w = (R)c_pw[MODEL][j];
                    if (w > R(0)) {
                        R z = (pos[k][j] - (R)c_pc[MODEL][j]) / w;
                        pri += R(0.5) * z * z;
                    }
                }
            }
            R c;
            if (F64) {
                c = ok[k] ? acc[k].s / nn : R(1e99);
                c = c + pri / (nn * nn);
            } else {
                c = ok[k] ? acc[k].s * inv_n : M::big();
                c = c + pri * inv_n * inv_n;
                if (!(c < M::big())) c = M::big();
            }
            cost[k] = c;
        }
    }

    // Box bounds lo/hi; with `seedp` particle 0 starts there with velocity
    // (hi - lo) 0.02 (2u - 1) (batch_pso_multi_bazin, src/gpu_cuda.rs:767-777).
    // Random numbers: the initial position and velocity from the Philox block
    // (band, particle, 0xFFFFFFFF, rng_tag + dimension pair); the velocity
    // updates from the particle's xoshiro128** stream, seeded by the block
    // (band, particle, 0xFFFFFFFE, rng_tag), r1 and r2 per dimension in order.
    __device__ __forceinline__ void run(const O* lds, const O* __restrict__ g, int n, int tile,
                                        uint32_t band, int P, int max_iters, int stall_iters,
                                        uint64_t seed, uint32_t flags, int& iters_out,
                                        const R (&lo)[NP], const R (&hi)[NP], const R* seedp = nullptr,
                                        uint32_t rng_tag = 0u, int nd_staged = 0, float tmin_staged = 0.0f,
                                        float tmax_staged = 0.0f) {
        const int lane = lane_id();
        if constexpr (FAST && MODEL < MB_MODEL0) {
#pragma unroll
            for (int j = 0; j < NP; j++) {
                const double w = c_pw[MODEL][j];
                ipw[j] = w > 0.0 ? (R)(0.70710678118654752440 / w) : R(0);
            }
        }
        if constexpr (FAST) inv_nf = 1.0f / (float)(n > 0 ? n : 1);
        const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
        constexpr int NPAIR = (NP + 1) / 2;
        Xoshiro128 rng[PPL];
        R vmax[NP];
#pragma unroll
        for (int d = 0; d < NP; d++) vmax[d] = R(0.5) * (hi[d] - lo[d]);
#pragma unroll
        for (int k = 0; k < PPL; k++) {
            uint32_t p = (uint32_t)(lane + WAVE * k);
            const bool seeded = seedp != nullptr && p == 0;
#pragma unroll
            for (int q = 0; q < NPAIR; q++) {
                U4 r = philox4x32_10(band, p, 0xFFFFFFFFu, rng_tag + (uint32_t)q, k0, k1);
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    int d = 2 * q + h;
                    if (d < NP) {
                        R ux = u24<R>(h ? r.z : r.x), uv = u24<R>(h ? r.w : r.y);
                        if (seeded) {
                            pos[k][d] = seedp[d];
                            vel[k][d] = (hi[d] - lo[d]) * R(0.02) * (R(2) * uv - R(1));
                        } else {
                            pos[k][d] = lo[d] + ux * (hi[d] - lo[d]);
                            vel[k][d] = vmax[d] * R(0.2) * (R(2) * uv - R(1));
                        }
                    }
                }
            }
            rng[k].seed(philox4x32_10(band, p, 0xFFFFFFFEu, rng_tag, k0, k1));
#pragma unroll
            for (int d = 0; d < NP; d++) pb[k][d] = R(0);
            pcost[k] = R(INFINITY);
        }
#pragma unroll
        for (int d = 0; d < NP; d++) gb[d] = R(0);
        gcost = R(INFINITY);

        const R w_max = R(0.9), w_min = R(0.4), c1 = R(1.5), c2 = R(1.5);
        const R c1s = c1 * R(1.0 / 16777216.0), c2s = c2 * R(1.0 / 16777216.0);
        const R inv_max = R(1) / R(max_iters);
        R prev = R(INFINITY);
        int stall = 0;
        int it = 0;
        for (; it < max_iters; it++) {
            R w = w_max - (w_max - w_min) * R(it) * inv_max;
            R cost[PPL];
            costs(cost, lds, g, n, tile, P, nd_staged, tmin_staged, tmax_staged);
            // personal bests; lowest cost this iteration (ties: lowest index)
            R bc = R(INFINITY);
            int bi = 0x7fffffff;
#pragma unroll
            for (int k = 0; k < PPL; k++) {
                int p = lane + WAVE * k;
                if (p < P) {
                    if (cost[k] < pcost[k]) {
                        pcost[k] = cost[k];
#pragma unroll
                        for (int d = 0; d < NP; d++) pb[k][d] = pos[k][d];
                    }
                    if (cost[k] < bc || (cost[k] == bc && p < bi)) { bc = cost[k]; bi = p; }
                }
            }
            int win;
            if constexpr (FAST && PPL == 1) {
                // costs are never NaN here (mapped to the sentinel) and the
                // particle index is the lane: the lowest lane holding the
                // minimum is the lexicographic (cost, index) minimum
                win = wave_argmin_lane(bc);
            } else {
                win = wave_argmin<R>(bc, bi);
            }
            R wcost = bcast(bc, win & (WAVE - 1));
            if (wcost < gcost) {
                gcost = wcost;
                int wl = win & (WAVE - 1), wk = win / WAVE;
#pragma unroll
                for (int k = 0; k < PPL; k++)
                    if (k == wk) {
#pragma unroll
                        for (int d = 0; d < NP; d++) gb[d] = bcast(pos[k][d], wl);
                    }
            }
            if ((flags & 1u) && it == 0) prev = gcost;
            // velocity / position update with clamping and wall absorption
#pragma unroll
            for (int k = 0; k < PPL; k++) {
#pragma unroll
                for (int d = 0; d < NP; d++) {
                    // c1 u24 and c2 u24 as one product each: (x >> 8) (1.5 2^-24)
                    // is c1 (x >> 8) 2^-24 exactly (a power-of-two scale)
                    R r1 = R(rng[k].next() >> 8) * c1s;
                    R r2 = R(rng[k].next() >> 8) * c2s;
                    R x = pos[k][d];
                    R v = w * vel[k][d] + r1 * (pb[k][d] - x) + r2 * (gb[d] - x);
                    // clamp to +-vmax; walls absorb (position on the wall, velocity 0)
                    v = fmin(fmax(v, -vmax[d]), vmax[d]);
                    R nx = x + v;
                    const bool inside = nx > lo[d] && nx < hi[d];
                    pos[k][d] = fmin(fmax(nx, lo[d]), hi[d]);
                    vel[k][d] = inside ? v : R(0);
                }
            }
            // stall rule (src/gpu_cuda.rs:537-552); prev starts at +inf as
            // in the reference unless the caller opts into flag bit 0.
            R ap = prev < R(0) ? -prev : prev;
            bool improved = prev - gcost > R(0.01) * (ap > R(1e-10) ? ap : R(1e-10));
            if (improved) { stall = 0; prev = gcost; }
            else if (++stall >= stall_iters) { it++; break; }
        }
        iters_out = it;
    }
};

// The reference's FP64 cost (the FP64 swarm's arithmetic: reference model
// form, 1e99 for a non-finite prediction, NLL / n + prior / n^2) of one
// wave-uniform parameter vector, lanes over the band's FP64 records.
template <int MODEL>
__device__ double cost64_at(const double* p, const double4* __restrict__ g64, int n) {
    constexpr int NP = ModelTraits<MODEL>::NP;
    const int lane = lane_id();
    Model<MODEL, double> mdl;
    mdl.setup(p);
    const double se = ::exp(p[NP - 1]);
    const double se2 = se * se;
    double part = 0.0;
    bool bad = false;
    for (int i = lane; i < n; i += WAVE) {
        const double4 o = g64[i];
        double f;
        const bool good = mdl.pred(o.x, f) && __builtin_isfinite(f);
        if (!good) {
            bad = true;
            continue;
        }
        const double tv = o.z + se2;
        if (o.w != 0.0) {
            part -= log_normal_cdf<double>((o.y - f) / ::sqrt(tv));
        } else {
            const double d = f - o.y;
            part += d * d / tv + ::log(tv);
        }
    }
    bad = __ballot(bad) != 0ull;
    const double sum = wave_sum(part);
    const double nn = (double)(n > 0 ? n : 1);
    double pri = 0.0;
#pragma unroll
    for (int j = 0; j < NP; j++) {
        const double w = c_pw[MODEL][j];
        if (w > 0.0) {
            const double z = (p[j] - c_pc[MODEL][j]) / w;
            pri += 0.5 * z * z;
        }
    }
    double c = (bad || n <= 0) ? 1e99 : sum / nn;
    c = c + pri / (nn * nn);
    return __builtin_isfinite(c) ? c : 1e99;
}

// Fits model MODEL on one band and writes its slot of the per-model outputs.
// With a gate, the band is skipped unless the gate model's cost >= gate
// (batch_model_select, src/gpu_cuda.rs:584-619).
template <int MODEL, typename R, int PPL>
__device__ void fit_band(const PsoArgs& a, typename ObsT<R>::T* lds, int band, int model_slot) {
    using O = typename ObsT<R>::T;
    const int64_t slot = (int64_t)model_slot * a.n_bands + band;
    const int lane = lane_id();
    if (a.gate_cost && !(a.gate_cost[band] >= a.gate)) {
        if (lane < 7) a.model_params[slot * 7 + lane] = 0.0;
        if (lane == 0) { a.model_cost[slot] = INFINITY; a.model_iters[slot] = 0; }
        return;
    }
    const O* g = reinterpret_cast<const O*>(a.obs) + a.offsets[band];
    const int n = (int)(a.offsets[band + 1] - a.offsets[band]);
    int nd = 0;
    float tmin = 0.0f, tmax = 0.0f;
    if constexpr (Swarm<MODEL, R, PPL>::FAST) {
        if (n <= a.tile)
            nd = f32::stage_split(reinterpret_cast<float4*>(lds), reinterpret_cast<const float4*>(g), n, tmin, tmax);
    } else {
        if (n <= a.tile) stage_obs<R>(lds, g, n);
    }
    constexpr int NP = Swarm<MODEL, R, PPL>::NP;
    R lo[NP], hi[NP];
#pragma unroll
    for (int d = 0; d < NP; d++) {
        lo[d] = (R)c_lo[MODEL][d];
        hi[d] = (R)c_hi[MODEL][d];
    }
    Swarm<MODEL, R, PPL> s;
    int iters = 0;
    s.run(lds, g, n, a.tile, (uint32_t)band, a.n_particles, a.max_iters, a.stall_iters, a.seed,
          a.flags, iters, lo, hi, nullptr, 0u, nd, tmin, tmax);
    double cst = (double)s.gcost;
    if (cst >= (double)Math<R>::big()) cst = 1e99;
    double prm = 0.0;  // lane d holds parameter d
#pragma unroll
    for (int d = 0; d < Swarm<MODEL, R, PPL>::NP; d++)
        if (lane == d) prm = (double)s.gb[d];
    if (lane < 7) a.model_params[slot * 7 + lane] = prm;
    if (lane == 0) {
        a.model_cost[slot] = cst;
        a.model_iters[slot] = iters;
    }
}

// After an FP32 swarm launch: the returned best position of every fitted
// band of slot `slot` re-costed in FP64 (cost64_at), so that the cost a fit
// returns -- and the model selection that compares costs -- is the
// reference's at the returned parameters, also where the reference's cost
// jumps (an FP64 exponential overflowing at a phase of 709.78 d, a threshold
// that FP32 times cannot place to better than 1e-4 d).  One wavefront per
// band; gated bands (no iterations) and empty bands keep their sentinel.
// A separate launch, so the FP64 registers do not lower the swarm kernel's
// occupancy.
template <int MODEL>
__global__ __launch_bounds__(64) void pso_cost64_kernel(PsoArgs a) {
    const int band = (int)blockIdx.x;
    if (band >= a.n_bands) return;
    const int64_t slot = (int64_t)a.slot * a.n_bands + band;
    const int n = (int)(a.offsets[band + 1] - a.offsets[band]);
    if (n <= 0 || a.model_iters[slot] <= 0) return;
    constexpr int NP = ModelTraits<MODEL>::NP;
    double p[NP];
#pragma unroll
    for (int d = 0; d < NP; d++) p[d] = a.model_params[slot * 7 + d];
    const double c = cost64_at<MODEL>(p, a.obs64 + a.offsets[band], n);
    if (lane_id() == 0) a.model_cost[slot] = c;
}

// Persistent form: each single-wave workgroup pulls band indices from a
// device queue ordered by decreasing band size (longest-processing-time
// first) until the queue is drained.  One-shot form (PERSISTENT = false):
// workgroup b fits band order[task_base + b]; used by the per-size-class
// schedule.
// Waves per SIMD the kernel is compiled for: the FP32 cost loops of models
// 0-6 with one particle per lane fit 96 registers, so five single-wave
// workgroups share a SIMD (their 8000-byte LDS tiles of 500-point bands fill
// the CU's 160 KB exactly); the other forms keep the compiler's choice.
template <int MODEL, typename R, int PPL>
constexpr int pso_waves_per_simd() {
    return (sizeof(R) == 4 && MODEL <= 6 && PPL == 1) ? 5 : 1;
}

template <int MODEL, typename R, int PPL, bool PERSISTENT>
__global__ __launch_bounds__(64, (pso_waves_per_simd<MODEL, R, PPL>())) void pso_model_kernel(PsoArgs a) {
    extern __shared__ unsigned char lds_raw[];
    using O = typename ObsT<R>::T;
    O* lds = reinterpret_cast<O*>(lds_raw);
    if (PERSISTENT) {
        for (;;) {
            const int t = wave_next_task(a.queue);
            if (t >= a.n_tasks) break;
            fit_band<MODEL, R, PPL>(a, lds, a.order[t], a.slot);
            wave_reconverge();
        }
    } else {
        int t = a.task_base + (int)blockIdx.x;
        if (t < a.n_tasks) fit_band<MODEL, R, PPL>(a, lds, a.order[t], a.slot);
    }
}

// One persistent launch over mixed model families: task t is (band
// order[t / n], model slot t % n), so the queue hands out every model of the
// longest band first and the bands in decreasing size, and a wave switches
// model family from one task to the next.  FP32 swarms of models 0-6 with
// one particle per lane (the instruction-lean cost loops, all within the
// same 96 registers, so the switch costs no occupancy); no gate (a gated
// model waits for its band's first model).
template <typename R, int PPL>
__global__ __launch_bounds__(64, (pso_waves_per_simd<0, R, PPL>())) void pso_mixed_kernel(PsoArgs a, PsoMix mx) {
    static_assert(pso_waves_per_simd<0, R, PPL>() == pso_waves_per_simd<6, R, PPL>(),
                  "the mixed launch shares one occupancy across models 0-6");
    extern __shared__ unsigned char lds_raw[];
    using O = typename ObsT<R>::T;
    O* lds = reinterpret_cast<O*>(lds_raw);
    for (;;) {
        const int t = wave_next_task(a.queue);
        if (t >= a.n_tasks) break;
        const int k = t / mx.n, mi = t - k * mx.n;
        const int band = a.order[k];
        switch (mx.model[mi]) {
            case 0: fit_band<0, R, PPL>(a, lds, band, mi); break;
            case 1: fit_band<1, R, PPL>(a, lds, band, mi); break;
            case 2: fit_band<2, R, PPL>(a, lds, band, mi); break;
            case 3: fit_band<3, R, PPL>(a, lds, band, mi); break;
            case 4: fit_band<4, R, PPL>(a, lds, band, mi); break;
            case 5: fit_band<5, R, PPL>(a, lds, band, mi); break;
            default: fit_band<6, R, PPL>(a, lds, band, mi); break;
        }
        wave_reconverge();
    }
}

// ---------------------------------------------------------------------------
// Greedy MultiBazin (batch_pso_multi_bazin, src/gpu_cuda.rs:650-963): for
// K = 1..4 a swarm over 4K + 2 parameters (no population prior); from K = 2
// particle 0 starts at the K - 1 solution plus a component seeded at the
// largest detection residual; BIC = 2 cost n + (4K + 2) ln n selects K
// (strict '<'), and a band stops once K's BIC exceeds K - 1's by more than 2.
// One wavefront per band runs all K in registers.
// ---------------------------------------------------------------------------
struct MbState {
    double prev[MB_MAX_PARAMS];   // gbest of the last K fitted (wave-uniform)
    double best[MB_MAX_PARAMS];
    double best_bic, best_cost;
    double pk_cost[MB_MAX_K], pk_bic[MB_MAX_K];
    int best_k;
    bool stopped;
};

template <int K, typename R, int PPL>
__device__ __forceinline__ void mb_fit_k(const MbArgs& a, typename ObsT<R>::T* lds, int band, int n,
                                         MbState& st, int nd, float tmin, float tmax) {
    using O = typename ObsT<R>::T;
    constexpr int MODEL = MB_MODEL0 + K - 1;
    constexpr int NP = 4 * K + 2;
    if (st.stopped) {
        st.pk_cost[K - 1] = NAN;
        st.pk_bic[K - 1] = NAN;
        return;
    }
    R lo[NP], hi[NP];
#pragma unroll
    for (int c = 0; c < K; c++) {
        lo[4 * c] = R(-3); lo[4 * c + 1] = (R)a.t_lo; lo[4 * c + 2] = R(-2); lo[4 * c + 3] = R(-2);
        hi[4 * c] = R(3); hi[4 * c + 1] = (R)a.t_hi; hi[4 * c + 2] = R(5); hi[4 * c + 3] = R(6);
    }
    lo[4 * K] = R(-0.3); hi[4 * K] = R(0.3);
    lo[4 * K + 1] = R(-5); hi[4 * K + 1] = R(0);
    R seedv[NP];
    if (K > 1) {
        // residual peak of the K - 1 fit over detections, in FP64 with the
        // reference's expression (src/gpu_cuda.rs:718-745); ties: first index
        const double4* g64 = a.obs64 + a.offsets[band];
        const double* pv = st.prev;
        double bv = INFINITY;
        int bi = 0x7fffffff;
        for (int i = lane_id(); i < n; i += WAVE) {
            double4 o = g64[i];
            if (o.w != 0.0) continue;
            double pred = 0.0;
#pragma unroll
            for (int c = 0; c < K - 1; c++) {
                double aa = ::exp(pv[4 * c]), t0 = pv[4 * c + 1];
                double tr = ::exp(pv[4 * c + 2]), tf = ::exp(pv[4 * c + 3]);
                double dt = o.x - t0;
                double sig = 1.0 / (1.0 + ::exp(-dt / tr));
                pred += aa * ::exp(-dt / tf) * sig;
            }
            pred += pv[4 * (K - 1)];
            double resid = o.y - pred;
            if (!(-resid < bv)) continue;   // NaN never wins; strict: first index kept
            bv = -resid;
            bi = i;
        }
        int win = wave_argmin<double>(bv, bi);
        double wv = bcast(bv, win & (WAVE - 1));
        int wi = bcast_int(bi, win & (WAVE - 1));
        double peak = -wv;
        int pidx = (wv < INFINITY) ? wi : 0;
        if (!(wv < INFINITY)) peak = -INFINITY;
        double tpk = n > 0 ? g64[pidx].x : 0.0;
#pragma unroll
        for (int d = 0; d < 4 * (K - 1); d++) seedv[d] = (R)pv[d];
        seedv[4 * K - 4] = (R)::log(fmax(peak, 1e-10));
        seedv[4 * K - 3] = (R)tpk;
        seedv[4 * K - 2] = R(1);
        seedv[4 * K - 1] = R(1);
        seedv[4 * K] = (R)pv[4 * (K - 1)];
        seedv[4 * K + 1] = (R)pv[4 * (K - 1) + 1];
#pragma unroll
        for (int d = 0; d < NP; d++) seedv[d] = seedv[d] < lo[d] ? lo[d] : (seedv[d] > hi[d] ? hi[d] : seedv[d]);
    }
    const O* g = reinterpret_cast<const O*>(a.obs) + a.offsets[band];
    Swarm<MODEL, R, PPL> s;
    int iters = 0;
    s.run(lds, g, n, a.tile, (uint32_t)band, a.n_particles, a.max_iters, a.stall_iters, a.seed, 0u, iters,
          lo, hi, K > 1 ? seedv : nullptr, 0x4D420000u + ((uint32_t)K << 8), nd, tmin, tmax);
    double cost = (double)s.gcost;
    if (sizeof(R) == 4 && cost >= (double)Math<R>::big()) cost = 1e99;
    const double nn = (double)n;
    const double bic = 2.0 * cost * nn + (double)NP * ::log(nn);
    st.pk_cost[K - 1] = cost;
    st.pk_bic[K - 1] = bic;
    if (bic < st.best_bic) {
        st.best_bic = bic;
        st.best_cost = cost;
        st.best_k = K;
#pragma unroll
        for (int d = 0; d < MB_MAX_PARAMS; d++) st.best[d] = d < NP ? (double)s.gb[d] : 0.0;
    }
#pragma unroll
    for (int d = 0; d < MB_MAX_PARAMS; d++) st.prev[d] = d < NP ? (double)s.gb[d] : 0.0;
    if (K > 1 && bic > st.pk_bic[K - 2] + 2.0) st.stopped = true;
}

template <typename R, int PPL>
__device__ void mb_band(const MbArgs& a, typename ObsT<R>::T* lds, int band) {
    using O = typename ObsT<R>::T;
    const int lane = lane_id();
    const O* g = reinterpret_cast<const O*>(a.obs) + a.offsets[band];
    const int n = (int)(a.offsets[band + 1] - a.offsets[band]);
    int nd = 0;
    float tmin = 0.0f, tmax = 0.0f;
    if constexpr (sizeof(R) == 4) {  // the FAST cost loop's detections-first tile
        if (n <= a.tile)
            nd = f32::stage_split(reinterpret_cast<float4*>(lds), reinterpret_cast<const float4*>(g), n, tmin, tmax);
    } else {
        if (n <= a.tile) stage_obs<R>(lds, g, n);
    }
    MbState st;
    st.best_bic = INFINITY;
    st.best_cost = INFINITY;
    st.best_k = 1;
    st.stopped = false;
#pragma unroll
    for (int d = 0; d < MB_MAX_PARAMS; d++) { st.prev[d] = 0.0; st.best[d] = 0.0; }
    mb_fit_k<1, R, PPL>(a, lds, band, n, st, nd, tmin, tmax);
    mb_fit_k<2, R, PPL>(a, lds, band, n, st, nd, tmin, tmax);
    mb_fit_k<3, R, PPL>(a, lds, band, n, st, nd, tmin, tmax);
    mb_fit_k<4, R, PPL>(a, lds, band, n, st, nd, tmin, tmax);
    double prm = 0.0;  // lane d holds parameter d
#pragma unroll
    for (int d = 0; d < MB_MAX_PARAMS; d++)
        if (lane == d) prm = st.best[d];
    if (lane < MB_MAX_PARAMS) a.params[(int64_t)band * MB_MAX_PARAMS + lane] = prm;
    double pc = 0.0, pb = 0.0;
#pragma unroll
    for (int k = 0; k < MB_MAX_K; k++)
        if (lane == k) { pc = st.pk_cost[k]; pb = st.pk_bic[k]; }
    if (lane < MB_MAX_K) {
        a.per_k_cost[(int64_t)band * MB_MAX_K + lane] = pc;
        a.per_k_bic[(int64_t)band * MB_MAX_K + lane] = pb;
    }
    if (lane == 0) {
        a.best_k[band] = st.best_k;
        a.cost[band] = st.best_cost;
        a.bic[band] = st.best_bic;
    }
}

template <typename R, int PPL>
__global__ __launch_bounds__(64) void multi_bazin_kernel(MbArgs a) {
    extern __shared__ unsigned char lds_raw[];
    using O = typename ObsT<R>::T;
    O* lds = reinterpret_cast<O*>(lds_raw);
    for (;;) {
        const int t = wave_next_task(a.queue);
        if (t >= a.n_tasks) break;
        mb_band<R, PPL>(a, lds, a.order[t]);
        wave_reconverge();
    }
}

template <typename R>
static const void* pick_mb_r(int ppl) {
    if (ppl == 1) return (const void*)multi_bazin_kernel<R, 1>;
    if (ppl == 2) return (const void*)multi_bazin_kernel<R, 2>;
    if (ppl == 4) return (const void*)multi_bazin_kernel<R, 4>;
    return nullptr;
}

template <int MODEL, typename R, int PPL>
static const void* kernel_ptr(bool persistent) {
    return persistent ? (const void*)pso_model_kernel<MODEL, R, PPL, true>
                      : (const void*)pso_model_kernel<MODEL, R, PPL, false>;
}

template <typename R, int PPL>
static const void* kernel_ptr_m(int model, bool persistent) {
    switch (model) {
        case 0: return kernel_ptr<0, R, PPL>(persistent);
        case 1: return kernel_ptr<1, R, PPL>(persistent);
        case 2: return kernel_ptr<2, R, PPL>(persistent);
        case 3: return kernel_ptr<3, R, PPL>(persistent);
        case 4: return kernel_ptr<4, R, PPL>(persistent);
        case 5: return kernel_ptr<5, R, PPL>(persistent);
        case 6: return kernel_ptr<6, R, PPL>(persistent);
        case 7: return kernel_ptr<7, R, PPL>(persistent);
        default: return nullptr;
    }
}

static const void* pick_cost64(int model) {
    switch (model) {
        case 0: return (const void*)pso_cost64_kernel<0>;
        case 1: return (const void*)pso_cost64_kernel<1>;
        case 2: return (const void*)pso_cost64_kernel<2>;
        case 3: return (const void*)pso_cost64_kernel<3>;
        case 4: return (const void*)pso_cost64_kernel<4>;
        case 5: return (const void*)pso_cost64_kernel<5>;
        case 6: return (const void*)pso_cost64_kernel<6>;
        default: return nullptr;
    }
}

template <typename R>
static const void* pick_kernel_r(int model, int ppl, bool persistent) {
    if (ppl == 1) return kernel_ptr_m<R, 1>(model, persistent);
    if (ppl == 2) return kernel_ptr_m<R, 2>(model, persistent);
    if (ppl == 4) return kernel_ptr_m<R, 4>(model, persistent);
    return nullptr;
}

}  // namespace lcf
