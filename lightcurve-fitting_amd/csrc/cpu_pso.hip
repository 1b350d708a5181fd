// cpu_pso.hip -- the CPU path's swarm fits on the device, FP64.
//
// The reference's CPU fits (fit_parametric, src/parametric.rs:3563-4131) use
// pso_minimize / pso_minimize_seeded (:1160-1266, 1646-1757): the global best
// is updated after every particle (asynchronous), so particle p moves with
// the best that particles 0..p-1 of the same iteration produced.  That order
// is kept here: one wavefront per job walks the particles in order, lanes
// take the search dimensions for the move and the observations for the cost
// (LDS-staged packed records), and the cost is the lanes' strided partial
// sums reduced by a fixed xor butterfly, so every job is deterministic and
// the oracle reproduces it bit for bit.  Job kinds:
//   CPSO_SWARM_BAND   PsoCost::cost_from_slice (:1303-1370): Villar
//                     constraint, MetzgerKN rescale, sum of
//                     (pred - f)^2 / v + ln v and -ln Phi for upper limits,
//                     + population prior / n, all / n; optionally with t0
//                     held fixed outside the search vector
//                     (profile_t0_search, :1767-1849);
//   CPSO_SWARM_JOINT  fit_parametric_multiband's joint Villar cost
//                     (:3905-4000): reference band + offsets, constraint per
//                     band, band-balanced likelihood, population prior on the
//                     reference and relative priors on the offsets;
//   CPSO_SWARM_MULTIBAZIN  multi_bazin_cost (:983-1006) of fit_multi_bazin
//                     (:1008-1157): K Bazin components as written
//                     (bazin_component_eval, :919-929) plus the baseline, no
//                     prior, 1e99 on a non-finite prediction, / n;
//   CPSO_POLISH       lbfgs_refine (:1453-1497): L-BFGS (7 pairs, at most
//                     max_iters iterations, gradient norm 1e-6) on the
//                     clamped cost with the projected analytic gradient of
//                     PsoCost::gradient (:1383-1446), kept only if it lowers
//                     the cost; the line search is backtracking Armijo.
// Random numbers: Philox4x32-10 keyed by the job's seed, counter (stream,
// tag, iteration, particle x 32 + dimension / 2).
#include "cpu_pso.hpp"
#include "metzger_batch.hpp"
#include "model_tables.hpp"
#include "models_grad.hpp"

namespace lcf {

namespace {

constexpr int NTH = 64;
constexpr int LBFGS_M = 7;

// relative_band_priors (src/parametric.rs:3771-3785): mean, sigma per Villar parameter
#define LN10_ 2.30258509299404568402
__constant__ double c_rel_mean[7] = {-0.015 * LN10_, -0.001, -0.095 * LN10_, -1.181, -0.195 * LN10_,
                                     -0.323 * LN10_, -0.025 * LN10_};
__constant__ double c_rel_sig[7] = {0.08 * LN10_, 0.008, 0.25 * LN10_, 1.5, 0.35 * LN10_, 0.25 * LN10_,
                                    0.30 * LN10_};
#undef LN10_

struct Smem {
    double xs[CPSO_MAX_DIM];   // position under evaluation
    double gt[MKN_GRID];
    double gl[MKN_GRID];
    double rho[LBFGS_M];
    int flag[4];
};

__device__ __forceinline__ bool wave_any(bool b) { return __ballot(b) != 0ull; }

__device__ __forceinline__ double wave_maxd(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, WAVE));
    return v;
}

__device__ __forceinline__ int t0_index(int m) { return (m == 0 || m == 2) ? 2 : (m == 1 || m == 7) ? 3 : 1; }

// Villar validity (villar_constraint, src/parametric.rs:3790-3798; the same
// expression inside PsoCost): 0 when valid.
__device__ __forceinline__ double villar_constraint(double lg, double beta, double ltr, double ltf) {
    const double gamma = ::exp(lg), tr = ::exp(ltr), tf = ::exp(ltf);
    const double c1 = fmax(gamma * beta - 1.0, 0.0);
    const double c2 = fmax(::exp(-gamma / tr) * (tf / tr - 1.0) - 1.0, 0.0);
    const double c3 = fmax(beta * tf - 1.0 + beta * gamma, 0.0);
    return c1 + c2 + c3;
}

// One band's NLL partial of this lane (observations lane, lane + 64, ...),
// with the non-finite flag; MetzgerKN in its batch form with the peak
// rescale.  Returns the butterfly sum (uniform), or NaN when a prediction is
// non-finite.
template <int M>
__device__ double band_nll(const double* p, const double4* o, int n, Smem& sm) {
    constexpr int NP = ModelTraits<M>::NP;
    const int lane = lane_id();
    const double se = ::exp(p[NP - 1]);
    const double se2 = se * se;
    double part = 0.0;
    bool bad = false;
    if constexpr (M == 7) {
        double ph = 0.01;
        for (int i = lane; i < n; i += WAVE) ph = fmax(ph, o[i].x - p[3]);
        ph = wave_maxd(ph);
        __syncthreads();
        if (lane == 0) sm.flag[1] = metzger_grid(p, ph, sm.gt, sm.gl) ? 1 : 0;
        __syncthreads();
        const bool live = sm.flag[1] != 0;
        double mx = -INFINITY;
        for (int i = lane; i < n; i += WAVE)
            if (o[i].w == 0.0) mx = fmax(mx, live ? metzger_interp(sm.gt, sm.gl, o[i].x - p[3]) : 0.0);
        mx = wave_maxd(mx);
        const double sc = (mx > 1e-10 && __builtin_isfinite(mx)) ? fmin(fmax(1.0 / mx, 0.1), 10.0) : 1.0;
        for (int i = lane; i < n; i += WAVE) {
            const double4 ob = o[i];
            const double f = (live ? metzger_interp(sm.gt, sm.gl, ob.x - p[3]) : 0.0) * sc;
            if (!__builtin_isfinite(f)) { bad = true; continue; }
            const double tv = ob.z + se2;
            if (ob.w != 0.0) part -= log_normal_cdf<double>((ob.y - f) / ::sqrt(tv));
            else {
                const double d = f - ob.y;
                part += d * d / tv + ::log(tv);
            }
        }
    } else {
        Model<M, double> mdl;
        mdl.setup(p);
        for (int i = lane; i < n; i += WAVE) {
            const double4 ob = o[i];
            double f;
            (void)mdl.pred(ob.x, f);
            if (!__builtin_isfinite(f)) { bad = true; continue; }
            const double tv = ob.z + se2;
            if (ob.w != 0.0) part -= log_normal_cdf<double>((ob.y - f) / ::sqrt(tv));
            else {
                const double d = f - ob.y;
                part += d * d / tv + ::log(tv);
            }
        }
    }
    if (wave_any(bad)) return NAN;
    return wave_sum(part);
}

// PsoCost::cost_from_slice of the full parameter vector p.
template <int M>
__device__ double pso_cost(const double* p, const double4* o, int n, bool prior, Smem& sm) {
    constexpr int NP = ModelTraits<M>::NP;
    if (M == 1) {
        const double c = villar_constraint(p[2], p[1], p[4], p[5]);
        if (c > 0.0) return 1e12 + 10000.0 * c;
    }
    double nll = band_nll<M>(p, o, n, sm);
    if (__builtin_isnan(nll)) return 1e99;
    const double nn = (double)(n > 1 ? n : 1);
    if (prior) {
        double lp = 0.0;
#pragma unroll
        for (int j = 0; j < NP; j++) {
            const double w = c_pw[M][j];
            if (w > 0.0) {
                const double z = (p[j] - c_pc[M][j]) / w;
                lp += 0.5 * z * z;
            }
        }
        nll += lp / nn;
    }
    return nll / nn;
}

// The search vector in sm.xs -> full parameters (t0 inserted when fixed).
template <int M>
__device__ void full_params(const CpuPsoJob& J, const double* xs, double (&p)[ModelTraits<M>::NP]) {
    constexpr int NP = ModelTraits<M>::NP;
    const int t0i = t0_index(M);
#pragma unroll
    for (int j = 0; j < NP; j++) {
        if (!J.fixed_t0) p[j] = xs[j];
        else p[j] = j < t0i ? xs[j] : (j == t0i ? J.t0_value : xs[j - 1]);
    }
}

template <int M>
__device__ double band_job_cost(const CpuPsoJob& J, const double4* o, int n, Smem& sm) {
    double p[ModelTraits<M>::NP];
    full_params<M>(J, sm.xs, p);
    return pso_cost<M>(p, o, n, J.use_prior != 0, sm);
}

__device__ double band_cost_dispatch(const CpuPsoJob& J, const double4* o, int n, Smem& sm) {
    switch (J.model) {
        case 0: return band_job_cost<0>(J, o, n, sm);
        case 1: return band_job_cost<1>(J, o, n, sm);
        case 2: return band_job_cost<2>(J, o, n, sm);
        case 3: return band_job_cost<3>(J, o, n, sm);
        case 4: return band_job_cost<4>(J, o, n, sm);
        case 5: return band_job_cost<5>(J, o, n, sm);
        case 6: return band_job_cost<6>(J, o, n, sm);
        default: return band_job_cost<7>(J, o, n, sm);
    }
}

// fit_parametric_multiband's joint cost (src/parametric.rs:3905-4000).
__device__ double joint_cost(const CpuPsoArgs& a, const CpuPsoJob& J, const double4* o, Smem& sm) {
    const int B = J.n_bands;
    const int64_t base = a.offsets[J.band];
    const int n_tot = (int)(a.offsets[J.band + B] - base);
    const double* xs = sm.xs;
    double ref[7];
#pragma unroll
    for (int j = 0; j < 7; j++) ref[j] = xs[j];
    const double rc = villar_constraint(ref[2], ref[1], ref[4], ref[5]);
    if (rc > 0.0) return 1e12 + 10000.0 * rc;
    double bal = 0.0;
    for (int b = 0; b < B; b++) {
        const int s0 = (int)(a.offsets[J.band + b] - base), nb = (int)(a.offsets[J.band + b + 1] - base) - s0;
        double p[7];
#pragma unroll
        for (int j = 0; j < 7; j++) p[j] = b == 0 ? ref[j] : ref[j] + xs[7 * b + j];
        if (b > 0) {
            const double c = villar_constraint(p[2], p[1], p[4], p[5]);
            if (c > 0.0) return 1e12 + 10000.0 * c;
        }
        const double nll = band_nll<1>(p, o + s0, nb, sm);
        if (__builtin_isnan(nll)) return 1e99;
        if (nb > 0) bal += (nll / (double)nb) * ((double)n_tot / (double)B);
    }
    const double nt = (double)(n_tot > 1 ? n_tot : 1);
    double lp = 0.0;
#pragma unroll
    for (int j = 0; j < 7; j++) {
        const double w = c_pw[1][j];
        if (w > 0.0) {
            const double z = (ref[j] - c_pc[1][j]) / w;
            lp += 0.5 * z * z;
        }
    }
    for (int b = 1; b < B; b++)
#pragma unroll
        for (int j = 0; j < 7; j++) {
            const double sg = c_rel_sig[j];
            if (sg > 0.0) {
                const double z = (xs[7 * b + j] - c_rel_mean[j]) / sg;
                lp += 0.5 * z * z;
            }
        }
    return (bal + lp) / nt;
}

// multi_bazin_cost (src/parametric.rs:983-1006) of K components: lanes over
// observations, the reference's component expression and summation order.
__device__ double multi_bazin_job_cost(int K, const double* xs, const double4* o, int n) {
    const int lane = lane_id();
    const double se = ::exp(xs[4 * K + 1]);
    const double se2 = se * se;
    const double B = xs[4 * K];
    double part = 0.0;
    bool bad = false;
    for (int i = lane; i < n; i += WAVE) {
        const double4 ob = o[i];
        double f = 0.0;
        for (int c = 0; c < K; c++) {
            const double a = ::exp(xs[4 * c]), tr = ::exp(xs[4 * c + 2]), tf = ::exp(xs[4 * c + 3]);
            const double dt = ob.x - xs[4 * c + 1];
            const double e_fall = ::exp(-dt / tf);
            const double sig = 1.0 / (1.0 + ::exp(-dt / tr));
            f += a * e_fall * sig;
        }
        f += B;
        if (!__builtin_isfinite(f)) { bad = true; continue; }
        const double tv = ob.z + se2;
        if (ob.w != 0.0) part -= log_normal_cdf<double>((ob.y - f) / ::sqrt(tv));
        else {
            const double d = f - ob.y;
            part += d * d / tv + ::log(tv);
        }
    }
    if (wave_any(bad)) return 1e99;
    return wave_sum(part) / (double)(n > 1 ? n : 1);
}

// Cost of the vector in sm.xs (all lanes, after a barrier).
__device__ double job_cost(const CpuPsoArgs& a, const CpuPsoJob& J, const double4* o, int n, Smem& sm) {
    if (J.kind == CPSO_SWARM_MULTIBAZIN) return multi_bazin_job_cost(J.model, sm.xs, o, n);
    return J.kind == CPSO_SWARM_JOINT ? joint_cost(a, J, o, sm) : band_cost_dispatch(J, o, n, sm);
}

__device__ __forceinline__ void draw(const CpuPsoJob& J, uint32_t slot, int p, int d, double& u1, double& u2) {
    const uint32_t tag = (uint32_t)J.model | ((uint32_t)J.kind << 8) | (J.fixed_t0 ? 0x10000u : 0u);
    const U4 r = philox4x32_10(J.stream, tag, slot, (uint32_t)p * 32u + (uint32_t)(d >> 1), (uint32_t)J.seed,
                               (uint32_t)(J.seed >> 32));
    u1 = u24<double>((d & 1) ? r.z : r.x);
    u2 = u24<double>((d & 1) ? r.w : r.y);
}

// pso_minimize / pso_minimize_seeded (src/parametric.rs:1160-1266, 1646-1757).
__device__ void swarm(const CpuPsoArgs& a, const CpuPsoJob& J, int jid, const double4* o, int n, double* slot,
                      Smem& sm) {
    const int lane = lane_id();
    const int D = J.dim, P = J.n_particles;
    const bool own = lane < D;
    const double lo = own ? a.lo[(int64_t)jid * CPSO_MAX_DIM + lane] : 0.0;
    const double hi = own ? a.hi[(int64_t)jid * CPSO_MAX_DIM + lane] : 0.0;
    const double vmax = 0.5 * (hi - lo);
    double* pos = slot;
    double* vel = slot + (int64_t)P * D;
    double* pb = vel + (int64_t)P * D;
    double* pbc = pb + (int64_t)P * D;
    double gb = 0.0, gcost = INFINITY;
    for (int p = 0; p < P; p++) {
        if (own) {
            double u1, u2;
            draw(J, 0u, p, lane, u1, u2);
            const double x = (J.seeded && p == 0) ? a.x0[(int64_t)jid * CPSO_MAX_DIM + lane] : lo + u1 * (hi - lo);
            const double v = vmax * 0.2 * (2.0 * u2 - 1.0);
            pos[p * D + lane] = x;
            vel[p * D + lane] = v;
            pb[p * D + lane] = x;
            sm.xs[lane] = x;
        }
        __syncthreads();
        const double c = job_cost(a, J, o, n, sm);
        if (lane == 0) pbc[p] = c;
        if (c < gcost) {
            gcost = c;
            if (own) gb = sm.xs[lane];
        }
        __syncthreads();
    }
    double prev = gcost;
    int stall = 0, it = 0;
    const double inv = 1.0 / (double)J.max_iters;
    for (; it < J.max_iters;) {
        const double w = 0.9 - (0.9 - 0.4) * (double)it * inv;
        for (int p = 0; p < P; p++) {
            if (own) {
                double u1, u2;
                draw(J, (uint32_t)it + 1u, p, lane, u1, u2);
                const double x = pos[p * D + lane], v = vel[p * D + lane], b = pb[p * D + lane];
                double vn = w * v + 1.5 * u1 * (b - x) + 1.5 * u2 * (gb - x);
                vn = vn < -vmax ? -vmax : (vn > vmax ? vmax : vn);
                const double np_ = x + vn;
                double xn, vv;
                if (np_ <= lo) { xn = lo; vv = 0.0; }
                else if (np_ >= hi) { xn = hi; vv = 0.0; }
                else { xn = np_; vv = vn; }
                pos[p * D + lane] = xn;
                vel[p * D + lane] = vv;
                sm.xs[lane] = xn;
            }
            __syncthreads();
            const double c = job_cost(a, J, o, n, sm);
            const double pc = pbc[p];
            if (c < pc) {
                if (own) pb[p * D + lane] = sm.xs[lane];
                if (lane == 0) pbc[p] = c;
                if (c < gcost) {
                    gcost = c;
                    if (own) gb = sm.xs[lane];
                }
            }
            __syncthreads();
        }
        it++;
        const bool improved = prev - gcost > 0.01 * fmax(fabs(prev), 1e-10);
        if (improved) {
            stall = 0;
            prev = gcost;
        } else if (++stall >= J.stall_iters) {
            break;
        }
    }
    // every lane writes its column: zeros beyond the search dimensions
    a.out_x[(int64_t)jid * CPSO_MAX_DIM + lane] = own ? gb : 0.0;
    if (lane == 0) {
        a.out_cost[jid] = gcost;
        a.out_iters[jid] = it;
    }
}

// PsoCost::gradient (src/parametric.rs:1383-1446) at p; zero vector when a
// prediction is non-finite.  Lane j < NP returns component j.
template <int M>
__device__ double pso_grad_lane(const double* p, const double4* o, int n) {
    constexpr int NP = ModelTraits<M>::NP;
    constexpr int SE = NP - 1;
    const int lane = lane_id();
    const double se = ::exp(p[SE]);
    const double se2 = se * se;
    ModelGrad<M, double> mdl;
    mdl.setup(p);
    double g[NP];
#pragma unroll
    for (int j = 0; j < NP; j++) g[j] = 0.0;
    bool bad = false;
    for (int i = lane; i < n; i += WAVE) {
        const double4 ob = o[i];
        double f, mg[NP];
        (void)mdl.pred_grad(ob.x, f, mg);
        if (!__builtin_isfinite(f)) { bad = true; continue; }
        const double tv = ob.z + se2;
        if (ob.w != 0.0) {
            const double sv = ::sqrt(tv);
            const double z = (ob.y - f) / sv;
            const double phi = ::exp(-0.5 * z * z) / ::sqrt(2.0 * 3.14159265358979323846);
            const double zz = z / 1.41421356237309504880;
            const double az = fabs(zz);
            const double t = 1.0 / (1.0 + 0.3275911 * az);
            const double poly = t * (0.254829592 + t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 +
                                t * 1.061405429))));
            const double val = 1.0 - poly * ::exp(-az * az);
            const double cdf = 0.5 * (1.0 + (zz >= 0.0 ? val : -val));
            const double ratio = phi / fmax(cdf, 1e-300);
#pragma unroll
            for (int j = 0; j < NP; j++) g[j] += ratio * mg[j] / sv;
            g[SE] += ratio * z * se2 / tv;
        } else {
            const double d = f - ob.y;
            const double sc = 2.0 * d / tv;
#pragma unroll
            for (int j = 0; j < NP; j++) g[j] += sc * mg[j];
            g[SE] += (-d * d / (tv * tv) + 1.0 / tv) * (2.0 * se2);
        }
    }
    const bool any_bad = wave_any(bad);
    const double nn = (double)(n > 1 ? n : 1);
    double mine = 0.0;
#pragma unroll
    for (int j = 0; j < NP; j++) {
        const double s = wave_sum(g[j]);
        if (lane == j) mine = any_bad ? 0.0 : s / nn;
    }
    return mine;
}

__device__ double grad_dispatch(int m, const double* p7, const double4* o, int n) {
    switch (m) {
        case 0: return pso_grad_lane<0>(p7, o, n);
        case 1: return pso_grad_lane<1>(p7, o, n);
        case 2: return pso_grad_lane<2>(p7, o, n);
        case 3: return pso_grad_lane<3>(p7, o, n);
        case 4: return pso_grad_lane<4>(p7, o, n);
        case 5: return pso_grad_lane<5>(p7, o, n);
        default: return pso_grad_lane<6>(p7, o, n);
    }
}

// lbfgs_refine (src/parametric.rs:1453-1497) for one band and model.
__device__ void polish(const CpuPsoArgs& a, const CpuPsoJob& J, int jid, const double4* o, int n, double* slot,
                       Smem& sm) {
    const int lane = lane_id();
    const int D = J.dim;
    const bool own = lane < D;
    const int64_t r0 = (int64_t)jid * CPSO_MAX_DIM;
    const double lo = own ? a.lo[r0 + lane] : 0.0, hi = own ? a.hi[r0 + lane] : 0.0;
    const double x0 = own ? a.x0[r0 + lane] : 0.0;
    double* S = slot;                    // [LBFGS_M][64]
    double* Y = slot + LBFGS_M * CPSO_MAX_DIM;
    // bounded cost / projected gradient at x (this lane's component)
    auto bcost = [&](double x) {
        __syncthreads();
        if (own) sm.xs[lane] = fmin(fmax(x, lo), hi);
        __syncthreads();
        return band_cost_dispatch(J, o, n, sm);
    };
    auto bgrad = [&](double x) {
        __syncthreads();
        if (own) sm.xs[lane] = fmin(fmax(x, lo), hi);
        __syncthreads();
        double p7[7];
#pragma unroll
        for (int j = 0; j < 7; j++) p7[j] = j < D ? sm.xs[j] : 0.0;
        double g = grad_dispatch(J.model, p7, o, n);
        if (!own) g = 0.0;
        else if ((x <= lo && g > 0.0) || (x >= hi && g < 0.0)) g = 0.0;
        return g;
    };
    auto dot = [&](double u, double v) { return wave_sum(own ? u * v : 0.0); };
    double x = x0;
    double f = bcost(x);
    double g = bgrad(x);
    double bx = x, bf = f;
    int nm = 0, head = 0, it = 0;
    for (; it < J.max_iters; it++) {
        if (::sqrt(dot(g, g)) < 1e-6) break;
        // two-loop recursion, newest pair first
        double q = g, al[LBFGS_M];
        for (int k = 0; k < nm; k++) {
            const int pk = (head - 1 - k + LBFGS_M) % LBFGS_M;
            al[k] = sm.rho[pk] * dot(S[pk * CPSO_MAX_DIM + lane], q);
            q -= al[k] * Y[pk * CPSO_MAX_DIM + lane];
        }
        if (nm > 0) {
            const int pk = (head - 1 + LBFGS_M) % LBFGS_M;
            const double yv = Y[pk * CPSO_MAX_DIM + lane];
            q *= dot(S[pk * CPSO_MAX_DIM + lane], yv) / dot(yv, yv);
        }
        for (int k = nm - 1; k >= 0; k--) {
            const int pk = (head - 1 - k + LBFGS_M) % LBFGS_M;
            const double b = sm.rho[pk] * dot(Y[pk * CPSO_MAX_DIM + lane], q);
            q += S[pk * CPSO_MAX_DIM + lane] * (al[k] - b);
        }
        double d = -q;
        double gd = dot(g, d);
        if (!(gd < 0.0)) {
            nm = 0;
            d = -g;
            gd = dot(g, d);
        }
        double step = 1.0, xt = x, ft = f;
        bool ok = false;
        for (int ls = 0; ls < 20; ls++) {
            xt = x + step * d;
            ft = bcost(xt);
            if (__builtin_isfinite(ft) && ft <= f + 1e-4 * step * gd) { ok = true; break; }
            step *= 0.5;
        }
        if (!ok) break;
        const double gt = bgrad(xt);
        const double s = xt - x, y = gt - g;
        const double sy = dot(s, y);
        if (sy > 1e-10) {
            if (own) {
                S[head * CPSO_MAX_DIM + lane] = s;
                Y[head * CPSO_MAX_DIM + lane] = y;
            }
            __syncthreads();
            if (lane == 0) sm.rho[head] = 1.0 / sy;
            __syncthreads();
            head = (head + 1) % LBFGS_M;
            nm = nm + 1 < LBFGS_M ? nm + 1 : LBFGS_M;
        }
        x = xt;
        f = ft;
        g = gt;
        if (f < bf) { bf = f; bx = x; }
    }
    const double xc = fmin(fmax(bx, lo), hi);
    const double fc = bcost(xc);
    const bool keep = fc < J.start_cost && __builtin_isfinite(fc);
    a.out_x[r0 + lane] = own ? (keep ? xc : x0) : 0.0;
    if (lane == 0) {
        a.out_cost[jid] = keep ? fc : J.start_cost;
        a.out_iters[jid] = it;
    }
}

}  // namespace

__global__ __launch_bounds__(NTH) void cpu_pso_kernel(CpuPsoArgs a) {
    extern __shared__ double4 cp_obs[];
    __shared__ Smem sm;
    double* slot = a.scratch + (int64_t)blockIdx.x * a.slot_doubles;
    for (;;) {
        const int t = wave_next_task(a.queue);
        if (t >= a.n_jobs) break;
        const CpuPsoJob J = a.jobs[t];
        const int64_t base = a.offsets[J.band];
        const int nb = J.kind == CPSO_SWARM_JOINT ? J.n_bands : 1;
        const int n = (int)(a.offsets[J.band + nb] - base);
        const double4* g = reinterpret_cast<const double4*>(a.obs) + base;
        __syncthreads();
        for (int i = lane_id(); i < n; i += WAVE) cp_obs[i] = g[i];
        __syncthreads();
        if (J.kind == CPSO_POLISH) polish(a, J, t, cp_obs, n, slot, sm);
        else swarm(a, J, t, cp_obs, n, slot, sm);
        wave_reconverge();
    }
}

hipError_t launch_cpu_pso(const CpuPsoArgs& a, int grid, hipStream_t st) {
    if (grid <= 0 || a.max_obs < 0 || a.max_obs > CPSO_MAX_OBS) return hipErrorInvalidValue;
    void* args[] = {const_cast<CpuPsoArgs*>(&a)};
    return hipLaunchKernel((const void*)cpu_pso_kernel, dim3(grid), dim3(NTH), args,
                           sizeof(double4) * (size_t)(a.max_obs > 0 ? a.max_obs : 1), st);
}

hipError_t cpu_pso_occupancy(int max_obs, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)cpu_pso_kernel, NTH,
                                                        sizeof(double4) * (size_t)(max_obs > 0 ? max_obs : 1));
}

}  // namespace lcf
