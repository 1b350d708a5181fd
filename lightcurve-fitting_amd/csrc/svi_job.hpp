// svi_job.hpp -- batched stochastic variational inference for gfx950: the
// device code, instantiated per precision by svi_f32.hip and svi_f64.hip.
//
// One wavefront owns one (band, model) job for its whole optimisation: the
// variational state (mu, log sigma, Adam moments) stays in registers across
// all steps, so the step loop costs no launches and no HBM traffic.  The 64
// lanes are split into S' = pow2 >= n_samples sample slots x G = 64/S'
// observation groups: lane (s, g) draws the reparameterised sample
// theta_s = mu + sigma * eps_s and walks observations g, g+G, ... of the
// band's LDS copy, accumulating the log-likelihood (Kahan) and its adjoint
// dLL/dtheta.  The adjoint is formed per observation from dLL/df and the
// model's parameter gradient, recomputed with the forward value (nothing per
// observation is stored).  One reduce-scatter per step (permlane swaps and
// DPP, no LDS) forms sum LL, sum dF/dtheta and sum dF/dtheta * sigma * eps,
// which are exactly the reparameterisation gradients w.r.t. mu and log
// sigma, and leaves each total in the lanes that own that component's Adam
// state; the group lanes of a sample draw its Philox pairs once between
// them.  Per step the wave thus spends its time on observations, not on
// replicated per-job bookkeeping.
//
// Semantics follow the reference's GPU SVI (cuda/svi.cu:48-330): PSO point as
// the initial mean, log sigma = -1, Gaussian prior with the given centres and
// widths, entropy sum(log sigma) + np/2 ln(2 pi e), Adam (0.9, 0.999, 1e-8)
// skipping non-finite components, log sigma clamped to [-6, 2], EMA early stop
// after 50 steps with a 20-step stall.  The noise is Philox4x32-10 keyed by the
// seed with counter (job, step, sample, pair) instead of the reference's
// per-job xorshift64, so every (job, step, sample) draw is independent of how
// lanes and jobs are scheduled.
#pragma once
#include "svi_kernels.hpp"
#include "models_grad.hpp"

namespace lcf {

template <typename R> struct ObsR;
template <> struct ObsR<float> { using T = float4; };
template <> struct ObsR<double> { using T = double4; };

// The FP32 upper-limit term of the exact-tail adjoint, log_normal_cdf(z) and
// dlog_normal_cdf(z) (models_fit.hpp, the reference's A&S form and its own
// derivative), tabulated on [-30, 8] in SVI_TAIL_N intervals of width 1/8:
// interval i holds two float4 of cubic coefficients in u = 8 (z - z_i), one
// per function, through the FP64 values at u = 0, 1/3, 2/3, 1 (built once per
// device by svi_f32_prepare).  z = 0 is a node, so the |z| in the A&S form
// never falls inside an interval.  Below -30 the reference's asymptote is
// evaluated; above 8 the last interval's cubics (below 1e-14) stand for the
// reference's exact 0.  One copy per translation unit: only svi_f32.hip
// (LCF_SVI_F32_UNIT) fills and reads it.
constexpr int SVI_TAIL_N = 304;
constexpr float SVI_TAIL_Z0 = -30.0f;
constexpr float SVI_TAIL_SCALE = 8.0f;
static __device__ float4 g_svi_tail[2 * SVI_TAIL_N];

// phi(z) / max(Phi(z), 1e-300) with the A&S 7.1.26 erfc.  FP64 evaluates the
// reference's expression as written (cuda/svi.cu:200-204); FP32 reproduces
// the value that expression takes in binary64, including its rounding of
// 1 - (1 - tiny) and the 1e-300 floor far in the tail.
template <typename R> __device__ __forceinline__ R mills_ratio(R z);
template <> __device__ __forceinline__ double mills_ratio<double>(double z) {
    double a = fabs(z * 0.70710678118654752440);
    double t = 1.0 / (1.0 + 0.3275911 * a);
    double poly = t * (0.254829592 + t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429))));
    double val = 1.0 - poly * ::exp(-a * a);
    double erfv = z >= 0.0 ? val : -val;
    double phi = ::exp(-0.5 * z * z) * 0.39894228040143267794;
    double cdf = fmax(0.5 * (1.0 + erfv), 1e-300);
    return phi / cdf;
}
template <> __device__ __forceinline__ float mills_ratio<float>(float z) {
    using M = Math<float>;
    float a = fabsf(z) * 0.707106781f;
    float t = M::rcp(1.0f + 0.3275911f * a);
    float poly = t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
    if (z >= 0.0f) {
        float e = M::exp(-0.5f * z * z);
        return 0.398942280f * e * M::rcp(1.0f - 0.5f * poly * e);
    }
    // z < 0: the reference's binary64 cdf is 0.5 (1 - fl(1 - x)), x = poly e^{-a^2},
    // floored at 1e-300.  For x >= 2^-20 that is 0.5 x to FP32 accuracy; below,
    // 1 - x rounds to a multiple k of 2^-53, so cdf = 0.5 k 2^-53 (k = 0: the
    // floor).  The ratio is formed in the log domain; it is 0 where the
    // binary64 phi(z) underflows (z^2 > 1490.27) and +inf where the
    // reference's value exceeds the FP32 range (then the step is skipped).
    float ln_x = M::log(poly) - a * a;
    if (ln_x > -13.8629436f) return 0.797884561f * M::rcp(poly);  // (1/sqrt(2 pi)) / (poly / 2)
    if (z * z > 1490.26644f) return 0.0f;
    float k = rintf(M::exp(ln_x + 36.7368005f));                   // x 2^53
    float ln_phi = -0.5f * z * z - 0.918938533f;
    float ln_cdf = k >= 1.0f ? M::log(0.5f * k) - 36.7368005f : -690.775528f;
    return M::exp(ln_phi - ln_cdf);
}

// phi(z) and max(Phi(z), 1e-300) as the reference forms them for the
// upper-limit adjoint (cuda/svi.cu:200-203, erf_approx of src/parametric.rs:867-880).
template <typename R>
__device__ __forceinline__ void phi_cdf(R z, R& phi, R& cdf) {
    using M = Math<R>;
    phi = M::exp(R(-0.5) * z * z) / M::sqrt(R(2.0 * 3.14159265358979323846));
    R x = z * R(0.70710678118654752440);
    R ax = x < R(0) ? -x : x;
    R t = R(1) / (R(1) + R(0.3275911) * ax);
    R poly = t * (R(0.254829592) + t * (R(-0.284496736) + t * (R(1.421413741) + t * (R(-1.453152027) + t * R(1.061405429)))));
    R val = R(1) - poly * M::exp(-ax * ax);
    R erfv = x >= R(0) ? val : -val;
    R c = R(0.5) * (R(1) + erfv);
    cdf = c > R(1e-300) ? c : R(1e-300);
}

// 2 pi u in revolutions: v_cos_f32 / v_sin_f32 take their argument in units of 2 pi.
__device__ __forceinline__ void sincos_rev(float u, float& s, float& c) {
    s = __builtin_amdgcn_sinf(u);
    c = __builtin_amdgcn_cosf(u);
}
__device__ __forceinline__ void sincos_rev(double u, double& s, double& c) {
    ::sincos(6.28318530717958647692 * u, &s, &c);
}

// Box-Muller of one Philox block: the pair of normals from (x, y).
template <typename R>
__device__ __forceinline__ void box_muller(U4 r, R& e0, R& e1) {
    R u1 = (R((r.x >> 9)) + R(0.5)) * R(1.0 / 8388608.0);
    R u2 = R(r.y >> 8) * R(1.0 / 16777216.0);
    R rad = Math<R>::sqrt(R(-2) * Math<R>::log(u1));
    R s, c;
    sincos_rev(u2, s, c);
    e0 = rad * c;
    e1 = rad * s;
}

// Philox normals of (job, step, sample): eps[2q], eps[2q+1] from pair q.
template <typename R, int NP>
__device__ __forceinline__ void draw_normals(uint64_t seed, uint32_t job, uint32_t step, uint32_t sample,
                                             R (&eps)[NP]) {
#pragma unroll
    for (int q = 0; q < (NP + 1) / 2; q++) {
        R e0, e1;
        box_muller(philox4x32_10(job, step, sample, (uint32_t)q, (uint32_t)seed, (uint32_t)(seed >> 32)), e0, e1);
        eps[2 * q] = e0;
        if (2 * q + 1 < NP) eps[2 * q + 1] = e1;
    }
}

// The same normals with each pair drawn once: lane (s, g) of the G >= pairs
// group lanes s G .. s G + G - 1 draws pair g and the group exchanges them.
template <typename R, int NP>
__device__ __forceinline__ void draw_normals_split(uint64_t seed, uint32_t job, uint32_t step, int s, int grp,
                                                   int G, R (&eps)[NP]) {
    R e0, e1;
    box_muller(philox4x32_10(job, step, (uint32_t)s, (uint32_t)grp, (uint32_t)seed, (uint32_t)(seed >> 32)), e0, e1);
    const int base = s * G;
#pragma unroll
    for (int q = 0; q < (NP + 1) / 2; q++) {
        eps[2 * q] = __shfl(e0, base + q, WAVE);
        if (2 * q + 1 < NP) eps[2 * q + 1] = __shfl(e1, base + q, WAVE);
    }
}

// log_normal_cdf(z) and its derivative from the tail table (any z: the index
// is clamped before the load, so NaN and +-inf read inside the table).
__device__ __forceinline__ float svi_tail_tab(float z, float& d) {
    const float ug = __builtin_fmaf(z, SVI_TAIL_SCALE, -SVI_TAIL_Z0 * SVI_TAIL_SCALE);
    const unsigned i = (unsigned)__builtin_amdgcn_fmed3f(ug, 0.0f, (float)(SVI_TAIL_N - 1));
    const float u = __builtin_amdgcn_fractf(ug);
    const float4 a = g_svi_tail[2 * i], b = g_svi_tail[2 * i + 1];
    d = __builtin_fmaf(u, __builtin_fmaf(u, __builtin_fmaf(u, b.w, b.z), b.y), b.x);
    return __builtin_fmaf(u, __builtin_fmaf(u, __builtin_fmaf(u, a.w, a.z), a.y), a.x);
}
// The reference's asymptote below z = -30: -z^2/2 - ln sqrt(2 pi) - ln(-z)
// and its derivative -z - 1/z.
__device__ __forceinline__ float svi_tail_asym(float z, float& d) {
    d = -z - __builtin_amdgcn_rcpf(z);
    return __builtin_fmaf(-0.5f * z, z, -0.918938533204672742f) -
           0.693147180559945309f * __builtin_amdgcn_logf(-z);
}

// Copies cnt observations into LDS with the measured ones first and the
// upper limits after them, each kind in its original order, and returns the
// number of measured ones.  One pass over a tile hands G consecutive
// observations to the wave (one per group, all samples), so after this
// partition only the pass that straddles the two kinds takes both branches
// of the likelihood (FP64), and the FP32 walk runs the two kinds as two
// loops.
template <typename R>
__device__ __forceinline__ int stage(typename ObsR<R>::T* lds, const typename ObsR<R>::T* __restrict__ g, int cnt) {
    using O = typename ObsR<R>::T;
    const int lane = lane_id();
    __syncthreads();
    int n_meas = 0;
    for (int b = 0; b < cnt; b += WAVE) {
        const int i = b + lane;
        n_meas += __popcll(__ballot(i < cnt && g[i].w == R(0)));
    }
    int km = 0, ku = n_meas;
    for (int b = 0; b < cnt; b += WAVE) {
        const int i = b + lane;
        const bool valid = i < cnt;
        const O v = g[valid ? i : cnt - 1];  // cnt >= 1
        const bool meas = valid && v.w == R(0), up = valid && !meas;
        const unsigned long long bm = __ballot(meas), bu = __ballot(up);
        const int rm = __builtin_amdgcn_mbcnt_hi((uint32_t)(bm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u));
        const int ru = __builtin_amdgcn_mbcnt_hi((uint32_t)(bu >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bu, 0u));
        if (meas) lds[km + rm] = v;
        if (up) lds[ku + ru] = v;
        km += __popcll(bm);
        ku += __popcll(bu);
    }
    __syncthreads();
    return n_meas;
}

template <int MODEL, typename R, bool FD>
struct SviJob {
    static constexpr int NP = ModelTraits<MODEL>::NP;
    static constexpr int SE = NP - 1;
    using M = Math<R>;
    using O = typename ObsR<R>::T;

    // Model value and gradient: analytic adjoint or forward differences with
    // the reference's step h = max(1e-5, 1e-5 |theta_j|).
    struct Eval {
        ModelGrad<MODEL, R> base;
        typename ValueForm<MODEL, R>::T pert[FD ? NP - 1 : 1];
        R h[NP];
        __device__ __forceinline__ void setup(const R (&th)[NP]) {
            base.setup(th);
            if (FD) {
#pragma unroll
                for (int j = 0; j < NP - 1; j++) {
                    R tp[NP];
#pragma unroll
                    for (int k = 0; k < NP; k++) tp[k] = th[k];
                    R a = th[j] < R(0) ? -th[j] : th[j];
                    h[j] = R(1e-5) * (a > R(1) ? a : R(1));
                    tp[j] = th[j] + h[j];
                    pert[j].setup(tp);
                }
            }
        }
        __device__ __forceinline__ bool operator()(R t, R& f, R (&g)[NP]) const {
            if (!FD) return base.pred_grad(t, f, g);
            if (!base.pred(t, f)) return false;
#pragma unroll
            for (int j = 0; j < NP - 1; j++) {
                R fp;
                bool ok = pert[j].pred(t, fp) && M::finite(fp);
                g[j] = ok ? (fp - f) / h[j] : R(0);
            }
            g[NP - 1] = R(0);
            return true;
        }
    };

    // LL of this lane's observations for sample theta and its gradient.
    __device__ __forceinline__ R loglik(const R (&th)[NP], const O* lds, const O* __restrict__ gobs, int n,
                                        int tile, int grp, int G, bool exact_tail, R (&dll)[NP]) {
        constexpr bool F64 = sizeof(R) == 8;
        Eval ev;
        ev.setup(th);
        R se2;
        if (F64) {  // the reference squares exp(ln sigma_x) (cuda/svi.cu:164-165)
            R se = M::exp(th[SE]);
            se2 = se * se;
        } else {
            se2 = M::exp(R(2) * th[SE]);
        }
        Kahan<R> ll;
        ll.init();
#pragma unroll
        for (int j = 0; j < NP; j++) dll[j] = R(0);
        for (int base = 0; base < n; base += tile) {
            int cnt = n - base < tile ? n - base : tile;
            if (n > tile) stage<R>(const_cast<O*>(lds), gobs + base, cnt);
            for (int i = grp; i < cnt; i += G) {
                O o = lds[i];
                R f, g[NP];
                const bool ok = ev(o.x, f, g) && M::finite(f);
                R tv = o.z + se2;
                R term, wgt, dse;  // LL term, dLL/df, dLL/d ln sigma_x
                if (o.w != R(0)) {
                    const R r = o.y - f;
                    R st, rst, z, ratio;
                    if (F64) {
                        st = M::sqrt(tv);
                        z = r / st;
                    } else {
                        rst = M::rsqrt(tv);
                        z = r * rst;
                    }
                    if (exact_tail) {  // the term and its own derivative
                        term = log_normal_cdf_d<R>(z, ratio);
                        wgt = F64 ? -ratio / st : -ratio * rst;
                    } else if (F64) {  // cuda/svi.cu:200-204 as written
                        term = log_normal_cdf<R>(z);
                        R phi, cdf;
                        phi_cdf(z, phi, cdf);
                        ratio = phi / cdf;
                        wgt = -phi / (cdf * st);
                    } else {
                        term = log_normal_cdf<R>(z);
                        ratio = mills_ratio<R>(z);
                        wgt = -ratio * rst;
                    }
                    dse = F64 ? ratio * (-r * se2 / (st * tv)) : -ratio * z * se2 * (rst * rst);
                } else {
                    R inv = M::rcp(tv);
                    R r = o.y - f, r2 = r * r;
                    term = R(-0.5) * (r2 * inv + M::log(R(6.28318530717958647692) * tv));
                    wgt = r * inv;
                    dse = (r2 * inv * inv - inv) * se2;
                }
                // An observation whose model value is not finite is skipped,
                // and so are non-finite gradient components.  One test of the
                // components' sum clears the common case; the rare one zeroes
                // what is skipped, so the accumulators have one update site.
                R gs = g[0];
#pragma unroll
                for (int j = 1; j < NP - 1; j++) gs += g[j];
                if (!ok || !M::finite(gs)) {
#pragma unroll
                    for (int j = 0; j < NP - 1; j++) g[j] = ok && M::finite(g[j]) ? g[j] : R(0);
                    if (!ok) term = wgt = dse = R(0);
                }
                if (F64) ll.s += term; else ll.add(term);
#pragma unroll
                for (int j = 0; j < NP - 1; j++) dll[j] += wgt * g[j];
                dll[SE] += dse;
            }
        }
        return ll.s;
    }

    // FP32 walk of this lane's observations: the measured ones, then the
    // upper limits (two loops over the staged tile, no branch on the kind),
    // with no per-observation finiteness tests.  A non-finite model value or
    // gradient makes the lane's sums non-finite; `fin` reports it and the
    // caller then repeats the step's walk with loglik(), which skips such
    // observations and components as the reference does -- so the result
    // is loglik()'s whenever every term is finite, and loglik()'s exactly
    // otherwise.  The exact-tail upper-limit term and its derivative come
    // from the tail table (the asymptote only when some lane needs it).
    __device__ __forceinline__ R loglik_fast(const R (&th)[NP], const O* lds, const O* __restrict__ gobs, int n,
                                             int nd_staged, int tile, int grp, int G, bool exact_tail,
                                             R (&dll)[NP], bool& fin) {
#ifndef LCF_SVI_F32_UNIT
        static_assert(sizeof(R) != 4, "FP32 SVI is instantiated in svi_f32.hip only (it builds the tail table)");
#endif
        Eval ev;
        ev.setup(th);
        const R se2 = M::exp(R(2) * th[SE]);
        Kahan<R> ll;
        ll.init();
#pragma unroll
        for (int j = 0; j < NP; j++) dll[j] = R(0);
        for (int base = 0; base < n; base += tile) {
            const int cnt = n - base < tile ? n - base : tile;
            const int nd = n > tile ? stage<R>(const_cast<O*>(lds), gobs + base, cnt) : nd_staged;
            for (int i = grp; i < nd; i += G) {
                const O o = lds[i];
                R f, g[NP];
                const bool ok = ev(o.x, f, g);
                f = ok ? f : R(INFINITY);
                const R tv = o.z + se2, inv = M::rcp(tv);
                const R r = o.y - f, ri = r * inv;
                // -(r^2/tv + ln(2 pi tv)) / 2, dLL/df = r/tv, dLL/d ln sigma_x = (r^2/tv^2 - 1/tv) se2
                const R term = __builtin_fmaf(R(-0.5) * r, ri,
                                              __builtin_fmaf(__builtin_amdgcn_logf(tv), R(-0.346573590279972655),
                                                             R(-0.918938533204672742)));
                ll.add(term);
#pragma unroll
                for (int j = 0; j < NP - 1; j++) dll[j] = __builtin_fmaf(ri, g[j], dll[j]);
                dll[SE] = __builtin_fmaf(__builtin_fmaf(ri, ri, -inv), se2, dll[SE]);
            }
            for (int i = nd + grp; i < cnt; i += G) {
                const O o = lds[i];
                R f, g[NP];
                const bool ok = ev(o.x, f, g);
                f = ok ? f : R(INFINITY);
                const R rst = M::rsqrt(o.z + se2);
                const R z = (o.y - f) * rst;
                R term, ratio;
                if (exact_tail) {
                    term = svi_tail_tab(z, ratio);
                    if (__builtin_expect(__any(z < SVI_TAIL_Z0), 0)) {
                        R ra, ta = svi_tail_asym(z, ra);
                        term = z < SVI_TAIL_Z0 ? ta : term;
                        ratio = z < SVI_TAIL_Z0 ? ra : ratio;
                    }
                } else {
                    term = log_normal_cdf<R>(z);
                    ratio = mills_ratio<R>(z);
                }
                ll.add(term);
                const R wgt = -ratio * rst;
#pragma unroll
                for (int j = 0; j < NP - 1; j++) dll[j] = __builtin_fmaf(wgt, g[j], dll[j]);
                dll[SE] = __builtin_fmaf(wgt * z, se2 * rst, dll[SE]);
            }
        }
        bool f_ok = M::finite(ll.s);
#pragma unroll
        for (int j = 0; j < NP; j++) f_ok = f_ok && M::finite(dll[j]);
        fin = f_ok;
        return ll.s;
    }

    // Monte-Carlo ELBO of step `step` at the wave-uniform (mu, sigma =
    // e^ls): returns the ELBO and, in lane l, the wave total of value l >> 2
    // of {sum LL + log prior, sum dL/dtheta_j (j < NP), sum dL/dtheta_j
    // sigma_j eps_j (j < NP)}, the unnormalised reparameterisation gradients
    // w.r.t. mu_j and ls_j.
    __device__ __forceinline__ R elbo_grad(const R (&mu)[NP], const R (&sig)[NP], const R (&ls)[NP],
                                           const R (&pc)[NP], const R (&pw2)[NP], const O* lds,
                                           const O* __restrict__ gobs, int n, int nd_staged, int tile,
                                           uint32_t job, uint32_t step, int S, int G, uint64_t seed,
                                           bool exact_tail, R& red) {
        constexpr bool F64 = sizeof(R) == 8;
        const int lane = lane_id();
        const int s = lane / G, grp = lane - s * G;
        const bool active = s < S;
        R eps[NP], th[NP], dll[NP];
        if (G >= (NP + 1) / 2) draw_normals_split<R, NP>(seed, job, step, s, grp, G, eps);
        else draw_normals<R, NP>(seed, job, step, (uint32_t)s, eps);
#pragma unroll
        for (int j = 0; j < NP; j++) th[j] = mu[j] + sig[j] * eps[j];
        R L;
        if constexpr (F64) {
            L = loglik(th, lds, gobs, n, tile, grp, G, exact_tail, dll);
        } else {
            bool fin;
            L = loglik_fast(th, lds, gobs, n, nd_staged, tile, grp, G, exact_tail, dll, fin);
            if (__builtin_expect(__any(!fin), 0)) L = loglik(th, lds, gobs, n, tile, grp, G, exact_tail, dll);
        }
        if (grp == 0) {
            R lp = R(0);
#pragma unroll
            for (int j = 0; j < NP; j++) {
                R d = th[j] - pc[j];
                if (F64) {  // as the reference forms it (cuda/svi.cu:236-240)
                    lp += R(-0.5) * d * d / pw2[j];
                    dll[j] -= d / pw2[j];
                } else {  // pw2 holds 1 / width^2
                    R q = d * pw2[j];
                    lp += R(-0.5) * d * q;
                    dll[j] -= q;
                }
            }
            L += lp;
        }
        R v[16];
        v[0] = active ? L : R(0);
#pragma unroll
        for (int j = 0; j < NP; j++) {
            v[1 + j] = active ? dll[j] : R(0);
            v[1 + NP + j] = active ? dll[j] * sig[j] * eps[j] : R(0);
        }
#pragma unroll
        for (int k = 1 + 2 * NP; k < 16; k++) v[k] = R(0);
        red = wave_sum16(v);
        const R tot = bcast(red, 0);
        R ent = R(0);
#pragma unroll
        for (int j = 0; j < NP; j++) ent += ls[j];
        // 0.5 np ln(2 pi e), as the reference forms it (cuda/svi.cu:264)
        ent += R(0.5 * NP) * (R)::log(2.0 * 3.14159265358979323846 * 2.71828182845904523536);
        return (sizeof(R) == 8 ? tot / R(S) : tot * (R(1) / R(S))) + ent;
    }

    __device__ void run(const SviArgs& a, O* lds, int job) {
        const int lane = lane_id();
        const O* g = reinterpret_cast<const O*>(a.obs) + a.offsets[job];
        const int n = (int)(a.offsets[job + 1] - a.offsets[job]);
        const int S = a.n_samples;
        int sp = 1;
        while (sp < S) sp <<= 1;
        const int G = WAVE / sp;
        if (n <= 0) {
            if (lane < 7) {
                a.out_mu[(int64_t)job * 7 + lane] = 0.0;
                a.out_ls[(int64_t)job * 7 + lane] = 0.0;
            }
            if (lane == 0) {
                a.out_elbo[job] = -1e99;
                if (a.out_steps) a.out_steps[job] = 0;
            }
            return;
        }
        // Lane roles: the reduction leaves value v in lanes 4v .. 4v + 3, so
        // lanes 4 (1 + j) carry mu_j and its Adam moments and lanes
        // 4 (1 + NP + j) carry ls_j and its moments; the wave reads the
        // current (mu, ls) from them with readlane once per step.
        const int vi = lane >> 2;
        const bool own_mu = vi >= 1 && vi <= NP, own_ls = vi > NP && vi <= 2 * NP;
        const bool writer = (lane & 3) == 0 && (own_mu || own_ls);
        const int pj = own_mu ? vi - 1 : (own_ls ? vi - 1 - NP : 0);
        const int64_t row = (int64_t)job * 7;
        R x = R(0);
        if (own_mu) x = (R)a.mu_in[row + pj];
        if (own_ls) x = a.ls_in ? (R)a.ls_in[row + pj] : R(-1);
        R pc[NP], pw2[NP];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            pc[j] = bcast((R)a.prior_c[row + j], 0);
            R w = (R)a.prior_w[row + j];
            pw2[j] = bcast(sizeof(R) == 8 ? w * w : R(1) / (w * w), 0);
        }
        auto current = [&](R (&mu)[NP], R (&sig)[NP], R (&ls)[NP]) {
            const R sx = own_ls ? M::exp(x) : x;
#pragma unroll
            for (int j = 0; j < NP; j++) {
                mu[j] = bcast(x, 4 * (1 + j));
                ls[j] = bcast(x, 4 * (1 + NP + j));
                sig[j] = bcast(sx, 4 * (1 + NP + j));
            }
        };
        auto zero_unused = [&]() {  // padding components of the 7-wide rows
            if (lane >= NP && lane < 7) {
                a.out_mu[row + lane] = 0.0;
                a.out_ls[row + lane] = 0.0;
            }
        };
        const int nd_staged = n <= a.tile ? stage<R>(lds, g, n) : 0;
        // the noise stream's job word: the job's index, or 0 for every job
        // (the CPU fit seeds a fresh SmallRng(42) per band, src/parametric.rs:2239)
        const uint32_t rng_job = (a.flags & LCF_SVI_SHARED_STREAM_FLAG) ? 0u : (uint32_t)job;
        const bool exact_tail = (a.flags & LCF_SVI_EXACT_TAIL_FLAG) != 0;
        const R rS = R(S);
        const R inv_s = R(1) / rS;
        R red;
        if (a.grad_only) {
            R mu[NP], sig[NP], ls[NP];
            current(mu, sig, ls);
            R e = elbo_grad(mu, sig, ls, pc, pw2, lds, g, n, nd_staged, a.tile, rng_job, (uint32_t)a.step0, S,
                            G, a.seed, exact_tail, red);
            if (writer) {  // + 1 on ls: derivative of the entropy term
                if (own_mu) a.out_mu[row + pj] = (double)(red / rS);
                else a.out_ls[row + pj] = (double)(red / rS) + 1.0;
            }
            zero_unused();
            if (lane == 0) a.out_elbo[job] = (double)e;
            return;
        }
        R m1 = R(0), m2 = R(0);
        const int min_it = (a.flags & LCF_SVI_CPU_STOPPING_FLAG) ? 200 : 50;
        const int stall_it = (a.flags & LCF_SVI_CPU_STOPPING_FLAG) ? 50 : 20;
        const R lr = (R)a.lr;
        R b1t = R(1), b2t = R(1);
        R final_elbo = R(-1e30);
        double ema = -1e300, best = -1e300;
        int stall = 0, step = 0;
        for (; step < a.n_steps;) {
            R mu[NP], sig[NP], ls[NP];
            current(mu, sig, ls);
            R e = elbo_grad(mu, sig, ls, pc, pw2, lds, g, n, nd_staged, a.tile, rng_job, (uint32_t)step, S, G,
                            a.seed, exact_tail, red);
            final_elbo = e;
            step++;
            // Adam (cuda/svi.cu:266-291) on the owning lanes, non-finite
            // components skipped, ls clamped to [-6, 2]
            b1t *= R(0.9);
            b2t *= R(0.999);
            R bc1 = R(1) - b1t, bc2 = R(1) - b2t;
            if (sizeof(R) == 8) {  // cuda/svi.cu:269-270 as written
                bc1 = (R)(1.0 - ::pow(0.9, (double)step));
                bc2 = (R)(1.0 - ::pow(0.999, (double)step));
            }
            if constexpr (sizeof(R) == 8) {  // the reference's divisions
                R gg = -(red / rS + (own_ls ? R(1) : R(0)));
                if (M::finite(gg)) {
                    m1 = R(0.9) * m1 + R(0.1) * gg;
                    m2 = R(0.999) * m2 + R(0.001) * gg * gg;
                    x -= lr * (m1 / bc1) / (M::sqrt(m2 / bc2) + R(1e-8));
                }
            } else {  // FP32: hardware reciprocals in place of the divisions
                R gg = -(red * inv_s + (own_ls ? R(1) : R(0)));
                if (M::finite(gg)) {
                    m1 = R(0.9) * m1 + R(0.1) * gg;
                    m2 = R(0.999) * m2 + R(0.001) * gg * gg;
                    const R mh = m1 * M::rcp(bc1), vh = m2 * M::rcp(bc2);
                    x -= lr * mh * M::rcp(M::sqrt(vh) + R(1e-8));
                }
            }
            if (own_ls) x = x < R(-6) ? R(-6) : (x > R(2) ? R(2) : x);
            // EMA early stop (cuda/svi.cu:297-312), unless all steps are asked for
            if (!(a.flags & LCF_SVI_FIXED_STEPS_FLAG) && step - 1 >= min_it) {
                double fe = (double)final_elbo;
                ema = ema < -1e290 ? fe : 0.1 * fe + 0.9 * ema;
                if (ema - best > 0.01 * fmax(fabs(best), 1e-10)) { best = ema; stall = 0; }
                else if (++stall >= stall_it) break;
            }
        }
        const double infl = (a.flags & LCF_SVI_INFLATE_FLAG) ? 1.3862943611198906 : 0.0;
        if (writer) {
            if (own_mu) a.out_mu[row + pj] = (double)x;
            else a.out_ls[row + pj] = (double)x + infl;
        }
        zero_unused();
        if (lane == 0) {
            a.out_elbo[job] = step > 0 ? (double)final_elbo : -1e300;
            if (a.out_steps) a.out_steps[job] = step;
        }
    }
};

// FP32 kernels are compiled for four waves per SIMD (at most 128 registers;
// the spills this costs sit outside the observation walk), except MetzgerKN,
// whose 200-point grid would spill inside it.
template <int MODEL, typename R>
constexpr int svi_waves_per_simd() { return sizeof(R) == 4 && MODEL != 7 ? 4 : 1; }

template <int MODEL, typename R, bool FD, bool PERSISTENT>
__global__ __launch_bounds__(64, (svi_waves_per_simd<MODEL, R>())) void svi_kernel(SviArgs a) {
    extern __shared__ unsigned char lds_raw[];
    using O = typename ObsR<R>::T;
    O* lds = reinterpret_cast<O*>(lds_raw);
    SviJob<MODEL, R, FD> job;
    if (PERSISTENT) {
        for (;;) {
            const int t = wave_next_task(a.queue);
            if (t >= a.n_tasks) break;
            job.run(a, lds, a.order[t]);
            wave_reconverge();
        }
    } else {
        int t = a.task_base + (int)blockIdx.x;
        if (t < a.n_tasks) job.run(a, lds, a.order[t]);
    }
}

template <int MODEL, typename R>
static const void* svi_ptr(bool fd, bool persistent) {
    if (fd) return persistent ? (const void*)svi_kernel<MODEL, R, true, true> : (const void*)svi_kernel<MODEL, R, true, false>;
    return persistent ? (const void*)svi_kernel<MODEL, R, false, true> : (const void*)svi_kernel<MODEL, R, false, false>;
}

template <typename R>
static const void* svi_ptr_m(int model, bool fd, bool persistent) {
    switch (model) {
        case 0: return svi_ptr<0, R>(fd, persistent);
        case 1: return svi_ptr<1, R>(fd, persistent);
        case 2: return svi_ptr<2, R>(fd, persistent);
        case 3: return svi_ptr<3, R>(fd, persistent);
        case 4: return svi_ptr<4, R>(fd, persistent);
        case 5: return svi_ptr<5, R>(fd, persistent);
        case 6: return svi_ptr<6, R>(fd, persistent);
        case 7: return svi_ptr<7, R>(fd, persistent);
        default: return nullptr;
    }
}

}  // namespace lcf
