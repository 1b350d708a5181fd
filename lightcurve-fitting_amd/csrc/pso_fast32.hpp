// pso_fast32.hpp -- FP32 forms of models 0-6 and MultiBazin for the PSO cost loop, written
// for the gfx950 instruction stream rather than for the textbook formula.
//
// The swarm cost is Sum_i [(f(t_i) - y_i)^2 / v_i + ln v_i] (plus the censored
// term for upper limits), evaluated for every (particle, observation) pair;
// it is the whole of the PSO kernels' work.  The forms below compute the same
// predictions as the overflow-free forms of models_fit.hpp with fewer
// instructions:
//   * time is staged in LDS as t log2(e), so every exponential is a bare
//     v_exp_f32 (2^x) and every logarithm a bare v_log_f32 (log2 x): the
//     scale factors live in the per-particle constants set up once per
//     iteration (t0 log2(e), ln 2 / tau, ...);
//   * |x| and -|x| are source modifiers (fabsf), not compare + select;
//   * a logistic 1 / (1 + 2^-x) is one exponential and one reciprocal (2^-x
//     overflows to +inf and the reciprocal gives the exact limit 0);
//   * the reference's non-finite cases ("prediction non-finite => cost
//     1e99", cuda/models.cu:570) and its exact limits (a logistic weight of
//     exactly 0, a phase beyond the FP64 exponential range) are one compare
//     of the phase against a per-particle threshold: pred() returns +inf for
//     a non-finite reference prediction, and the +inf reaches the cost sum,
//     which the caller maps to the invalid-cost sentinel.
// The softplus log(1 + 2^-|x|) is taken without the small-argument series
// where only an absolute error of ~1e-7 in the exponent reaches the
// prediction (Bazin, TDE, Arnett, Magnetar); ShockCooling and Afterglow raise
// softplus(ph) to a power and keep the series.
//
// Likelihood (cost_quad / cost_pair / cost_one): observations are consumed
// four at a time from the detections-first LDS tile; a pair's terms share one
// reciprocal and one logarithm, d_a^2/v_a + d_b^2/v_b = (d_a^2 v_b + d_b^2 v_a)
// / (v_a v_b) and ln v_a + ln v_b = log2(v_a v_b) ln 2 (v >= e^-12 + var, so a
// product of two never leaves the normal FP32 range for normalised variances
// below 1e19), and the four terms enter one Kahan update.  Upper limits follow, four at a time, with the censored term read
// from a table (neg_ln_phi_tab).
#pragma once
#include "lcf_device.hpp"
#include "models_fit.hpp"

namespace lcf {
namespace f32 {

constexpr float L2E = 1.44269504088896341f;     // log2(e)
constexpr float LN2 = 0.693147180559945309f;    // ln 2
constexpr float SQL2E = 1.20112240878644981f;   // sqrt(log2(e))
constexpr float OVF2 = 1024.0f;                 // EXP_OVF64 log2(e): 2^x overflows FP64 above this
constexpr float UNF2 = -1075.0f;                // ln(denorm_min) log2(e)

__device__ __forceinline__ float ex2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float lg2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float ex(float x) { return ex2(x * L2E); }  // e^x, setup only

// log2(1 + 2^-|x|) + max(x, 0) = softplus(x ln 2) / ln 2
__device__ __forceinline__ float softplus2(float x) {
    return __builtin_fmaxf(x, 0.0f) + lg2(1.0f + ex2(-__builtin_fabsf(x)));
}
// The same with the series log2(1 + e) = (e - e^2/2 + e^3/3 - e^4/4) / ln 2
// below e = 2^-7, for forms that take a logarithm of the result.
__device__ __forceinline__ float softplus2_acc(float x) {
    float e = ex2(-__builtin_fabsf(x));
    float s = e * (L2E - e * (0.5f * L2E - e * (L2E / 3.0f - e * (0.25f * L2E))));
    float l = lg2(1.0f + e);
    return __builtin_fmaxf(x, 0.0f) + (e < 0.0078125f ? s : l);
}

template <int MODEL> struct Fast;

// Bazin: f = A e^{-dt/tf} / (1 + e^{-dt/tr}) + B  (src/parametric.rs:200-211)
// as 2^(la - ph/tf - softplus(-ph/tr)) + B; non-finite when A e^{-dt/tf}
// overflows FP64, exactly B when the FP64 logistic underflows to 0.
template <> struct Fast<0> {
    float la, b, t0, itr, itf, p_inf, p_b;
    __device__ __forceinline__ void setup(const float* p) {
        la = p[0] * L2E; b = p[1]; t0 = p[2] * L2E;
        itr = ex(-p[3]); itf = ex(-p[4]);
        p_inf = -(OVF2 - __builtin_fmaxf(la, 0.0f)) * ex(p[4]);
        p_b = -OVF2 * ex(p[3]);
    }
    __device__ __forceinline__ bool safe(float tmin, float) const {
        return tmin - t0 >= __builtin_fmaxf(p_b, p_inf);
    }
    template <bool CHECK = true>
    __device__ __forceinline__ float pred(float tl) const {
        float ph = tl - t0;
        float e1 = __builtin_fmaf(-ph, itf, la);
        float arg = e1 - softplus2(-ph * itr);
        if (CHECK) arg = ph < p_b ? -INFINITY : arg;
        float f = ex2(arg) + b;
        return CHECK && ph < p_inf ? INFINITY : f;
    }
};

// Villar (src/parametric.rs:231-247): A s(ph/tr) [(1-w)(1-beta ph) +
// w (1-beta gamma) e^{(gamma-ph)/tf}], w = s(10(ph-gamma)), with w e^{y}
// merged into one exponential; non-finite when e^{(gamma-ph)/tf} overflows.
// Both logistic weights share one reciprocal: with e_r = 2^{-ph/tr} and
// e_z = 2^{-|z|}, s(ph/tr) w = 1 / ((1 + e_r)(1 + e_z)) times 1 (z >= 0) or
// e_z (z < 0), and s(ph/tr)(1 - w) the same with the factors swapped.  The
// product (1 + e_r)(1 + e_z) is formed as a plain product, so e_r = +inf
// with e_z = 0 gives +inf (weight 0), never inf * 0.
template <> struct Fast<1> {
    float a, abn, zc, t0, itr, itf, yc, aomb, p_inf;  // A folded into the bracket
    __device__ __forceinline__ void setup(const float* p) {
        float beta = p[1], gamma = ex(p[2]);
        a = ex(p[0]); t0 = p[3] * L2E;
        itr = ex(-p[4]); itf = ex(-p[5]);
        float gl = gamma * L2E;
        abn = a * beta * LN2;
        zc = -10.0f * gl;
        yc = gl * itf;
        aomb = a * (1.0f - beta * gamma);
        p_inf = gl - OVF2 * ex(p[5]);
    }
    __device__ __forceinline__ bool safe(float tmin, float) const { return tmin - t0 >= p_inf; }
    template <bool CHECK = true>
    __device__ __forceinline__ float pred(float tl) const {
        float ph = tl - t0;
        float er1 = 1.0f + ex2(-ph * itr);
        float z = __builtin_fmaf(ph, 10.0f, zc);
        float ez = ex2(-__builtin_fabsf(z));
        float q = rcp(er1 * (1.0f + ez));  // s(ph/tr) s(|z|)
        float sel = z >= 0.0f ? ez : 1.0f;   // (1 - w) / s(|z|)
        float y = __builtin_fmaf(-ph, itf, yc);
        float e = ex2(y + __builtin_fminf(z, 0.0f));  // w e^{y} / s(|z|)
        float inner = __builtin_fmaf(sel, __builtin_fmaf(-abn, ph, a), aomb * e);
        float f = q * inner;
        return CHECK && ph < p_inf ? INFINITY : f;
    }
};

// TDE (src/parametric.rs:297-311): A s(ph/tr) (1 + (softplus(ph) + 1e-6)/tf)^-alpha + B,
// exactly B beyond the FP64 exponential range of the phase.
template <> struct Fast<2> {
    float la, b, t0, itr, wa, wb, al;
    __device__ __forceinline__ void setup(const float* p) {
        la = p[0] * L2E; b = p[1]; t0 = p[2] * L2E;
        itr = ex(-p[3]);
        float itf = ex(-p[4]);
        wa = itf * LN2; wb = 1.0f + 1e-6f * itf;
        al = p[5];
    }
    __device__ __forceinline__ bool safe(float, float tmax) const { return tmax - t0 <= OVF2; }
    template <bool CHECK = true>
    __device__ __forceinline__ float pred(float tl) const {
        float ph = tl - t0;
        float sr = rcp(1.0f + ex2(-ph * itr));
        float w = __builtin_fmaf(softplus2(ph), wa, wb);
        float f = __builtin_fmaf(ex2(__builtin_fmaf(-al, lg2(w), la)), sr, b);
        return CHECK && ph > OVF2 ? b : f;
    }
};

// Arnett (src/parametric.rs:339-356): A [f e^{-ps/8.8} + (1-f) e^{-ps/111.3}]
// (1 - e^{-(ps/tm)^2}), ps = softplus(ph) + 1e-6; 0 beyond the FP64 range.
template <> struct Fast<3> {
    float t0, afr, aomfr, xa, xb;  // A folded into the heating weights
    __device__ __forceinline__ void setup(const float* p) {
        float a = ex(p[0]);
        t0 = p[1] * L2E;
        float itm = ex(-p[2]);
        float fr = logistic(p[3]);
        afr = a * fr;
        aomfr = a * (1.0f - fr);
        xa = LN2 * itm * SQL2E; xb = 1e-6f * itm * SQL2E;
    }
    __device__ __forceinline__ bool safe(float, float tmax) const { return tmax - t0 <= OVF2; }
    template <bool CHECK = true>
    __device__ __forceinline__ float pred(float tl) const {
        float ph = tl - t0;
        float sp = softplus2(ph);  // ps = sp ln2 + 1e-6
        float h1 = ex2(__builtin_fmaf(sp, -1.0f / 8.8f, -1e-6f * L2E / 8.8f));
        float h2 = ex2(__builtin_fmaf(sp, -1.0f / 111.3f, -1e-6f * L2E / 111.3f));
        float heat = __builtin_fmaf(afr, h1, aomfr * h2);
        float x = __builtin_fmaf(sp, xa, xb);  // ps / tm sqrt(log2 e)
        float f = heat * (1.0f - ex2(-x * x));
        return CHECK && ph > OVF2 ? 0.0f : f;
    }
};

// Magnetar (src/parametric.rs:387-400): A (1 + ps/tsd)^-2 (1 - e^{-(ps/td)^2}).
template <> struct Fast<4> {
    float t0, wa, wb, xa, xb;  // w scaled by A^-1/2, so A / w^2 = 1 / (w')^2
    __device__ __forceinline__ void setup(const float* p) {
        float isa = ex(-0.5f * p[0]);
        t0 = p[1] * L2E;
        float itsd = ex(-p[2]), itd = ex(-p[3]);
        wa = LN2 * itsd * isa; wb = (1.0f + 1e-6f * itsd) * isa;
        xa = LN2 * itd * SQL2E; xb = 1e-6f * itd * SQL2E;
    }
    __device__ __forceinline__ bool safe(float, float tmax) const { return tmax - t0 <= OVF2; }
    template <bool CHECK = true>
    __device__ __forceinline__ float pred(float tl) const {
        float ph = tl - t0;
        float sp = softplus2(ph);
        float w = __builtin_fmaf(sp, wa, wb);
        float x = __builtin_fmaf(sp, xa, xb);
        float f = rcp(w * w) * (1.0f - ex2(-x * x));
        return CHECK && ph > OVF2 ? 0.0f : f;
    }
};

// ShockCooling (src/parametric.rs:427-440): A s(5 ph) ps^-n e^{-(ps/ttr)^2}.
template <> struct Fast<5> {
    float la, t0, n, rl;
    __device__ __forceinline__ void setup(const float* p) {
        la = p[0] * L2E; t0 = p[1] * L2E; n = p[2];
        rl = ex(-p[3]) * SQL2E;
    }
    __device__ __forceinline__ bool safe(float, float tmax) const { return tmax - t0 <= OVF2; }
    template <bool CHECK = true>
    __device__ __forceinline__ float pred(float tl) const {
        float ph = tl - t0;
        float s5 = rcp(1.0f + ex2(-5.0f * ph));
        float ps = __builtin_fmaf(softplus2_acc(ph), LN2, 1e-6f);
        float r = ps * rl;
        float f = s5 * ex2(__builtin_fmaf(-r, r, __builtin_fmaf(-n, lg2(ps), la)));
        return CHECK && ph > OVF2 ? 0.0f : f;
    }
};

// Afterglow (src/parametric.rs:468-483): A (r^{2 a1} + r^{2 a2})^{-1/2},
// r = ps / t_b, as 2^(la - (max + log2(1 + 2^-|e1-e2|)) / 2); non-finite when
// the FP64 sum underflows to 0, and beyond the FP64 phase range as the
// reference's limit of the two powers (see ModelStable<6>).
template <> struct Fast<6> {
    float la, t0, ltb, a1, a2, f_ovf;
    __device__ __forceinline__ void setup(const float* p) {
        la = p[0] * L2E; t0 = p[1] * L2E; ltb = p[2] * L2E;
        a1 = 2.0f * p[3]; a2 = 2.0f * p[4];
        f_ovf = (p[3] == 0.0f || p[4] == 0.0f || (p[3] < 0.0f && p[4] < 0.0f)) ? INFINITY : 0.0f;
    }
    __device__ __forceinline__ bool safe(float, float) const { return false; }
    template <bool CHECK = true>
    __device__ __forceinline__ float pred(float tl) const {
        float ph = tl - t0;
        float ps = __builtin_fmaf(softplus2_acc(ph), LN2, 1e-6f);
        float lr = lg2(ps) - ltb;
        float e1 = a1 * lr, e2 = a2 * lr;
        float mx = __builtin_fmaxf(e1, e2);
        float d = -__builtin_fabsf(e1 - e2);
        float f = ex2(__builtin_fmaf(-0.5f, mx + lg2(1.0f + ex2(d)), la));
        f = mx < UNF2 ? INFINITY : f;
        return ph > OVF2 ? f_ovf : f;
    }
};

// MultiBazin with K components (internal ids 8..11, src/parametric.rs:905-965):
// B + sum_c of the Bazin form above without its baseline, components added
// in order; a component whose FP64 amplitude term overflows makes the
// prediction non-finite, one whose logistic underflows to 0 adds nothing.
template <int K> struct FastMB {
    float la[K], t0[K], itr[K], itf[K], p_inf[K], p_b[K], b;
    __device__ __forceinline__ void setup(const float* p) {
#pragma unroll
        for (int c = 0; c < K; c++) {
            la[c] = p[4 * c] * L2E;
            t0[c] = p[4 * c + 1] * L2E;
            itr[c] = ex(-p[4 * c + 2]);
            itf[c] = ex(-p[4 * c + 3]);
            p_inf[c] = -(OVF2 - __builtin_fmaxf(la[c], 0.0f)) * ex(p[4 * c + 3]);
            p_b[c] = -OVF2 * ex(p[4 * c + 2]);
        }
        b = p[4 * K];
    }
    __device__ __forceinline__ bool safe(float tmin, float) const {
        bool ok = true;
#pragma unroll
        for (int c = 0; c < K; c++) ok = ok && tmin - t0[c] >= __builtin_fmaxf(p_b[c], p_inf[c]);
        return ok;
    }
    template <bool CHECK = true>
    __device__ __forceinline__ float pred(float tl) const {
        float sum = 0.0f;
        bool bad = false;
#pragma unroll
        for (int c = 0; c < K; c++) {
            float ph = tl - t0[c];
            float arg = __builtin_fmaf(-ph, itf[c], la[c]) - softplus2(-ph * itr[c]);
            if (CHECK) {
                arg = ph < p_b[c] ? -INFINITY : arg;
                bad = bad || ph < p_inf[c];
            }
            sum += ex2(arg);
        }
        float f = sum + b;
        return CHECK && bad ? INFINITY : f;
    }
};
template <> struct Fast<8> : FastMB<1> {};
template <> struct Fast<9> : FastMB<2> {};
template <> struct Fast<10> : FastMB<3> {};
template <> struct Fast<11> : FastMB<4> {};

// Gaussian terms of 4, 2 and 1 detections for one particle: records are
// {t log2 e, y, var, 0}; se2 = sigma_x^2.  CHECK = false drops the phase
// tests of pred() (the caller has shown that no observation of the tile
// reaches them).
template <bool CHECK, int MODEL>
__device__ __forceinline__ float cost_quad(const Fast<MODEL>& m, float se2, const float4& r0,
                                           const float4& r1, const float4& r2, const float4& r3) {
    float v0 = r0.z + se2, v1 = r1.z + se2, v2 = r2.z + se2, v3 = r3.z + se2;
    float d0 = m.template pred<CHECK>(r0.x) - r0.y, d1 = m.template pred<CHECK>(r1.x) - r1.y;
    float d2 = m.template pred<CHECK>(r2.x) - r2.y, d3 = m.template pred<CHECK>(r3.x) - r3.y;
    float v01 = v0 * v1, v23 = v2 * v3;
    float n01 = __builtin_fmaf(d0 * d0, v1, (d1 * d1) * v0);
    float n23 = __builtin_fmaf(d2 * d2, v3, (d3 * d3) * v2);
    float q = __builtin_fmaf(n23, rcp(v23), n01 * rcp(v01));
    float l = lg2(v01) + lg2(v23);
    return __builtin_fmaf(l, LN2, q);
}
template <bool CHECK, int MODEL>
__device__ __forceinline__ float cost_pair(const Fast<MODEL>& m, float se2, const float4& r0,
                                           const float4& r1) {
    float v0 = r0.z + se2, v1 = r1.z + se2;
    float d0 = m.template pred<CHECK>(r0.x) - r0.y, d1 = m.template pred<CHECK>(r1.x) - r1.y;
    float v01 = v0 * v1;
    float q = __builtin_fmaf(d0 * d0, v1, (d1 * d1) * v0) * rcp(v01);
    return __builtin_fmaf(lg2(v01), LN2, q);
}
template <bool CHECK, int MODEL>
__device__ __forceinline__ float cost_one(const Fast<MODEL>& m, float se2, const float4& r) {
    float v = r.z + se2;
    float d = m.template pred<CHECK>(r.x) - r.y;
    return __builtin_fmaf(lg2(v), LN2, d * (d * rcp(v)));
}

// -ln Phi(x) of the reference's log_normal_cdf (A&S 7.1.26 erfc rational,
// src/parametric.rs:882-901) tabulated on [-30, 8] in LNPHI_N intervals of
// width 1/8: entry i holds {c0, c1, c2, c3} of the cubic in u = 8 (x - x_i)
// through the FP64 values at u = 0, 1/3, 2/3, 1 (built once per device by
// build_lnphi_table in pso_f32.hip).  Interpolation error < 1e-7 absolute.
constexpr int LNPHI_N = 304;
constexpr float LNPHI_X0 = -30.0f;
constexpr float LNPHI_SCALE = 8.0f;
// One copy per translation unit (no device linking across units): the
// table is filled only in pso_f32.hip, the one unit that instantiates the
// FP32 swarm kernels (it defines LCF_PSO_F32_UNIT before including this
// header); a unit without it may not reach the FAST cost loop.
static __device__ float4 g_lnphi[LNPHI_N];

// Censored term -ln Phi(x), x = (y_up - f) / sqrt(v), of an upper limit: the
// table on [-30, 8] and the reference's asymptote x^2/2 + ln sqrt(2 pi) +
// ln(-x) below -30.  The loops work in the table coordinate u = 8 (x + 30):
// upper-limit records are staged with var / 64 and the particle's sigma_x^2
// enters as sigma_x^2 / 64 (both exact), so rsqrt(v / 64) = 8 / sqrt(v) and
// u = (y_up - f) rsqrt(v / 64) + 240 is one fused multiply-add.  The table
// index is clamped before the load, so any u (NaN, +-inf) reads inside the
// table; a +inf prediction (u = -inf) gives +inf.
constexpr float LNPHI_U0 = -LNPHI_X0 * LNPHI_SCALE;  // u at x = 0
template <bool CHECK, int MODEL>
__device__ __forceinline__ float upper_u(const Fast<MODEL>& m, float se2u, const float4& r) {
    return __builtin_fmaf(r.y - m.template pred<CHECK>(r.x), __builtin_amdgcn_rsqf(r.z + se2u), LNPHI_U0);
}
// Table part: interval index clamped to the table (v_med3), position in it
// by v_fract; beyond x = 8 the last interval's cubic, < 1e-15 on all of
// [0, 1], stands for the reference's exact 0.
__device__ __forceinline__ float neg_ln_phi_tab(float ug) {
    unsigned i = (unsigned)__builtin_amdgcn_fmed3f(ug, 0.0f, (float)(LNPHI_N - 1));
    float u = __builtin_amdgcn_fractf(ug);
    float4 c = g_lnphi[i];
    return __builtin_fmaf(u, __builtin_fmaf(u, __builtin_fmaf(u, c.w, c.z), c.y), c.x);
}
// The asymptote at x = u / 8 - 30 (u < 0).
__device__ __forceinline__ float neg_ln_phi_asym(float ug) {
    float x = __builtin_fmaf(ug, 1.0f / LNPHI_SCALE, LNPHI_X0);
    return __builtin_fmaf(LN2, lg2(-x), __builtin_fmaf(0.5f * x, x, 0.918938533204672742f));
}
__device__ __forceinline__ float neg_ln_phi_u(float ug) {
    return ug < 0.0f ? neg_ln_phi_asym(ug) : neg_ln_phi_tab(ug);
}
// Four upper limits: the asymptote is evaluated only when some lane has an
// observation below x = -30 (a wave-uniform branch).
template <bool CHECK, int MODEL>
__device__ __forceinline__ float cost_upper_quad(const Fast<MODEL>& m, float se2u, const float4& r0,
                                                 const float4& r1, const float4& r2, const float4& r3) {
    float u0 = upper_u<CHECK>(m, se2u, r0), u1 = upper_u<CHECK>(m, se2u, r1);
    float u2 = upper_u<CHECK>(m, se2u, r2), u3 = upper_u<CHECK>(m, se2u, r3);
    float c0 = neg_ln_phi_tab(u0), c1 = neg_ln_phi_tab(u1);
    float c2 = neg_ln_phi_tab(u2), c3 = neg_ln_phi_tab(u3);
    float um = __builtin_fminf(__builtin_fminf(u0, u1), __builtin_fminf(u2, u3));
    if (__builtin_expect(__any(um < 0.0f), 0)) {
        c0 = u0 < 0.0f ? neg_ln_phi_asym(u0) : c0;
        c1 = u1 < 0.0f ? neg_ln_phi_asym(u1) : c1;
        c2 = u2 < 0.0f ? neg_ln_phi_asym(u2) : c2;
        c3 = u3 < 0.0f ? neg_ln_phi_asym(u3) : c3;
    }
    return (c0 + c1) + (c2 + c3);
}
template <bool CHECK, int MODEL>
__device__ __forceinline__ float cost_upper(const Fast<MODEL>& m, float se2u, const float4& r) {
    return neg_ln_phi_u(upper_u<CHECK>(m, se2u, r));
}

// Stage `cnt` packed records {t, y, var, flag} from global memory into LDS as
// {t log2 e, y, var, flag} (var / 64 for an upper limit, see upper_u),
// detections first and upper limits after them,
// each group in its original order.  Returns the number of detections and
// the range [tmin, tmax] of t log2 e (wave-uniform).  Single-wave workgroups.
__device__ __forceinline__ int stage_split(float4* lds, const float4* __restrict__ g, int cnt,
                                           float& tmin, float& tmax) {
    const int lane = lane_id();
    __syncthreads();
    int nd = 0;
    float lo = INFINITY, hi = -INFINITY;
    for (int i0 = 0; i0 < cnt; i0 += WAVE) {
        int i = i0 + lane;
        bool valid = i < cnt;
        float4 r = valid ? g[i] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        if (valid) {
            lo = __builtin_fminf(lo, r.x);
            hi = __builtin_fmaxf(hi, r.x);
        }
        nd += __popcll(__ballot(valid && r.w == 0.0f));
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        lo = __builtin_fminf(lo, __shfl_xor(lo, off, WAVE));
        hi = __builtin_fmaxf(hi, __shfl_xor(hi, off, WAVE));
    }
    tmin = lo * L2E;
    tmax = hi * L2E;
    int kd = 0, ku = nd;
    for (int i0 = 0; i0 < cnt; i0 += WAVE) {
        int i = i0 + lane;
        bool valid = i < cnt;
        float4 r = valid ? g[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        bool det = valid && r.w == 0.0f, up = valid && r.w != 0.0f;
        unsigned long long bd = __ballot(det), bu = __ballot(up);
        r.x *= L2E;
        if (up) r.z *= 1.0f / 64.0f;  // upper limits: var / 64 (the table coordinate's scale)
        int rd = __builtin_amdgcn_mbcnt_hi((uint32_t)(bd >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bd, 0u));
        int ru = __builtin_amdgcn_mbcnt_hi((uint32_t)(bu >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bu, 0u));
        if (det) lds[kd + rd] = r;
        if (up) lds[ku + ru] = r;
        kd += __popcll(bd);
        ku += __popcll(bu);
    }
    __syncthreads();
    return nd;
}

}  // namespace f32
}  // namespace lcf
