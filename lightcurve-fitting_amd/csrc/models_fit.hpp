// models_fit.hpp -- parametric light-curve models as used inside the fitting
// kernels (PSO cost, SVI likelihood), templated on the arithmetic type.
//
// Each model is a small functor: setup() hoists everything that depends only
// on the particle's parameters (exponentials of log-scales, logistic weights),
// pred() evaluates one observation time.  The formulas are the reference's
// (src/parametric.rs:199-510, cuda/models_device.h:60-145) rewritten in a
// form that cannot overflow in FP32:
//   * logistic terms as q = 1/(1+e^-|x|) with the tail factor e^-|x| applied
//     on the side where it belongs;
//   * products of exponentials merged into one exponential of a sum;
//   * softplus as max(x,0) + log1p(e^-|x|).
// Where the reference's direct FP64 expression is non-finite (an exponential
// argument above ~709.78), pred() returns false, reproducing the reference's
// "non-finite prediction => cost 1e99" path (cuda/models.cu:570) exactly; where
// the reference's FP64 expression underflows to an exact limit (a logistic
// weight of exactly 0, softplus of exactly +inf), pred() returns that limit.
#pragma once
#include "lcf_device.hpp"

namespace lcf {

template <int MODEL> struct ModelTraits;
template <> struct ModelTraits<0> { static constexpr int NP = 6; };  // Bazin
template <> struct ModelTraits<1> { static constexpr int NP = 7; };  // Villar
template <> struct ModelTraits<2> { static constexpr int NP = 7; };  // Tde
template <> struct ModelTraits<3> { static constexpr int NP = 5; };  // Arnett
template <> struct ModelTraits<4> { static constexpr int NP = 5; };  // Magnetar
template <> struct ModelTraits<5> { static constexpr int NP = 5; };  // ShockCooling
template <> struct ModelTraits<6> { static constexpr int NP = 6; };  // Afterglow
template <> struct ModelTraits<7> { static constexpr int NP = 5; };  // MetzgerKN
// MultiBazin with K = 1..4 components (internal ids 8..11): 4K + 2 parameters
// [ln A_c, t0_c, ln tau_rise_c, ln tau_fall_c]_c, B, ln sigma_x
// (cuda/models.cu:286-317, src/parametric.rs:905-965).
constexpr int MB_MODEL0 = 8;
template <> struct ModelTraits<8> { static constexpr int NP = 6; };
template <> struct ModelTraits<9> { static constexpr int NP = 10; };
template <> struct ModelTraits<10> { static constexpr int NP = 14; };
template <> struct ModelTraits<11> { static constexpr int NP = 18; };

// logistic(x) = 1/(1+e^-x) without overflow: returns (q, e) with
// e = e^-|x|, q = 1/(1+e); logistic = x>=0 ? q : e*q, 1-logistic = x>=0 ? e*q : q.
template <typename R>
__device__ __forceinline__ void logistic_parts(R x, R& q, R& e) {
    using M = Math<R>;
    R ax = x < R(0) ? -x : x;
    e = M::exp(-ax);
    q = M::rcp(R(1) + e);
}
template <typename R>
__device__ __forceinline__ R logistic(R x) {
    R q, e;
    logistic_parts(x, q, e);
    return x >= R(0) ? q : e * q;
}

template <int MODEL, typename R> struct ModelStable;

// Bazin: [ln A, B, t0, ln tau_rise, ln tau_fall, ln sigma_x]
// f = A e^{-dt/tf} / (1 + e^{-dt/tr}) + B      (src/parametric.rs:200-211)
template <typename R> struct ModelStable<0, R> {
    R la, b, t0, itr, itf;
    __device__ __forceinline__ void setup(const R* p) {
        la = p[0]; b = p[1]; t0 = p[2];
        itr = Math<R>::exp(-p[3]);
        itf = Math<R>::exp(-p[4]);
    }
    __device__ __forceinline__ bool pred(R t, R& f) const {
        R dt = t - t0;
        R x1 = -dt * itf;  // ln of the decay factor
        R x2 = -dt * itr;  // logistic denominator exponent
        // reference: (A * e^{x1}) overflows => non-finite
        if (x1 > R(EXP_OVF64) || la + x1 > R(EXP_OVF64)) return false;
        if (x2 > R(EXP_OVF64)) { f = b; return true; }  // logistic is exactly 0
        f = Math<R>::exp(la + x1 - softplus(x2)) + b;
        return true;
    }
};

// Villar: [ln A, beta, ln gamma, t0, ln tau_rise, ln tau_fall, ln sigma_x]
// f = A s(ph/tr) [(1-w)(1-beta ph) + w (1-beta gamma) e^{(gamma-ph)/tf}],
// w = s(10 (ph - gamma))                        (src/parametric.rs:231-247)
template <typename R> struct ModelStable<1, R> {
    R a, beta, gamma, t0, itr, itf, one_m_bg;
    __device__ __forceinline__ void setup(const R* p) {
        using M = Math<R>;
        a = M::exp(p[0]); beta = p[1]; gamma = M::exp(p[2]); t0 = p[3];
        itr = M::exp(-p[4]);
        itf = M::exp(-p[5]);
        one_m_bg = R(1) - beta * gamma;
    }
    __device__ __forceinline__ bool pred(R t, R& f) const {
        using M = Math<R>;
        R ph = t - t0;
        R y = (gamma - ph) * itf;
        if (y > R(EXP_OVF64)) return false;  // reference e^{y} overflows
        R xr = ph * itr, qr, er;
        logistic_parts(xr, qr, er);
        R sr = xr >= R(0) ? qr : er * qr;
        R z = R(10) * (ph - gamma), qz, ez;
        logistic_parts(z, qz, ez);
        R omw = z >= R(0) ? ez * qz : qz;  // 1 - w
        // w e^{y} = qz e^{y + min(z, 0)}
        R wexp = qz * M::exp(y + (z < R(0) ? z : R(0)));
        f = a * sr * (omw * (R(1) - beta * ph) + one_m_bg * wexp);
        return true;
    }
};

// TDE: [ln A, B, t0, ln tau_rise, ln tau_fall, alpha, ln sigma_x]
// f = A s(ph/tr) (1 + softplus(ph)/tf)^{-alpha} + B   (src/parametric.rs:297-311)
template <typename R> struct ModelStable<2, R> {
    R la, b, t0, itr, itf, alpha;
    __device__ __forceinline__ void setup(const R* p) {
        la = p[0]; b = p[1]; t0 = p[2];
        itr = Math<R>::exp(-p[3]);
        itf = Math<R>::exp(-p[4]);
        alpha = p[5];
    }
    __device__ __forceinline__ bool pred(R t, R& f) const {
        using M = Math<R>;
        R ph = t - t0;
        if (ph > R(EXP_OVF64)) { f = b; return true; }  // reference: w = inf, decay 0
        R sg = logistic(ph * itr);
        R ps = softplus(ph) + R(1e-6);
        R w = R(1) + ps * itf;
        f = M::exp(la - alpha * M::log(w)) * sg + b;
        return true;
    }
};

// Arnett: [ln A, t0, ln tau_m, logit f, ln sigma_x]
// f = A [f e^{-ps/8.8} + (1-f) e^{-ps/111.3}] (1 - e^{-(ps/tm)^2})
//                                               (src/parametric.rs:339-356)
template <typename R> struct ModelStable<3, R> {
    R a, t0, itm, fr;
    __device__ __forceinline__ void setup(const R* p) {
        using M = Math<R>;
        a = M::exp(p[0]); t0 = p[1];
        itm = M::exp(-p[2]);
        fr = logistic(p[3]);
    }
    __device__ __forceinline__ bool pred(R t, R& f) const {
        using M = Math<R>;
        R ph = t - t0;
        if (ph > R(EXP_OVF64)) { f = R(0); return true; }  // reference: heat 0
        R ps = softplus(ph) + R(1e-6);
        R heat = fr * M::exp(-ps * R(1.0 / 8.8)) + (R(1) - fr) * M::exp(-ps * R(1.0 / 111.3));
        R x = ps * itm;
        f = a * heat * one_minus_exp_neg(x * x);
        return true;
    }
};

// Magnetar: [ln A, t0, ln tau_sd, ln tau_diff, ln sigma_x]
// f = A (1 + ps/tsd)^{-2} (1 - e^{-(ps/td)^2})  (src/parametric.rs:387-400)
template <typename R> struct ModelStable<4, R> {
    R a, t0, itsd, itd;
    __device__ __forceinline__ void setup(const R* p) {
        using M = Math<R>;
        a = M::exp(p[0]); t0 = p[1];
        itsd = M::exp(-p[2]);
        itd = M::exp(-p[3]);
    }
    __device__ __forceinline__ bool pred(R t, R& f) const {
        using M = Math<R>;
        R ph = t - t0;
        if (ph > R(EXP_OVF64)) { f = R(0); return true; }  // reference: spindown 0
        R ps = softplus(ph) + R(1e-6);
        R w = R(1) + ps * itsd;
        R x = ps * itd;
        f = a * M::rcp(w * w) * one_minus_exp_neg(x * x);
        return true;
    }
};

// ShockCooling: [ln A, t0, n, ln tau_tr, ln sigma_x]
// f = A s(5 ph) ps^{-n} e^{-(ps/ttr)^2}         (src/parametric.rs:427-440)
template <typename R> struct ModelStable<5, R> {
    R la, t0, n, ittr;
    __device__ __forceinline__ void setup(const R* p) {
        la = p[0]; t0 = p[1]; n = p[2];
        ittr = Math<R>::exp(-p[3]);
    }
    __device__ __forceinline__ bool pred(R t, R& f) const {
        using M = Math<R>;
        R ph = t - t0;
        if (ph > R(EXP_OVF64)) { f = R(0); return true; }  // reference: cutoff 0
        R s5 = logistic(R(5) * ph);
        R ps = softplus(ph) + R(1e-6);
        R r = ps * ittr;
        f = s5 * M::exp(la - n * M::log(ps) - r * r);
        return true;
    }
};

// Afterglow: [ln A, t0, ln t_b, alpha1, alpha2, ln sigma_x]
// f = A (r^{2 a1} + r^{2 a2})^{-1/2}, r = ps/t_b  (src/parametric.rs:468-483)
template <typename R> struct ModelStable<6, R> {
    R la, t0, ltb, a1, a2;
    __device__ __forceinline__ void setup(const R* p) {
        la = p[0]; t0 = p[1]; ltb = p[2]; a1 = p[3]; a2 = p[4];
    }
    __device__ __forceinline__ bool pred(R t, R& f) const {
        using M = Math<R>;
        R ph = t - t0;
        if (ph > R(EXP_OVF64)) {
            // reference: ln r = +inf; u = inf (finite result 0) unless a term
            // is exp(0 * inf) = NaN or both exponents are negative (u = 0).
            if (a1 == R(0) || a2 == R(0) || (a1 < R(0) && a2 < R(0))) return false;
            f = R(0);
            return true;
        }
        R ps = softplus(ph) + R(1e-6);
        R lr = M::log(ps) - ltb;
        R e1 = R(2) * a1 * lr, e2 = R(2) * a2 * lr;
        R mx = e1 > e2 ? e1 : e2;
        if (mx < R(-745.1332191019412)) return false;  // reference: u = 0 -> inf
        R d = e1 > e2 ? e2 - e1 : e1 - e2;             // -|e1 - e2|
        f = M::exp(la - R(0.5) * (mx + log1p_unit<R>(M::exp(d))));
        return true;
    }
};

// The model used by the fitting kernels: the overflow-free form, except in
// FP64 (specializations below).
template <int MODEL, typename R> struct Model : ModelStable<MODEL, R> {};

// ---------------------------------------------------------------------------
// FP64: the reference's expressions evaluated as written (src/parametric.rs:
// 199-510; the same formulas as cuda/models.cu:46-131), with only the
// parameter-only exponentials hoisted into setup().  The FP64 kernels then
// round as the reference does (up to the device math library), and every
// non-finite prediction is the reference's by construction.  These full
// specializations take precedence over the overflow-free forms above, which
// serve FP32.
// ---------------------------------------------------------------------------
template <> struct Model<0, double> {
    double a, b, t0, tr, tf;
    __device__ __forceinline__ void setup(const double* p) {
        a = ::exp(p[0]); b = p[1]; t0 = p[2]; tr = ::exp(p[3]); tf = ::exp(p[4]);
    }
    __device__ __forceinline__ bool pred(double t, double& f) const {
        double dt = t - t0;
        double e_fall = ::exp(-dt / tf);
        double sig = 1.0 / (1.0 + ::exp(-dt / tr));
        f = a * e_fall * sig + b;
        return __builtin_isfinite(f);
    }
};

template <> struct Model<1, double> {
    double a, beta, gamma, t0, tr, tf;
    __device__ __forceinline__ void setup(const double* p) {
        a = ::exp(p[0]); beta = p[1]; gamma = ::exp(p[2]); t0 = p[3]; tr = ::exp(p[4]); tf = ::exp(p[5]);
    }
    __device__ __forceinline__ bool pred(double t, double& f) const {
        double ph = t - t0;
        double sr = 1.0 / (1.0 + ::exp(-ph / tr));
        double w = 1.0 / (1.0 + ::exp(-10.0 * (ph - gamma)));
        double pl = 1.0 - beta * ph;
        double pr = (1.0 - beta * gamma) * ::exp((gamma - ph) / tf);
        f = a * sr * ((1.0 - w) * pl + w * pr);
        return __builtin_isfinite(f);
    }
};

template <> struct Model<2, double> {
    double a, b, t0, tr, tf, alpha;
    __device__ __forceinline__ void setup(const double* p) {
        a = ::exp(p[0]); b = p[1]; t0 = p[2]; tr = ::exp(p[3]); tf = ::exp(p[4]); alpha = p[5];
    }
    __device__ __forceinline__ bool pred(double t, double& f) const {
        double ph = t - t0;
        double sig = 1.0 / (1.0 + ::exp(-ph / tr));
        double ps = ::log(1.0 + ::exp(ph)) + 1e-6;
        double w = 1.0 + ps / tf;
        f = a * sig * ::pow(w, -alpha) + b;
        return __builtin_isfinite(f);
    }
};

template <> struct Model<3, double> {
    double a, t0, tm, fr;
    __device__ __forceinline__ void setup(const double* p) {
        a = ::exp(p[0]); t0 = p[1]; tm = ::exp(p[2]); fr = 1.0 / (1.0 + ::exp(-p[3]));
    }
    __device__ __forceinline__ bool pred(double t, double& f) const {
        double ps = ::log(1.0 + ::exp(t - t0)) + 1e-6;
        double e_ni = ::exp(-ps / 8.8), e_co = ::exp(-ps / 111.3);
        double heat = fr * e_ni + (1.0 - fr) * e_co;
        double x = ps / tm;
        f = a * heat * (1.0 - ::exp(-x * x));
        return __builtin_isfinite(f);
    }
};

template <> struct Model<4, double> {
    double a, t0, tsd, tdf;
    __device__ __forceinline__ void setup(const double* p) {
        a = ::exp(p[0]); t0 = p[1]; tsd = ::exp(p[2]); tdf = ::exp(p[3]);
    }
    __device__ __forceinline__ bool pred(double t, double& f) const {
        double ps = ::log(1.0 + ::exp(t - t0)) + 1e-6;
        double w = 1.0 + ps / tsd;
        double spindown = 1.0 / (w * w);
        double x = ps / tdf;
        f = a * spindown * (1.0 - ::exp(-x * x));
        return __builtin_isfinite(f);
    }
};

template <> struct Model<5, double> {
    double a, t0, n, ttr;
    __device__ __forceinline__ void setup(const double* p) {
        a = ::exp(p[0]); t0 = p[1]; n = p[2]; ttr = ::exp(p[3]);
    }
    __device__ __forceinline__ bool pred(double t, double& f) const {
        double ph = t - t0;
        double sig5 = 1.0 / (1.0 + ::exp(-ph * 5.0));
        double ps = ::log(1.0 + ::exp(ph)) + 1e-6;
        double ratio = ps / ttr;
        f = a * sig5 * ::pow(ps, -n) * ::exp(-ratio * ratio);
        return __builtin_isfinite(f);
    }
};

template <> struct Model<6, double> {
    double a, t0, tb, a1, a2;
    __device__ __forceinline__ void setup(const double* p) {
        a = ::exp(p[0]); t0 = p[1]; tb = ::exp(p[2]); a1 = p[3]; a2 = p[4];
    }
    __device__ __forceinline__ bool pred(double t, double& f) const {
        double ps = ::log(1.0 + ::exp(t - t0)) + 1e-6;
        double ln_r = ::log(ps / tb);
        double u1 = ::exp(2.0 * a1 * ln_r), u2 = ::exp(2.0 * a2 * ln_r);
        f = a * ::pow(u1 + u2, -0.5);
        return __builtin_isfinite(f);
    }
};

// MetzgerKN per-observation ODE solve (cuda/models_device.h:376-455).
// Always integrated in FP64: the CGS energies (~1e50 erg) exceed FP32 range.
__device__ inline double metzger_kn_single(const double* p, double obs_time) {
    const double MSUN = 1.989e33, CCGS = 2.998e10, DAY = 86400.0, PI = 3.14159265358979323846;
    double m_ej = ::pow(10.0, p[0]) * MSUN;
    double v_ej = ::pow(10.0, p[1]) * CCGS;
    double kappa_r = ::pow(10.0, p[2]);
    double phase = obs_time - p[3];
    if (phase <= 0.0) return 0.0;
    double phase_max = phase * 1.05;
    if (phase_max < 0.02) phase_max = 0.02;
    const double lmin = ::log(0.01);
    double lmax = ::log(phase_max);
    const int NG = 100;
    const double xn0 = 0.8, scale = 1e40;
    double e0 = 0.5 * m_ej * v_ej * v_ej;
    double e_th = e0 / scale, e_kin = e0 / scale, v = v_ej;
    double step = (lmax - lmin) / (double)(NG - 1);
    double g0 = ::exp(lmin);
    double r = g0 * DAY * v;
    double l_peak = 0.0, l_at = 0.0, prev_t = g0, prev_l = 0.0;
    double t_day = g0;
    for (int i = 0; i < NG; i++) {
        double t_sec = t_day * DAY;
        double ef = 0.34 * ::pow(t_day, 0.74);
        double eth = 0.36 * (::exp(-0.56 * t_day) + (ef > 1e-10 ? ::log(1.0 + ef) / ef : 1.0));
        double xn = xn0 * ::exp(-t_sec / 900.0);
        double tt = 0.5 - ::atan((t_sec - 1.3) / 0.11) / PI;
        if (tt < 1e-30) tt = 1e-30;
        double l_heat = m_ej * (3.2e14 * xn + 2e18 * eth * ::pow(tt, 1.3)) / scale;
        double kappa = 0.4 * (xn0 - xn) + kappa_r * (1.0 - xn0);
        double t_diff = 3.0 * kappa * m_ej / (4.0 * PI * CCGS * v * t_sec) + r / CCGS;
        double l_rad = (e_th > 0.0 && t_diff > 0.0) ? e_th / t_diff : 0.0;
        if (l_rad > l_peak) l_peak = l_rad;
        if (i > 0 && phase >= prev_t && phase <= t_day)
            l_at = prev_l + (phase - prev_t) / (t_day - prev_t) * (l_rad - prev_l);
        if (i == NG - 1 && phase >= t_day) l_at = l_rad;
        if (i == 0 && phase <= t_day) l_at = l_rad;
        prev_t = t_day;
        prev_l = l_rad;
        double l_pdv = r > 0.0 ? e_th * v / r : 0.0;
        if (i < NG - 1) {
            double tn = ::exp(lmin + step * (double)(i + 1));
            double dt_sec = (tn - t_day) * DAY;
            e_th += (l_heat - l_pdv - l_rad) * dt_sec;
            if (e_th < 0.0) e_th = 0.0;
            e_kin += l_pdv * dt_sec;
            v = ::sqrt(2.0 * e_kin * scale / m_ej);
            if (v > CCGS) v = CCGS;
            r += v * dt_sec;
            t_day = tn;
        }
    }
    if (l_peak <= 0.0 || !__builtin_isfinite(l_peak)) return 0.0;
    return l_at / l_peak;
}

// MetzgerKN: [log10 M_ej, log10 v_ej, log10 kappa, t0, ln sigma_x]
template <typename R> struct ModelStable<7, R> {
    double p[4];
    __device__ __forceinline__ void setup(const R* q) {
        for (int i = 0; i < 4; i++) p[i] = (double)q[i];
    }
    __device__ __forceinline__ bool pred(R t, R& f) const {
        f = (R)metzger_kn_single(p, (double)t);
        return true;
    }
};

// MultiBazin, overflow-free form: each component as the Bazin model above
// without its baseline; a component whose reference value is non-finite makes
// the prediction non-finite, one whose logistic is exactly 0 contributes 0.
template <int K, typename R> struct MultiBazinStable {
    R la[K], t0[K], itr[K], itf[K], b;
    __device__ __forceinline__ void setup(const R* p) {
#pragma unroll
        for (int c = 0; c < K; c++) {
            la[c] = p[4 * c]; t0[c] = p[4 * c + 1];
            itr[c] = Math<R>::exp(-p[4 * c + 2]);
            itf[c] = Math<R>::exp(-p[4 * c + 3]);
        }
        b = p[4 * K];
    }
    __device__ __forceinline__ bool pred(R t, R& f) const {
        R sum = R(0);
#pragma unroll
        for (int c = 0; c < K; c++) {
            R dt = t - t0[c];
            R x1 = -dt * itf[c], x2 = -dt * itr[c];
            if (x1 > R(EXP_OVF64) || la[c] + x1 > R(EXP_OVF64)) return false;
            if (x2 > R(EXP_OVF64)) continue;
            sum += Math<R>::exp(la[c] + x1 - softplus(x2));
        }
        f = sum + b;
        return true;
    }
};
template <typename R> struct ModelStable<8, R> : MultiBazinStable<1, R> {};
template <typename R> struct ModelStable<9, R> : MultiBazinStable<2, R> {};
template <typename R> struct ModelStable<10, R> : MultiBazinStable<3, R> {};
template <typename R> struct ModelStable<11, R> : MultiBazinStable<4, R> {};

// MultiBazin in FP64, the reference's expression as written
// (bazin_component_at / multi_bazin_at, cuda/models.cu:297-313).
template <int K> struct MultiBazinRef {
    double a[K], t0[K], tr[K], tf[K], b;
    __device__ __forceinline__ void setup(const double* p) {
#pragma unroll
        for (int c = 0; c < K; c++) {
            a[c] = ::exp(p[4 * c]); t0[c] = p[4 * c + 1];
            tr[c] = ::exp(p[4 * c + 2]); tf[c] = ::exp(p[4 * c + 3]);
        }
        b = p[4 * K];
    }
    __device__ __forceinline__ bool pred(double t, double& f) const {
        double sum = 0.0;
#pragma unroll
        for (int c = 0; c < K; c++) {
            double dt = t - t0[c];
            sum += a[c] * ::exp(-dt / tf[c]) * (1.0 / (1.0 + ::exp(-(dt / tr[c]))));
        }
        f = sum + b;
        return __builtin_isfinite(f);
    }
};
template <> struct Model<8, double> : MultiBazinRef<1> {};
template <> struct Model<9, double> : MultiBazinRef<2> {};
template <> struct Model<10, double> : MultiBazinRef<3> {};
template <> struct Model<11, double> : MultiBazinRef<4> {};

// ln Phi(x) with the Abramowitz & Stegun 7.1.26 erfc polynomial
// (src/parametric.rs:882-901).
template <typename R>
__device__ __forceinline__ R log_normal_cdf(R x) {
    using M = Math<R>;
    if (x > R(8)) return R(0);
    if (x < R(-30)) return R(-0.5) * x * x - R(0.91893853320467274178) - M::log(-x);
    R z = -x * R(0.70710678118654752440);
    R az = z < R(0) ? -z : z;
    R t = M::rcp(R(1) + R(0.3275911) * az);
    R poly = t * (R(0.254829592) + t * (R(-0.284496736) + t * (R(1.421413741) +
             t * (R(-1.453152027) + t * R(1.061405429)))));
    R erfc_z = poly * M::exp(-z * z);
    R phi = z >= R(0) ? R(0.5) * erfc_z : R(1) - R(0.5) * erfc_z;
    return M::log(phi > R(1e-300) ? phi : R(1e-300));
}

template <>
__device__ __forceinline__ float log_normal_cdf<float>(float x) {
    using M = Math<float>;
    if (x > 8.0f) return 0.0f;
    if (x < -30.0f) return -0.5f * x * x - 0.918938533f - M::log(-x);
    float z = -x * 0.707106781f;
    float az = fabsf(z);
    float t = M::rcp(1.0f + 0.3275911f * az);
    float poly = t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f +
                 t * (-1.453152027f + t * 1.061405429f))));
    // ln(0.5 erfc) computed in the log domain for z >= 0 so tiny tails stay
    // representable in FP32 (the reference floors Phi at 1e-300 in FP64).
    if (z >= 0.0f) return M::log(0.5f * poly) - z * z;
    return M::log(1.0f - 0.5f * poly * M::exp(-z * z));
}

// d/dx log_normal_cdf(x), differentiated from the A&S form above rather than
// from the exact normal CDF, so that it is the derivative of what the
// likelihood sums: Q(t) = t P(t), t = 1 / (1 + p |z|), z = -x / sqrt 2;
// 0 above 8, -x - 1/x on the asymptote below -30.  Free of the
// 1 - (1 - tiny) cancellation of phi / Phi far in the lower tail.
template <typename R>
__device__ __forceinline__ R dlog_normal_cdf(R x) {
    using M = Math<R>;
    const R S2 = R(0.70710678118654752440), p = R(0.3275911);
    if (x > R(8)) return R(0);
    if (x < R(-30)) return -x - M::rcp(x);
    const R z = -x * S2, az = z < R(0) ? -z : z;
    const R t = M::rcp(R(1) + p * az);
    const R q = t * (R(0.254829592) + t * (R(-0.284496736) + t * (R(1.421413741) +
                t * (R(-1.453152027) + t * R(1.061405429)))));
    const R dq = R(0.254829592) + t * (R(2.0 * -0.284496736) + t * (R(3.0 * 1.421413741) +
                 t * (R(4.0 * -1.453152027) + t * R(5.0 * 1.061405429))));
    const R dt = -p * t * t;
    if (z >= R(0)) return -S2 * (dq * dt / q - R(2) * az);
    const R e = M::exp(-az * az);
    const R E = R(0.5) * q * e;
    const R dE = R(0.5) * e * (dq * dt - R(2) * az * q);
    return -S2 * dE / (R(1) - E);
}

// log_normal_cdf(x) and its derivative dlog_normal_cdf(x) from one
// evaluation of the shared terms (t, Q, e^{-z^2}).  FP64 returns exactly the
// two functions above; FP32 forms the quotients with v_rcp_f32.
template <typename R>
__device__ __forceinline__ R log_normal_cdf_d(R x, R& d) {
    d = dlog_normal_cdf<R>(x);
    return log_normal_cdf<R>(x);
}

template <>
__device__ __forceinline__ float log_normal_cdf_d<float>(float x, float& d) {
    using M = Math<float>;
    const float S2 = 0.707106781f, p = 0.3275911f;
    if (x > 8.0f) { d = 0.0f; return 0.0f; }
    if (x < -30.0f) {
        d = -x - M::rcp(x);
        return -0.5f * x * x - 0.918938533f - M::log(-x);
    }
    const float z = -x * S2, az = fabsf(z);
    const float t = M::rcp(1.0f + p * az);
    const float q = t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f +
                    t * (-1.453152027f + t * 1.061405429f))));
    const float dq = 0.254829592f + t * (2.0f * -0.284496736f + t * (3.0f * 1.421413741f +
                     t * (4.0f * -1.453152027f + t * (5.0f * 1.061405429f))));
    const float dt = -p * t * t;
    if (z >= 0.0f) {
        d = -S2 * (dq * dt * M::rcp(q) - 2.0f * az);
        return M::log(0.5f * q) - z * z;
    }
    const float e = M::exp(-z * z);
    const float E = 0.5f * q * e;
    const float ome = 1.0f - E;
    d = -S2 * (0.5f * e * (dq * dt - 2.0f * az * q)) * M::rcp(ome);
    return M::log(ome);
}

}  // namespace lcf
