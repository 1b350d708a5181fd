// models_grad.hpp -- value and parameter gradient of every model in one pass.
//
// pred_grad(t, f, g) returns the flux f and g[j] = df/dp_j for the internal
// parameters (g[sigma_x] = 0), i.e. the reference's analytic gradients
// (src/parametric.rs:213-510, cuda/models_device.h:152-347), in the same
// overflow-free form as models_fit.hpp: every product of an exponential with
// a logistic weight is evaluated as one exponential of a sum, and logistic
// pairs s(x), 1-s(x), s(x)(1-s(x)) come from a single e^{-|x|}.  The
// likelihood adjoint multiplies these by dLL/df, so the backward pass needs
// no per-observation storage: it is recomputed with the forward value.
// Returns false where the reference's FP64 value is non-finite.  In the
// regimes where the reference's own gradient entries are NaN (softplus of a
// phase above ~709.78 overflowing), the reference skips those entries, so
// they are returned as 0 here.
#pragma once
#include "models_fit.hpp"

namespace lcf {

// logistic(x) with its complement and product from one exponential.
template <typename R>
struct Logi {
    R s, oms, s1;  // s(x), 1 - s(x), s(x)(1 - s(x))
    R e, q;        // e^{-|x|}, 1/(1 + e^{-|x|})
    __device__ __forceinline__ void set(R x) {
        logistic_parts(x, q, e);
        R eq = e * q;
        s = x >= R(0) ? q : eq;
        oms = x >= R(0) ? eq : q;
        s1 = eq * q;
    }
};

// softplus(ph) + 1e-6 and its derivative s(ph), sharing one exponential.
template <typename R>
__device__ __forceinline__ void softplus_d(R ph, R& ps, R& sig_p) {
    Logi<R> l;
    l.set(ph);
    ps = (ph > R(0) ? ph : R(0)) + log1p_unit<R>(l.e) + R(1e-6);
    sig_p = l.s;
}

template <int MODEL, typename R> struct ModelGrad;

template <typename R> struct ModelGrad<0, R> : ModelStable<0, R> {
    using B = ModelStable<0, R>;
    __device__ __forceinline__ bool pred_grad(R t, R& f, R* g) const {
        R dt = t - B::t0;
        R x1 = -dt * B::itf, x2 = -dt * B::itr;
        if (x1 > R(EXP_OVF64) || B::la + x1 > R(EXP_OVF64)) return false;
        g[1] = R(1); g[5] = R(0);
        if (x2 > R(EXP_OVF64)) { f = B::b; g[0] = g[2] = g[3] = g[4] = R(0); return true; }
        Logi<R> l;
        l.set(-x2);  // s(dt/tr)
        R sp = (x2 > R(0) ? x2 : R(0)) + log1p_unit<R>(l.e);
        R base = Math<R>::exp(B::la + x1 - sp);
        f = base + B::b;
        g[0] = base;
        g[2] = base * (B::itf - l.oms * B::itr);
        g[3] = -base * l.oms * dt * B::itr;
        g[4] = base * dt * B::itf;
        return true;
    }
};

template <typename R> struct ModelGrad<1, R> : ModelStable<1, R> {
    using B = ModelStable<1, R>;
    __device__ __forceinline__ bool pred_grad(R t, R& f, R* g) const {
        R ph = t - B::t0;
        R y = (B::gamma - ph) * B::itf;
        if (y > R(EXP_OVF64)) return false;
        Logi<R> lr, lz;
        lr.set(ph * B::itr);
        R z = R(10) * (ph - B::gamma);
        lz.set(z);
        R wexp = lz.q * Math<R>::exp(y + (z < R(0) ? z : R(0)));  // w e^{y}
        R pl = R(1) - B::beta * ph;
        R piece = lz.oms * pl + B::one_m_bg * wexp;
        R as = B::a * lr.s;
        f = as * piece;
        R kterm = R(10) * (lz.s1 * pl - lz.oms * B::one_m_bg * wexp);  // k[w(1-w) pl - (1-w)(1-bg) w e^y]
        g[0] = f;
        g[1] = as * (-lz.oms * ph - B::gamma * wexp);
        g[2] = as * B::gamma * (kterm + wexp * (-B::beta + B::one_m_bg * B::itf));
        g[3] = B::a * (-lr.s1 * B::itr * piece + lr.s * (kterm + lz.oms * B::beta + B::one_m_bg * wexp * B::itf));
        g[4] = -B::a * piece * lr.s1 * ph * B::itr;
        g[5] = as * B::one_m_bg * wexp * (ph - B::gamma) * B::itf;
        g[6] = R(0);
        return true;
    }
};

// Villar in FP32 for the SVI walk (src/parametric.rs:231-247 and its
// gradient), with one reciprocal for both logistic weights and the
// exponentials as bare exp2 of pre-scaled arguments.  With x_r = ph/tr and
// z = 10 (ph - gamma) in log2 units, e_r = 2^-|x_r|, e_z = 2^-|z|,
// q = 1 / ((1 + e_r)(1 + e_z)) and N / M the numerators of a logistic and its
// complement (N = 1, M = e at a non-negative argument, N = e, M = 1 below):
//   s_r (1 - w) = N_r M_z q,  s_r w e^y = N_r q E_z,  E_z = 2^{y + min(z, 0)},
//   s_r (1 - s_r) = N_r M_r q^2 (1 + e_z)^2,
// so with A = a N_r q and P = M_z pl + c1 E_z (pl = 1 - beta ph,
// c1 = 1 - beta gamma): f = A P, and every component of the gradient is A
// times a short expression (derivation beside each line).  e_r, e_z <= 1, so
// (1 + e_r)(1 + e_z) lies in [1, 4].  Not finite where the reference's FP64
// e^{(gamma - ph)/tf} overflows (y log2 e > 1024): returns false, as the
// generic form does.
struct VillarGrad32 {
    float a, beta, gamma, t0, itr, itr2, itf2, zc, yc, c1, c1itf, c2;
    static constexpr float L2E = 1.44269504088896341f, K10 = 14.4269504088896341f, OVF2 = 1024.0f;
    __device__ __forceinline__ void setup(const float* p) {
        using M = Math<float>;
        a = M::exp(p[0]); beta = p[1]; gamma = M::exp(p[2]); t0 = p[3];
        itr = M::exp(-p[4]);
        const float itf = M::exp(-p[5]);
        itr2 = itr * L2E; itf2 = itf * L2E;
        zc = -K10 * gamma; yc = gamma * itf2;
        c1 = 1.0f - beta * gamma; c1itf = c1 * itf; c2 = c1itf - beta;
    }
    __device__ __forceinline__ bool pred(float t, float& f) const {
        const float ph = t - t0;
        const float er = __builtin_amdgcn_exp2f(-__builtin_fabsf(ph * itr2));
        const float z = __builtin_fmaf(ph, K10, zc);
        const float ez = __builtin_amdgcn_exp2f(-__builtin_fabsf(z));
        const float y = __builtin_fmaf(-ph, itf2, yc);
        const float ez_y = __builtin_amdgcn_exp2f(y + __builtin_fminf(z, 0.0f));
        const float q = __builtin_amdgcn_rcpf((1.0f + er) * (1.0f + ez));
        const float nr = ph >= 0.0f ? 1.0f : er, mz = z >= 0.0f ? ez : 1.0f;
        f = (a * q) * nr * __builtin_fmaf(mz, __builtin_fmaf(-beta, ph, 1.0f), c1 * ez_y);
        return !(y > OVF2);
    }
    __device__ __forceinline__ bool pred_grad(float t, float& f, float* g) const {
        const float ph = t - t0;
        const float er = __builtin_amdgcn_exp2f(-__builtin_fabsf(ph * itr2));
        const float z = __builtin_fmaf(ph, K10, zc);
        const float ez = __builtin_amdgcn_exp2f(-__builtin_fabsf(z));
        const float y = __builtin_fmaf(-ph, itf2, yc);
        const float ez_y = __builtin_amdgcn_exp2f(y + __builtin_fminf(z, 0.0f));
        const float er1 = 1.0f + er, ez1 = 1.0f + ez;
        const float q = __builtin_amdgcn_rcpf(er1 * ez1);
        const bool pr = ph >= 0.0f, pz = z >= 0.0f;
        const float nr = pr ? 1.0f : er, mr = pr ? er : 1.0f;
        const float nz = pz ? 1.0f : ez, mz = pz ? ez : 1.0f;
        const float pl = __builtin_fmaf(-beta, ph, 1.0f);
        const float c1e = c1 * ez_y;
        const float P = __builtin_fmaf(mz, pl, c1e);
        const float A = (a * q) * nr;
        f = A * P;
        // a s_r kterm = A K: kterm = 10 (1-w) (w pl - c1 w e^y)
        const float K = (10.0f * q) * er1 * (mz * __builtin_fmaf(nz, pl, -c1e));
        // a s_r (1 - s_r) piece / tr = A B
        const float B = itr * ((mr * ez1) * q * P);
        g[0] = f;                                                             // d/d ln A
        g[1] = -A * __builtin_fmaf(mz, ph, gamma * ez_y);                     // d/d beta
        g[2] = (gamma * A) * __builtin_fmaf(c2, ez_y, K);                     // d/d ln gamma
        g[3] = A * ((K - B) + __builtin_fmaf(beta, mz, c1itf * ez_y));        // d/d t0
        g[4] = -ph * (A * B);                                                 // d/d ln tau_rise
        g[5] = (c1itf * ez_y) * A * (ph - gamma);                             // d/d ln tau_fall
        g[6] = 0.0f;
        return !(y > OVF2);
    }
};
template <> struct ModelGrad<1, float> : VillarGrad32 {};

// The value-only form the numerical-gradient SVI evaluates at each perturbed
// parameter vector: the kernels' model form, and for Villar in FP32 the
// value of the form above (same operations as its gradient uses).
template <int MODEL, typename R> struct ValueForm { using T = Model<MODEL, R>; };
template <> struct ValueForm<1, float> { using T = VillarGrad32; };

template <typename R> struct ModelGrad<2, R> : ModelStable<2, R> {
    using B = ModelStable<2, R>;
    __device__ __forceinline__ bool pred_grad(R t, R& f, R* g) const {
        using M = Math<R>;
        R ph = t - B::t0;
        g[1] = R(1); g[6] = R(0);
        if (ph > R(EXP_OVF64)) { f = B::b; g[0] = g[2] = g[3] = g[4] = g[5] = R(0); return true; }
        Logi<R> l;
        l.set(ph * B::itr);
        R ps, sp;
        softplus_d(ph, ps, sp);
        R w = R(1) + ps * B::itf;
        R lw = M::log(w), iw = M::rcp(w);
        R D = M::exp(B::la - B::alpha * lw);
        R Ds = D * l.s;
        f = Ds + B::b;
        g[0] = Ds;
        g[2] = D * (-l.s1 * B::itr + l.s * B::alpha * sp * B::itf * iw);
        g[3] = -D * l.s1 * ph * B::itr;
        g[4] = Ds * B::alpha * ps * B::itf * iw;
        g[5] = -Ds * lw;
        return true;
    }
};

template <typename R> struct ModelGrad<3, R> : ModelStable<3, R> {
    using B = ModelStable<3, R>;
    __device__ __forceinline__ bool pred_grad(R t, R& f, R* g) const {
        using M = Math<R>;
        R ph = t - B::t0;
        g[4] = R(0);
        if (ph > R(EXP_OVF64)) { f = R(0); g[0] = g[1] = g[2] = g[3] = R(0); return true; }
        R ps, sp;
        softplus_d(ph, ps, sp);
        R eni = M::exp(-ps * R(1.0 / 8.8)), eco = M::exp(-ps * R(1.0 / 111.3));
        R heat = B::fr * eni + (R(1) - B::fr) * eco;
        R x = ps * B::itm, x2 = x * x;
        R ex2 = M::exp(-x2);
        R trap = one_minus_exp_neg(x2);
        f = B::a * heat * trap;
        R dheat = -B::fr * eni * R(1.0 / 8.8) - (R(1) - B::fr) * eco * R(1.0 / 111.3);
        R dtrap = R(2) * ps * ex2 * B::itm * B::itm;
        g[0] = f;
        g[1] = -B::a * sp * (dheat * trap + heat * dtrap);
        g[2] = R(-2) * B::a * heat * ex2 * x2;
        g[3] = B::a * trap * (eni - eco) * B::fr * (R(1) - B::fr);
        return true;
    }
};

template <typename R> struct ModelGrad<4, R> : ModelStable<4, R> {
    using B = ModelStable<4, R>;
    __device__ __forceinline__ bool pred_grad(R t, R& f, R* g) const {
        using M = Math<R>;
        R ph = t - B::t0;
        g[4] = R(0);
        if (ph > R(EXP_OVF64)) { f = R(0); g[0] = g[1] = g[2] = g[3] = R(0); return true; }
        R ps, sp;
        softplus_d(ph, ps, sp);
        R w = R(1) + ps * B::itsd, iw = M::rcp(w);
        R sd = iw * iw, iw3 = sd * iw;
        R x = ps * B::itd, x2 = x * x;
        R ex2 = M::exp(-x2);
        R trap = one_minus_exp_neg(x2);
        f = B::a * sd * trap;
        g[0] = f;
        g[1] = -B::a * sp * (R(-2) * iw3 * B::itsd * trap + sd * R(2) * ps * ex2 * B::itd * B::itd);
        g[2] = B::a * trap * R(2) * ps * iw3 * B::itsd;
        g[3] = R(-2) * B::a * sd * ex2 * x2;
        return true;
    }
};

template <typename R> struct ModelGrad<5, R> : ModelStable<5, R> {
    using B = ModelStable<5, R>;
    __device__ __forceinline__ bool pred_grad(R t, R& f, R* g) const {
        using M = Math<R>;
        R ph = t - B::t0;
        g[4] = R(0);
        if (ph > R(EXP_OVF64)) { f = R(0); g[0] = g[1] = g[2] = g[3] = R(0); return true; }
        Logi<R> l5;
        l5.set(R(5) * ph);
        R ps, sp;
        softplus_d(ph, ps, sp);
        R lps = M::log(ps), r = ps * B::ittr;
        R E = M::exp(B::la - B::n * lps - r * r);
        f = l5.s * E;
        g[0] = f;
        g[1] = E * (R(-5) * l5.s1 + l5.s * sp * (B::n * M::rcp(ps) + R(2) * ps * B::ittr * B::ittr));
        g[2] = -f * lps;
        g[3] = R(2) * f * r * r;
        return true;
    }
};

template <typename R> struct ModelGrad<6, R> : ModelStable<6, R> {
    using B = ModelStable<6, R>;
    __device__ __forceinline__ bool pred_grad(R t, R& f, R* g) const {
        using M = Math<R>;
        R ph = t - B::t0;
        g[5] = R(0);
        if (ph > R(EXP_OVF64)) {
            g[0] = g[1] = g[2] = g[3] = g[4] = R(0);
            return B::pred(t, f);
        }
        R ps, sp;
        softplus_d(ph, ps, sp);
        R lr = M::log(ps) - B::ltb;
        R e1 = R(2) * B::a1 * lr, e2 = R(2) * B::a2 * lr;
        R mx = e1 > e2 ? e1 : e2;
        if (mx < R(-745.1332191019412)) return false;
        R ed = M::exp(e1 > e2 ? e2 - e1 : e1 - e2);
        R ql = M::rcp(R(1) + ed);
        R wb = ql, ws = ed * ql;  // weights of the larger / smaller term
        R w1 = e1 >= e2 ? wb : ws, w2 = e1 >= e2 ? ws : wb;
        f = M::exp(B::la - R(0.5) * (mx + log1p_unit<R>(ed)));
        R S = B::a1 * w1 + B::a2 * w2;
        g[0] = f;
        g[1] = f * S * sp * M::rcp(ps);
        g[2] = f * S;
        g[3] = -f * lr * w1;
        g[4] = -f * lr * w2;
        return true;
    }
};

// ---------------------------------------------------------------------------
// FP64: the reference's value and gradient expressions as written
// (src/parametric.rs:213-510, identical to cuda/models_device.h:152-347), on
// top of the reference-form Model<M, double>.  Entries the reference computes
// as inf/NaN come out inf/NaN here too, and the likelihood adjoint skips them
// one by one exactly as the reference does (cuda/svi.cu:209-227).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ref_softplus(double x) { return ::log(1.0 + ::exp(x)) + 1e-6; }
__device__ __forceinline__ double ref_softplus_grad(double x) { return ::exp(x) / (1.0 + ::exp(x)); }

template <> struct ModelGrad<0, double> : Model<0, double> {
    __device__ __forceinline__ bool pred_grad(double t, double& f, double* g) const {
        if (!pred(t, f)) return false;
        double dt = t - t0;
        double ef = ::exp(-dt / tf);
        double sig = 1.0 / (1.0 + ::exp(-dt / tr));
        double base = a * ef * sig;
        g[0] = base; g[1] = 1.0;
        g[2] = base * (1.0 / tf - (1.0 - sig) / tr);
        g[3] = -base * (1.0 - sig) * dt / tr;
        g[4] = base * dt / tf;
        g[5] = 0.0;
        return true;
    }
};

template <> struct ModelGrad<1, double> : Model<1, double> {
    __device__ __forceinline__ bool pred_grad(double t, double& f, double* g) const {
        if (!pred(t, f)) return false;
        const double k = 10.0;
        double ph = t - t0;
        double sr = 1.0 / (1.0 + ::exp(-ph / tr));
        double w = 1.0 / (1.0 + ::exp(-k * (ph - gamma)));
        double pl = 1.0 - beta * ph;
        double ed = ::exp((gamma - ph) / tf);
        double pr = (1.0 - beta * gamma) * ed;
        double piece = (1.0 - w) * pl + w * pr;
        g[0] = a * sr * piece;
        g[1] = a * sr * ((1.0 - w) * (-ph) + w * (-gamma * ed));
        double dw_dlg = -k * w * (1.0 - w) * gamma;
        double dpr_dlg = ed * (-beta + (1.0 - beta * gamma) / tf) * gamma;
        g[2] = a * sr * (dw_dlg * (pr - pl) + w * dpr_dlg);
        double dsig_dt0 = -sr * (1.0 - sr) / tr;
        double dw_dt0 = -k * w * (1.0 - w);
        double dpiece_dt0 = dw_dt0 * (pr - pl) + (1.0 - w) * beta + w * (1.0 - beta * gamma) * ed / tf;
        g[3] = a * (dsig_dt0 * piece + sr * dpiece_dt0);
        g[4] = a * piece * sr * (1.0 - sr) * (-ph / tr);
        g[5] = a * sr * w * (pr * (ph - gamma) / tf);
        g[6] = 0.0;
        return true;
    }
};

template <> struct ModelGrad<2, double> : Model<2, double> {
    __device__ __forceinline__ bool pred_grad(double t, double& f, double* g) const {
        if (!pred(t, f)) return false;
        double ph = t - t0;
        double sig = 1.0 / (1.0 + ::exp(-ph / tr));
        double ps = ref_softplus(ph), sp = ref_softplus_grad(ph);
        double w = 1.0 + ps / tf;
        double decay = ::pow(w, -alpha);
        g[0] = a * sig * decay; g[1] = 1.0;
        g[2] = a * ((-sig * (1.0 - sig) / tr) * decay + sig * (alpha * ::pow(w, -alpha - 1.0) * sp / tf));
        g[3] = a * decay * (-sig * (1.0 - sig) * ph / tr);
        g[4] = a * sig * alpha * ::pow(w, -alpha - 1.0) * ps / tf;
        g[5] = a * sig * (-::log(w)) * decay;
        g[6] = 0.0;
        return true;
    }
};

template <> struct ModelGrad<3, double> : Model<3, double> {
    __device__ __forceinline__ bool pred_grad(double t, double& f, double* g) const {
        if (!pred(t, f)) return false;
        double ph = t - t0;
        double ps = ref_softplus(ph), sp = ref_softplus_grad(ph);
        double e_ni = ::exp(-ps / 8.8), e_co = ::exp(-ps / 111.3);
        double heat = fr * e_ni + (1.0 - fr) * e_co;
        double x = ps / tm, ex2 = ::exp(-x * x), trap = 1.0 - ex2;
        g[0] = a * heat * trap;
        double dh = -fr * e_ni / 8.8 - (1.0 - fr) * e_co / 111.3;
        double dtr = 2.0 * ps * ex2 / (tm * tm);
        g[1] = a * (-sp) * (dh * trap + heat * dtr);
        g[2] = -2.0 * a * heat * ex2 * x * x;
        g[3] = a * trap * (e_ni - e_co) * fr * (1.0 - fr);
        g[4] = 0.0;
        return true;
    }
};

template <> struct ModelGrad<4, double> : Model<4, double> {
    __device__ __forceinline__ bool pred_grad(double t, double& f, double* g) const {
        if (!pred(t, f)) return false;
        double ph = t - t0;
        double ps = ref_softplus(ph), sp = ref_softplus_grad(ph);
        double w = 1.0 + ps / tsd;
        double sd = 1.0 / (w * w), w3 = w * w * w;
        double x = ps / tdf, ex2 = ::exp(-x * x), trap = 1.0 - ex2;
        g[0] = a * sd * trap;
        double dsd = -2.0 / (w3 * tsd);
        double dtr = 2.0 * ps * ex2 / (tdf * tdf);
        g[1] = a * (-sp) * (dsd * trap + sd * dtr);
        g[2] = a * trap * 2.0 * ps / (w3 * tsd);
        g[3] = -2.0 * a * sd * ex2 * x * x;
        g[4] = 0.0;
        return true;
    }
};

template <> struct ModelGrad<5, double> : Model<5, double> {
    __device__ __forceinline__ bool pred_grad(double t, double& f, double* g) const {
        if (!pred(t, f)) return false;
        double ph = t - t0;
        double s5 = 1.0 / (1.0 + ::exp(-ph * 5.0));
        double ps = ref_softplus(ph), sp = ref_softplus_grad(ph);
        double cool = ::pow(ps, -n), ratio = ps / ttr, cut = ::exp(-ratio * ratio);
        double base = cool * cut, flux = a * s5 * base;
        g[0] = flux;
        g[1] = a * base * (-5.0 * s5 * (1.0 - s5) + s5 * sp * (n / ps + 2.0 * ps / (ttr * ttr)));
        g[2] = -flux * ::log(ps);
        g[3] = flux * 2.0 * ratio * ratio;
        g[4] = 0.0;
        return true;
    }
};

template <> struct ModelGrad<6, double> : Model<6, double> {
    __device__ __forceinline__ bool pred_grad(double t, double& f, double* g) const {
        if (!pred(t, f)) return false;
        double ph = t - t0;
        double ps = ref_softplus(ph), sp = ref_softplus_grad(ph);
        double lr = ::log(ps / tb);
        double u1 = ::exp(2.0 * a1 * lr), u2 = ::exp(2.0 * a2 * lr), u = u1 + u2;
        double u15 = ::pow(u, -1.5);
        g[0] = a * ::pow(u, -0.5);
        double dps = (2.0 * a1 * u1 + 2.0 * a2 * u2) / ps;
        g[1] = a * (-0.5) * u15 * dps * (-sp);
        g[2] = a * (-0.5) * u15 * (-(2.0 * a1 * u1 + 2.0 * a2 * u2));
        g[3] = a * (-0.5) * u15 * 2.0 * lr * u1;
        g[4] = a * (-0.5) * u15 * 2.0 * lr * u2;
        g[5] = 0.0;
        return true;
    }
};

// MetzgerKN: forward differences of the per-observation ODE in FP64 with the
// reference's step h = max(1e-5, 1e-5 |theta_j|) (cuda/svi.cu:180-192).
template <typename R> struct ModelGrad<7, R> : ModelStable<7, R> {
    using B = ModelStable<7, R>;
    __device__ __forceinline__ bool pred_grad(R t, R& f, R* g) const {
        double p[4] = {B::p[0], B::p[1], B::p[2], B::p[3]};
        double f0 = metzger_kn_single(p, (double)t);
        f = (R)f0;
        for (int j = 0; j < 4; j++) {
            double h = fmax(1e-5, 1e-5 * fabs(p[j]));
            double keep = p[j];
            p[j] = keep + h;
            double fp = metzger_kn_single(p, (double)t);
            p[j] = keep;
            g[j] = __builtin_isfinite(fp) ? (R)((fp - f0) / h) : R(0);
        }
        g[4] = R(0);
        return true;
    }
};

}  // namespace lcf
