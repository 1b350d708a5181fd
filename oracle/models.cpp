// models.cpp -- oracle restatement of the parametric light-curve models.
//
// TEST INFRASTRUCTURE ONLY (parity checker for the HIP product).
// Formulas follow src/parametric.rs:199-648 (CPU) and cuda/models_device.h /
// cuda/models.cu (GPU) of the reference, evaluated directly in FP64 exactly as
// written there (no re-association, no overflow guards), so that the oracle
// reproduces the reference's non-finite cases too.
#include "lcf_oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace orc {

static inline double sigmoid(double x) { return 1.0 / (1.0 + std::exp(-x)); }
static inline double softplus(double x) { return std::log(1.0 + std::exp(x)) + 1e-6; }
static inline double softplus_grad(double x) { return std::exp(x) / (1.0 + std::exp(x)); }

int n_params(int m) {
    // GpuModelName::n_params (src/gpu_types.rs:23-34)
    static const int np[8] = {6, 7, 7, 5, 5, 5, 6, 5};
    return (m >= 0 && m < 8) ? np[m] : -1;
}

// bazin_flux_eval (src/parametric.rs:200-211)
static double bazin(const double* p, double t) {
    double a = std::exp(p[0]), b = p[1], t0 = p[2];
    double tr = std::exp(p[3]), tf = std::exp(p[4]);
    double dt = t - t0;
    double e_fall = std::exp(-dt / tf);
    double sig = 1.0 / (1.0 + std::exp(-dt / tr));
    return a * e_fall * sig + b;
}

// villar_flux_eval (src/parametric.rs:231-247)
static double villar(const double* p, double t) {
    double a = std::exp(p[0]), beta = p[1], gamma = std::exp(p[2]), t0 = p[3];
    double tr = std::exp(p[4]), tf = std::exp(p[5]);
    double ph = t - t0;
    double sr = 1.0 / (1.0 + std::exp(-ph / tr));
    double w = 1.0 / (1.0 + std::exp(-10.0 * (ph - gamma)));
    double pl = 1.0 - beta * ph;
    double pr = (1.0 - beta * gamma) * std::exp((gamma - ph) / tf);
    return a * sr * ((1.0 - w) * pl + w * pr);
}

// tde_flux_eval (src/parametric.rs:297-311)
static double tde(const double* p, double t) {
    double a = std::exp(p[0]), b = p[1], t0 = p[2];
    double tr = std::exp(p[3]), tf = std::exp(p[4]), alpha = p[5];
    double ph = t - t0;
    double sig = 1.0 / (1.0 + std::exp(-ph / tr));
    double ps = std::log(1.0 + std::exp(ph)) + 1e-6;
    double w = 1.0 + ps / tf;
    return a * sig * std::pow(w, -alpha) + b;
}

// arnett_flux_eval (src/parametric.rs:339-356)
static double arnett(const double* p, double t) {
    double a = std::exp(p[0]), t0 = p[1], tm = std::exp(p[2]), lf = p[3];
    double ps = std::log(1.0 + std::exp(t - t0)) + 1e-6;
    double f = 1.0 / (1.0 + std::exp(-lf));
    double e_ni = std::exp(-ps / 8.8), e_co = std::exp(-ps / 111.3);
    double heat = f * e_ni + (1.0 - f) * e_co;
    double x = ps / tm;
    return a * heat * (1.0 - std::exp(-x * x));
}

// magnetar_flux_eval (src/parametric.rs:387-400)
static double magnetar(const double* p, double t) {
    double a = std::exp(p[0]), t0 = p[1], tsd = std::exp(p[2]), tdf = std::exp(p[3]);
    double ps = std::log(1.0 + std::exp(t - t0)) + 1e-6;
    double w = 1.0 + ps / tsd;
    double spindown = 1.0 / (w * w);
    double x = ps / tdf;
    return a * spindown * (1.0 - std::exp(-x * x));
}

// shockcooling_flux_eval (src/parametric.rs:427-440)
static double shock_cooling(const double* p, double t) {
    double a = std::exp(p[0]), t0 = p[1], n = p[2], ttr = std::exp(p[3]);
    double ph = t - t0;
    double sig5 = 1.0 / (1.0 + std::exp(-ph * 5.0));
    double ps = std::log(1.0 + std::exp(ph)) + 1e-6;
    double ratio = ps / ttr;
    return a * sig5 * std::pow(ps, -n) * std::exp(-ratio * ratio);
}

// afterglow_flux_eval (src/parametric.rs:468-483)
static double afterglow(const double* p, double t) {
    double a = std::exp(p[0]), t0 = p[1], tb = std::exp(p[2]), a1 = p[3], a2 = p[4];
    double ps = std::log(1.0 + std::exp(t - t0)) + 1e-6;
    double ln_r = std::log(ps / tb);
    double u1 = std::exp(2.0 * a1 * ln_r), u2 = std::exp(2.0 * a2 * ln_r);
    return a * std::pow(u1 + u2, -0.5);
}

// Metzger 1-zone kilonova, shared Euler integrator.  The grid is log-spaced
// on [0.01, t_hi] days with n_grid points; see src/parametric.rs:544-648
// (batch form, 200 points, t_hi = 1.05 * max phase) and cuda/models.cu:429-514
// (per-observation form, 100 points, t_hi = max(1.05 * phase, 0.02)).
static const double MSUN = 1.989e33, CCGS = 2.998e10, DAY = 86400.0;

static void metzger_grid(const double* p, double log_t_max, int n_grid,
                         std::vector<double>& gt, std::vector<double>& gl) {
    const double PI = 3.14159265358979323846;
    double m_ej = std::pow(10.0, p[0]) * MSUN;
    double v_ej = std::pow(10.0, p[1]) * CCGS;
    double kappa_r = std::pow(10.0, p[2]);
    double log_t_min = std::log(0.01);
    gt.resize(n_grid);
    gl.resize(n_grid);
    for (int i = 0; i < n_grid; i++)
        gt[i] = std::exp(log_t_min + (log_t_max - log_t_min) * (double)i / (double)(n_grid - 1));
    double ye = 0.1, xn0 = 1.0 - 2.0 * ye, scale = 1e40;
    double e0 = 0.5 * m_ej * v_ej * v_ej;
    double e_th = e0 / scale, e_kin = e0 / scale, v = v_ej;
    double r = gt[0] * DAY * v;
    for (int i = 0; i < n_grid; i++) {
        double t_day = gt[i], t_sec = t_day * DAY;
        double ef = 0.34 * std::pow(t_day, 0.74);
        double eth = 0.36 * (std::exp(-0.56 * t_day) + (ef > 1e-10 ? std::log(1.0 + ef) / ef : 1.0));
        double xn = xn0 * std::exp(-t_sec / 900.0);
        double eps_n = 3.2e14 * xn;
        double tt = 0.5 - std::atan((t_sec - 1.3) / 0.11) / PI;
        if (tt < 1e-30) tt = 1e-30;
        double eps_rp = 2e18 * eth * std::pow(tt, 1.3);
        double l_heat = m_ej * (eps_n + eps_rp) / scale;
        double xr = 1.0 - xn0, xnd = xn0 - xn;
        double kappa = 0.4 * xnd + kappa_r * xr;
        double t_diff = 3.0 * kappa * m_ej / (4.0 * PI * CCGS * v * t_sec) + r / CCGS;
        double l_rad = (e_th > 0.0 && t_diff > 0.0) ? e_th / t_diff : 0.0;
        gl[i] = l_rad;
        double l_pdv = r > 0.0 ? e_th * v / r : 0.0;
        if (i < n_grid - 1) {
            double dt_sec = (gt[i + 1] - gt[i]) * DAY;
            e_th += (l_heat - l_pdv - l_rad) * dt_sec;
            if (e_th < 0.0) e_th = 0.0;
            e_kin += l_pdv * dt_sec;
            v = std::sqrt(2.0 * e_kin * scale / m_ej);
            if (v > CCGS) v = CCGS;
            r += v * dt_sec;
        }
    }
}

// metzger_kn_eval_batch (src/parametric.rs:544-648)
static void metzger_batch(const double* p, const double* times, int64_t n, double* out) {
    double t0 = p[3];
    double phase_max = 0.01;
    for (int64_t i = 0; i < n; i++) phase_max = std::max(phase_max, times[i] - t0);
    if (phase_max <= 0.01) {
        for (int64_t i = 0; i < n; i++) out[i] = 0.0;
        return;
    }
    const int NG = 200;
    std::vector<double> gt, gl;
    metzger_grid(p, std::log(phase_max * 1.05), NG, gt, gl);
    double lp = -INFINITY;
    for (double l : gl) lp = std::max(lp, l);
    if (lp <= 0.0 || !std::isfinite(lp)) {
        for (int64_t i = 0; i < n; i++) out[i] = 0.0;
        return;
    }
    for (double& l : gl) l /= lp;
    for (int64_t k = 0; k < n; k++) {
        double ph = times[k] - t0;
        double val;
        if (ph <= 0.0) val = 0.0;
        else if (ph <= gt[0]) val = gl[0];
        else if (ph >= gt[NG - 1]) val = gl[NG - 1];
        else {
            // first index with gt[idx] >= ph, clamped to [1, NG-1]
            int idx = (int)(std::lower_bound(gt.begin(), gt.end(), ph) - gt.begin());
            idx = std::max(1, std::min(idx, NG - 1));
            double fr = (ph - gt[idx - 1]) / (gt[idx] - gt[idx - 1]);
            val = gl[idx - 1] + fr * (gl[idx] - gl[idx - 1]);
        }
        out[k] = val;
    }
}

// lc_metzger_kn_at_single (cuda/models_device.h:376-455, models.cu:429-514).
// Written out as its own loop because the interpolation is fused with the
// integration there (the value at the phase is taken from the bracketing
// grid interval while integrating, and the last/first-point rules differ).
double metzger_single(const double* p, double obs_time) {
    const double PI = 3.14159265358979323846;
    double m_ej = std::pow(10.0, p[0]) * MSUN;
    double v_ej = std::pow(10.0, p[1]) * CCGS;
    double kappa_r = std::pow(10.0, p[2]);
    double phase = obs_time - p[3];
    if (phase <= 0.0) return 0.0;
    double phase_max = phase * 1.05;
    if (phase_max < 0.02) phase_max = 0.02;
    double lmin = std::log(0.01), lmax = std::log(phase_max);
    const int NG = 100;
    double ye = 0.1, xn0 = 1.0 - 2.0 * ye, scale = 1e40;
    double e0 = 0.5 * m_ej * v_ej * v_ej;
    double e_th = e0 / scale, e_kin = e0 / scale, v = v_ej;
    double g0 = std::exp(lmin);
    double r = g0 * DAY * v;
    double l_peak = 0.0, l_at = 0.0, prev_t = g0, prev_l = 0.0;
    for (int i = 0; i < NG; i++) {
        double t_day = std::exp(lmin + (lmax - lmin) * (double)i / (double)(NG - 1));
        double t_sec = t_day * DAY;
        double ef = 0.34 * std::pow(t_day, 0.74);
        double eth = 0.36 * (std::exp(-0.56 * t_day) + (ef > 1e-10 ? std::log(1.0 + ef) / ef : 1.0));
        double xn = xn0 * std::exp(-t_sec / 900.0);
        double eps_n = 3.2e14 * xn;
        double tt = 0.5 - std::atan((t_sec - 1.3) / 0.11) / PI;
        if (tt < 1e-30) tt = 1e-30;
        double eps_rp = 2e18 * eth * std::pow(tt, 1.3);
        double l_heat = m_ej * (eps_n + eps_rp) / scale;
        double xr = 1.0 - xn0, xnd = xn0 - xn;
        double kappa = 0.4 * xnd + kappa_r * xr;
        double t_diff = 3.0 * kappa * m_ej / (4.0 * PI * CCGS * v * t_sec) + r / CCGS;
        double l_rad = (e_th > 0.0 && t_diff > 0.0) ? e_th / t_diff : 0.0;
        if (l_rad > l_peak) l_peak = l_rad;
        if (i > 0 && phase >= prev_t && phase <= t_day) {
            double fr = (phase - prev_t) / (t_day - prev_t);
            l_at = prev_l + fr * (l_rad - prev_l);
        }
        if (i == NG - 1 && phase >= t_day) l_at = l_rad;
        if (i == 0 && phase <= t_day) l_at = l_rad;
        prev_t = t_day;
        prev_l = l_rad;
        double l_pdv = r > 0.0 ? e_th * v / r : 0.0;
        if (i < NG - 1) {
            double tn = std::exp(lmin + (lmax - lmin) * (double)(i + 1) / (double)(NG - 1));
            double dt_sec = (tn - t_day) * DAY;
            e_th += (l_heat - l_pdv - l_rad) * dt_sec;
            if (e_th < 0.0) e_th = 0.0;
            e_kin += l_pdv * dt_sec;
            v = std::sqrt(2.0 * e_kin * scale / m_ej);
            if (v > CCGS) v = CCGS;
            r += v * dt_sec;
        }
    }
    if (l_peak <= 0.0 || !std::isfinite(l_peak)) return 0.0;
    return l_at / l_peak;
}

double eval_scalar(int m, const double* p, double t) {
    switch (m) {
        case 0: return bazin(p, t);
        case 1: return villar(p, t);
        case 2: return tde(p, t);
        case 3: return arnett(p, t);
        case 4: return magnetar(p, t);
        case 5: return shock_cooling(p, t);
        case 6: return afterglow(p, t);
        default: return 0.0;
    }
}

void eval_times(int m, const double* p, const double* t, int64_t n, double* out) {
    if (m == 7) { metzger_batch(p, t, n, out); return; }
    for (int64_t i = 0; i < n; i++) out[i] = eval_scalar(m, p, t[i]);
}

// Analytic gradients, src/parametric.rs:213-510 (and the identical
// cuda/models_device.h:152-347).
int eval_grad(int m, const double* p, double t, double* o) {
    switch (m) {
    case 0: {  // src/parametric.rs:213-228
        double a = std::exp(p[0]), t0 = p[2], tr = std::exp(p[3]), tf = std::exp(p[4]);
        double dt = t - t0;
        double ef = std::exp(-dt / tf);
        double sig = 1.0 / (1.0 + std::exp(-dt / tr));
        double base = a * ef * sig;
        o[0] = base; o[1] = 1.0;
        o[2] = base * (1.0 / tf - (1.0 - sig) / tr);
        o[3] = -base * (1.0 - sig) * dt / tr;
        o[4] = base * dt / tf;
        o[5] = 0.0;
        return 0;
    }
    case 1: {  // src/parametric.rs:249-294
        double a = std::exp(p[0]), beta = p[1], gamma = std::exp(p[2]), t0 = p[3];
        double tr = std::exp(p[4]), tf = std::exp(p[5]);
        double ph = t - t0, k = 10.0;
        double sr = 1.0 / (1.0 + std::exp(-ph / tr));
        double w = 1.0 / (1.0 + std::exp(-k * (ph - gamma)));
        double pl = 1.0 - beta * ph;
        double ed = std::exp((gamma - ph) / tf);
        double pr = (1.0 - beta * gamma) * ed;
        double piece = (1.0 - w) * pl + w * pr;
        o[0] = a * sr * piece;
        o[1] = a * sr * ((1.0 - w) * (-ph) + w * (-gamma * ed));
        double dw_dlg = -k * w * (1.0 - w) * gamma;
        double dpr_dlg = ed * (-beta + (1.0 - beta * gamma) / tf) * gamma;
        o[2] = a * sr * (dw_dlg * (pr - pl) + w * dpr_dlg);
        double dsig_dt0 = -sr * (1.0 - sr) / tr;
        double dw_dt0 = -k * w * (1.0 - w);
        double dpiece_dt0 = dw_dt0 * (pr - pl) + (1.0 - w) * beta + w * (1.0 - beta * gamma) * ed / tf;
        o[3] = a * (dsig_dt0 * piece + sr * dpiece_dt0);
        o[4] = a * piece * sr * (1.0 - sr) * (-ph / tr);
        o[5] = a * sr * w * (pr * (ph - gamma) / tf);
        o[6] = 0.0;
        return 0;
    }
    case 2: {  // src/parametric.rs:313-336
        double a = std::exp(p[0]), t0 = p[2], tr = std::exp(p[3]), tf = std::exp(p[4]), al = p[5];
        double ph = t - t0;
        double sig = 1.0 / (1.0 + std::exp(-ph / tr));
        double ps = softplus(ph), sp = softplus_grad(ph);
        double w = 1.0 + ps / tf;
        double decay = std::pow(w, -al);
        o[0] = a * sig * decay; o[1] = 1.0;
        o[2] = a * ((-sig * (1.0 - sig) / tr) * decay + sig * (al * std::pow(w, -al - 1.0) * sp / tf));
        o[3] = a * decay * (-sig * (1.0 - sig) * ph / tr);
        o[4] = a * sig * al * std::pow(w, -al - 1.0) * ps / tf;
        o[5] = a * sig * (-std::log(w)) * decay;
        o[6] = 0.0;
        return 0;
    }
    case 3: {  // src/parametric.rs:358-384
        double a = std::exp(p[0]), t0 = p[1], tm = std::exp(p[2]), lf = p[3];
        double ph = t - t0;
        double ps = softplus(ph), sp = softplus_grad(ph);
        double f = 1.0 / (1.0 + std::exp(-lf));
        double e_ni = std::exp(-ps / 8.8), e_co = std::exp(-ps / 111.3);
        double heat = f * e_ni + (1.0 - f) * e_co;
        double x = ps / tm, ex2 = std::exp(-x * x), trap = 1.0 - ex2;
        o[0] = a * heat * trap;
        double dh = -f * e_ni / 8.8 - (1.0 - f) * e_co / 111.3;
        double dtr = 2.0 * ps * ex2 / (tm * tm);
        o[1] = a * (-sp) * (dh * trap + heat * dtr);
        o[2] = -2.0 * a * heat * ex2 * x * x;
        o[3] = a * trap * (e_ni - e_co) * f * (1.0 - f);
        o[4] = 0.0;
        return 0;
    }
    case 4: {  // src/parametric.rs:402-424
        double a = std::exp(p[0]), t0 = p[1], tsd = std::exp(p[2]), tdf = std::exp(p[3]);
        double ph = t - t0;
        double ps = softplus(ph), sp = softplus_grad(ph);
        double w = 1.0 + ps / tsd;
        double sd = 1.0 / (w * w), w3 = w * w * w;
        double x = ps / tdf, ex2 = std::exp(-x * x), trap = 1.0 - ex2;
        o[0] = a * sd * trap;
        double dsd = -2.0 / (w3 * tsd);
        double dtr = 2.0 * ps * ex2 / (tdf * tdf);
        o[1] = a * (-sp) * (dsd * trap + sd * dtr);
        o[2] = a * trap * 2.0 * ps / (w3 * tsd);
        o[3] = -2.0 * a * sd * ex2 * x * x;
        o[4] = 0.0;
        return 0;
    }
    case 5: {  // src/parametric.rs:442-465
        double a = std::exp(p[0]), t0 = p[1], n = p[2], ttr = std::exp(p[3]);
        double ph = t - t0;
        double s5 = 1.0 / (1.0 + std::exp(-ph * 5.0));
        double ps = softplus(ph), sp = softplus_grad(ph);
        double cool = std::pow(ps, -n), ratio = ps / ttr, cut = std::exp(-ratio * ratio);
        double base = cool * cut, flux = a * s5 * base;
        o[0] = flux;
        o[1] = a * base * (-5.0 * s5 * (1.0 - s5) + s5 * sp * (n / ps + 2.0 * ps / (ttr * ttr)));
        o[2] = -flux * std::log(ps);
        o[3] = flux * 2.0 * ratio * ratio;
        o[4] = 0.0;
        return 0;
    }
    case 6: {  // src/parametric.rs:485-510
        double a = std::exp(p[0]), t0 = p[1], tb = std::exp(p[2]), a1 = p[3], a2 = p[4];
        double ph = t - t0;
        double ps = softplus(ph), sp = softplus_grad(ph);
        double lr = std::log(ps / tb);
        double u1 = std::exp(2.0 * a1 * lr), u2 = std::exp(2.0 * a2 * lr), u = u1 + u2;
        double u15 = std::pow(u, -1.5);
        o[0] = a * std::pow(u, -0.5);
        double dps = (2.0 * a1 * u1 + 2.0 * a2 * u2) / ps;
        o[1] = a * (-0.5) * u15 * dps * (-sp);
        o[2] = a * (-0.5) * u15 * (-(2.0 * a1 * u1 + 2.0 * a2 * u2));
        o[3] = a * (-0.5) * u15 * 2.0 * lr * u1;
        o[4] = a * (-0.5) * u15 * 2.0 * lr * u2;
        o[5] = 0.0;
        return 0;
    }
    default:
        return -1;
    }
}

}  // namespace orc

extern "C" {

int orc_n_params(int m) { return orc::n_params(m); }

int orc_eval_batch(int m, const double* params, int64_t n_draws, const double* times,
                   int64_t n_times, double* out) {
    int np = orc::n_params(m);
    if (np < 0 || n_draws < 0 || n_times < 0) return -1;
    for (int64_t d = 0; d < n_draws; d++)
        orc::eval_times(m, params + d * np, times, n_times, out + d * n_times);
    return 0;
}

double orc_metzger_kn_single(const double* p, double t) { return orc::metzger_single(p, t); }

int orc_eval_grad(int m, const double* p, double t, double* out) { return orc::eval_grad(m, p, t, out); }

}  // extern "C"
