// nonparametric.cpp -- oracle restatement of the per-band nonparametric GP
// pipeline and its light-curve features.
//
// TEST INFRASTRUCTURE ONLY (parity checker for the HIP product, and the CPU
// baseline of the nonparametric workload).
//
// process_band (src/nonparametric.rs:1228-1525) in FP64 with the reference's
// expressions and loop orders:
//   * detections: flux = 10^(-0.4 (m - 23.9)), flux_err = err / 1.0857.. flux,
//     kept when flux > 0, flux_err > 0 and flux / flux_err >= 3
//     (src/common.rs:32-36, src/nonparametric.rs:46-70); bands with fewer than
//     5 observations or 5 detections give no result;
//   * minimum length scale max(2 median(positive gaps of the sorted detection
//     times), 0.1) (gaps empty: median 1);
//   * candidates amp-major over amps x ls_factors, l = max(duration / f, 0.1),
//     skipped below the minimum length scale; dense GP on every detection
//     (src/sparse_gp.rs:72-176) scored by train RMS when n_det <= dense_max,
//     else FITC (src/sparse_gp.rs:183-328, m = min(n_inducing, n_det))
//     scored by approx_nlml, plus a dense GP on the subsample of
//     src/gp.rs:5-27 (first candidate that factorises, no minimum length
//     scale) kept for thermal reuse; first strict minimum of a finite score;
//   * predictions on the n_grid-point grid t_lo .. t_hi, std rescale by
//     0.6 clamp(rms / mean std, 0.05, 5) with the std taken at (a 200-point
//     subsample of) the detection times, chi2 / baseline chi2, peak on the
//     grid, peak_mag >= brightest detection - 2;
//   * compute_fine_grid_metrics (src/nonparametric.rs:503-554) with the
//     helpers of src/common.rs:168-460, compute_predictive_features (:211-398),
//     fit_decay_power_law (:407-472), predict_mag_at_offset (:475-485) and the
//     variability features of :586-770.
// Extension: the analytic time derivative of the predictive mean on the grid.
// Features are written in the field order of NonparametricBandResult
// (src/nonparametric.rs:74-160), ORC_NP_NFEAT doubles per band, NaN for None.
#include "lcf_oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace orc {
namespace np {

static const double FACTOR = 1.0857362047581294;  // 2.5 / ln 10 (src/common.rs:30)

static inline double fmax_r(double a, double b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a > b ? a : b)); }
static inline double fmin_r(double a, double b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a < b ? a : b)); }

// detection_mask (src/nonparametric.rs:46-57) through mag2flux (src/common.rs:32-36)
static bool is_detection(double m, double e) {
    double flux = std::pow(10.0, -0.4 * (m - 23.9));
    double ferr = e / FACTOR * flux;
    return flux > 0.0 && ferr > 0.0 && (flux / ferr) >= 3.0;
}

// rbf (src/sparse_gp.rs:17-20)
static inline double rbf(double x1, double x2, double amp, double it) {
    double d = x1 - x2;
    return amp * std::exp(-d * d * it);
}

// cholesky / solve_l / solve_lt (src/sparse_gp.rs:27-65), row-major n x n
static bool chol(std::vector<double>& a, int n) {
    for (int j = 0; j < n; j++) {
        double s = a[j * n + j];
        for (int k = 0; k < j; k++) s -= a[j * n + k] * a[j * n + k];
        if (s <= 0.0) return false;
        a[j * n + j] = std::sqrt(s);
        double ljj = a[j * n + j];
        for (int i = j + 1; i < n; i++) {
            double si = a[i * n + j];
            for (int k = 0; k < j; k++) si -= a[i * n + k] * a[j * n + k];
            a[i * n + j] = si / ljj;
        }
        for (int i = 0; i < j; i++) a[i * n + j] = 0.0;
    }
    return true;
}
static void sol_l(const std::vector<double>& l, const double* b, double* x, int n) {
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int j = 0; j < i; j++) s -= l[i * n + j] * x[j];
        x[i] = s / l[i * n + i];
    }
}
static void sol_lt(const std::vector<double>& l, const double* b, double* x, int n) {
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < n; j++) s -= l[j * n + i] * x[j];
        x[i] = s / l[i * n + i];
    }
}

// A fitted predictor: dense (x = training times) or FITC (x = inducing points).
struct Gp {
    bool ok = false, sparse = false;
    int n = 0;
    std::vector<double> x, alpha, l, lb;
    double amp = 0, it = 0, ymean = 0;

    double mean(double t) const {
        double dot = 0.0;
        for (int i = 0; i < n; i++) dot += rbf(t, x[i], amp, it) * alpha[i];
        return dot + ymean;
    }
    double dmean(double t) const {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += -2.0 * it * (t - x[i]) * rbf(t, x[i], amp, it) * alpha[i];
        return s;
    }
    double stdv(double t) const {
        std::vector<double> ks(n), v(n);
        for (int i = 0; i < n; i++) ks[i] = rbf(t, x[i], amp, it);
        sol_l(l, ks.data(), v.data(), n);
        double q = 0.0;
        for (double z : v) q += z * z;
        if (!sparse) return std::sqrt(fmax_r(amp - q, 1e-10));
        sol_l(lb, ks.data(), v.data(), n);
        double qb = 0.0;
        for (double z : v) qb += z * z;
        return std::sqrt(fmax_r(amp - q + qb, 1e-10));
    }
};

// DenseGP::fit (src/sparse_gp.rs:83-116)
static Gp dense_fit(const std::vector<double>& t, const std::vector<double>& y, const std::vector<double>& nv,
                    double amp, double ls) {
    Gp g;
    const int n = (int)t.size();
    if (n == 0) return g;
    g.n = n;
    g.amp = amp;
    g.it = 0.5 / (ls * ls);
    double ys = 0.0;
    for (double v : y) ys += v;
    g.ymean = ys / n;
    std::vector<double> k(n * n);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j <= i; j++) {
            double v = rbf(t[i], t[j], amp, g.it);
            k[i * n + j] = v;
            k[j * n + i] = v;
        }
        k[i * n + i] += fmax_r(nv[i], 1e-10);
    }
    if (!chol(k, n)) return g;
    std::vector<double> yc(n), tmp(n);
    for (int i = 0; i < n; i++) yc[i] = y[i] - g.ymean;
    g.alpha.resize(n);
    sol_l(k, yc.data(), tmp.data(), n);
    sol_lt(k, tmp.data(), g.alpha.data(), n);
    g.x = t;
    g.l = std::move(k);
    g.ok = true;
    return g;
}

// DenseGP::train_rms (src/sparse_gp.rs:168-175)
static double train_rms(const Gp& g, const std::vector<double>& y) {
    double rss = 0.0;
    for (int i = 0; i < g.n; i++) {
        double p = g.mean(g.x[i]);
        rss += (p - y[i]) * (p - y[i]);
    }
    return std::sqrt(rss / (double)std::max(g.n, 1));
}

// SparseGP::fit (src/sparse_gp.rs:195-271)
static Gp sparse_fit(const std::vector<double>& t, const std::vector<double>& y, const std::vector<double>& nv,
                     double amp, double ls, int m_req) {
    Gp g;
    const int n = (int)t.size();
    if (n == 0 || m_req == 0) return g;
    const int m = std::min(m_req, n);
    g.sparse = true;
    g.n = m;
    g.amp = amp;
    g.it = 0.5 / (ls * ls);
    double ys = 0.0;
    for (double v : y) ys += v;
    g.ymean = ys / n;
    double tmin = INFINITY, tmax = -INFINITY;
    for (double v : t) { tmin = fmin_r(tmin, v); tmax = fmax_r(tmax, v); }
    g.x.resize(m);
    for (int i = 0; i < m; i++) g.x[i] = tmin + (double)i * (tmax - tmin) / (double)std::max(m - 1, 1);
    std::vector<double> kmm(m * m);
    for (int i = 0; i < m; i++) {
        for (int j = 0; j <= i; j++) {
            double v = rbf(g.x[i], g.x[j], amp, g.it);
            kmm[i * m + j] = v;
            kmm[j * m + i] = v;
        }
        kmm[i * m + i] += 1e-6;
    }
    std::vector<double> lmm = kmm;
    if (!chol(lmm, m)) return g;
    std::vector<double> b = kmm, w(m, 0.0), kc(m), vc(m);
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < m; i++) kc[i] = rbf(g.x[i], t[j], amp, g.it);
        sol_l(lmm, kc.data(), vc.data(), m);
        double q = 0.0;
        for (double z : vc) q += z * z;
        double lam = fmax_r(amp - q, 0.0) + nv[j];
        double li = 1.0 / fmax_r(lam, 1e-12);
        for (int r = 0; r < m; r++) {
            double sc = li * kc[r];
            for (int c = r; c < m; c++) b[r * m + c] += sc * kc[c];
        }
        double val = li * (y[j] - g.ymean);
        for (int i = 0; i < m; i++) w[i] += kc[i] * val;
    }
    for (int r = 0; r < m; r++)
        for (int c = 0; c < r; c++) b[r * m + c] = b[c * m + r];
    if (!chol(b, m)) return g;
    std::vector<double> tmp(m);
    g.alpha.resize(m);
    sol_l(b, w.data(), tmp.data(), m);
    sol_lt(b, tmp.data(), g.alpha.data(), m);
    g.l = std::move(lmm);
    g.lb = std::move(b);
    g.ok = true;
    return g;
}

// SparseGP::approx_nlml (src/sparse_gp.rs:315-327)
static double approx_nlml(const Gp& g, const std::vector<double>& t, const std::vector<double>& y,
                          const std::vector<double>& nv) {
    double s = 0.0;
    for (size_t i = 0; i < t.size(); i++) {
        double r = y[i] - g.mean(t[i]);
        s += r * r / nv[i] + std::log(nv[i]);
    }
    return s / (double)t.size();
}

// Index of the first minimum among finite values (0 when none).
static int argmin_finite(const double* v, int n) {
    int best = -1;
    for (int i = 0; i < n; i++)
        if (std::isfinite(v[i]) && (best < 0 || v[i] < v[best])) best = i;
    return best < 0 ? 0 : best;
}

// ---- src/common.rs:168-460 ------------------------------------------------

static double rise_timescale(const double* t, const double* m, int n, int pk) {  // :168-193
    if (pk == 0 || pk >= n) return NAN;
    double pm = m[pk];
    int k = std::min(pk, 3);
    double base = 0.0;
    for (int i = 0; i < k; i++) base += m[i];
    base /= (double)k;
    double target = base + (pm - base) * (1.0 - std::exp(-1.0));
    double best_t = NAN, best_d = INFINITY;
    for (int i = 0; i < pk; i++) {
        double d = std::fabs(m[i] - target);
        if (d < best_d) { best_d = d; best_t = t[i]; }
    }
    if (std::isnan(best_t)) return NAN;
    return t[pk] - best_t;
}

static double decay_crossing(const double* t, const double* m, int n, int pk, double off) {  // :251-275
    if (pk >= std::max(n - 1, 0)) return NAN;
    double thr = m[pk] + off;
    for (int i = pk + 1; i < n; i++)
        if (m[i] >= thr) {
            double dm = m[i] - m[i - 1];
            if (std::fabs(dm) < 1e-15) return t[i] - t[pk];
            double fr = (thr - m[i - 1]) / dm;
            return t[i - 1] + fr * (t[i] - t[i - 1]) - t[pk];
        }
    return NAN;
}

static double rise_crossing(const double* t, const double* m, int n, int pk, double off) {  // :280-304
    if (pk == 0) return NAN;
    double thr = m[pk] + off;
    for (int i = pk - 1; i >= 0; i--)
        if (m[i] >= thr) {
            double dm = m[i] - m[i + 1];
            if (std::fabs(dm) < 1e-15) return t[pk] - t[i];
            double fr = (thr - m[i + 1]) / dm;
            return t[pk] - (t[i + 1] + fr * (t[i] - t[i + 1]));
        }
    return NAN;
}

static double dm15(const double* t, const double* m, int n, int pk) {  // :210-236
    if (pk >= n) return NAN;
    double tt = t[pk] + 15.0;
    for (int i = pk + 1; i < n; i++)
        if (t[i] >= tt) {
            double dt = t[i] - t[i - 1];
            if (std::fabs(dt) < 1e-15) return m[i - 1] - m[pk];
            double fr = (tt - t[i - 1]) / dt;
            return m[i - 1] + fr * (m[i] - m[i - 1]) - m[pk];
        }
    return NAN;
}

static double fwhm(const double* t, const double* m, int n, int pk) {  // :320-344
    if (pk >= n) return NAN;
    double half = m[pk] + 0.75;
    double tb = NAN, ta = NAN;
    for (int i = pk - 1; i >= 0; i--)
        if (m[i] >= half) { tb = t[i]; break; }
    for (int i = pk + 1; i < n; i++)
        if (m[i] >= half) { ta = t[i]; break; }
    if (std::isnan(tb) || std::isnan(ta)) return NAN;
    return ta - tb;
}

static double lin_slope(const double* t, const double* m, int n, double tiny) {
    if (n < 2) return NAN;
    double nf = n, st = 0, sm = 0, stt = 0, stm = 0;
    for (int i = 0; i < n; i++) st += t[i];
    for (int i = 0; i < n; i++) sm += m[i];
    for (int i = 0; i < n; i++) stt += t[i] * t[i];
    for (int i = 0; i < n; i++) stm += t[i] * m[i];
    double den = nf * stt - st * st;
    if (std::fabs(den) < tiny) return NAN;
    return (nf * stm - st * sm) / den;
}

static double rise_rate(const double* t, const double* m, int n) {  // :346-370
    if (n < 2) return NAN;
    int k = (int)std::ceil((double)n * 0.25);
    k = std::min(std::max(k, 2), n - 1);
    return lin_slope(t, m, k, 1e-10);
}

static double decay_rate(const double* t, const double* m, int n) {  // :372-397
    if (n < 2) return NAN;
    int k = (int)std::ceil((double)n * 0.25);
    k = std::min(std::max(k, 2), n);
    return lin_slope(t + (n - k), m + (n - k), k, 1e-10);
}

static double near_rise(const double* t, const double* m, int n, int pk, double win) {  // :421-443
    if (pk == 0 || pk >= n) return NAN;
    double ts = t[pk] - win;
    std::vector<double> st, sm;
    for (int i = 0; i <= pk; i++)
        if (t[i] >= ts) { st.push_back(t[i]); sm.push_back(m[i]); }
    return lin_slope(st.data(), sm.data(), (int)st.size(), 1e-15);
}

static double near_decay(const double* t, const double* m, int n, int pk, double win) {  // :445-469
    if (pk >= std::max(n - 1, 0)) return NAN;
    double te = t[pk] + win;
    std::vector<double> st, sm;
    for (int i = pk; i < n; i++)
        if (t[i] <= te) { st.push_back(t[i]); sm.push_back(m[i]); }
    return lin_slope(st.data(), sm.data(), (int)st.size(), 1e-15);
}

// The eleven fine-grid metrics in FineGridMetrics order (decay_efold, dm15,
// decay_halfmax, fwhm, rise_time, rise_halfmax, rise_efold, rise_rate,
// decay_rate, near_peak_rise_rate, near_peak_decay_rate).
static void curve_metrics(const double* t, const double* m, int n, int pk, double* o) {
    o[0] = decay_crossing(t, m, n, pk, 1.0857362047581294);
    o[1] = dm15(t, m, n, pk);
    o[2] = decay_crossing(t, m, n, pk, 0.7525749891599529);
    o[3] = fwhm(t, m, n, pk);
    o[4] = rise_timescale(t, m, n, pk);
    o[5] = rise_crossing(t, m, n, pk, 0.7525749891599529);
    o[6] = rise_crossing(t, m, n, pk, 1.0857362047581294);
    o[7] = rise_rate(t, m, n);
    o[8] = decay_rate(t, m, n);
    o[9] = near_rise(t, m, n, pk, 30.0);
    o[10] = near_decay(t, m, n, pk, 30.0);
}

// ---- src/nonparametric.rs:586-770 -----------------------------------------

static double von_neumann(const double* t, const double* m, int n) {  // :586-606
    if (n < 5) return NAN;
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return t[a] < t[b]; });
    double nf = n, s = 0.0;
    for (int i = 0; i < n; i++) s += m[idx[i]];
    double mean = s / nf, var = 0.0;
    for (int i = 0; i < n; i++) var += (m[idx[i]] - mean) * (m[idx[i]] - mean);
    var /= (nf - 1.0);
    if (var < 1e-20) return NAN;
    double d2 = 0.0;
    for (int i = 0; i + 1 < n; i++) d2 += (m[idx[i + 1]] - m[idx[i]]) * (m[idx[i + 1]] - m[idx[i]]);
    return d2 / ((nf - 1.0) * var);
}

static double stetson_k(const double* m, const double* e, int n) {  // :611-637
    if (n < 5) return NAN;
    double ws = 0.0, wm = 0.0;
    std::vector<double> w(n);
    for (int i = 0; i < n; i++) w[i] = 1.0 / fmax_r(e[i] * e[i], 1e-30);
    for (int i = 0; i < n; i++) ws += w[i];
    for (int i = 0; i < n; i++) wm += m[i] * w[i];
    double mean = wm / ws, sa = 0.0, sq = 0.0;
    std::vector<double> d(n);
    for (int i = 0; i < n; i++) d[i] = (m[i] - mean) / fmax_r(e[i], 1e-15);
    for (int i = 0; i < n; i++) sa += std::fabs(d[i]);
    for (int i = 0; i < n; i++) sq += d[i] * d[i];
    double rms = std::sqrt(sq / n);
    if (rms < 1e-20) return NAN;
    return (sa / n) / rms;
}

static double excess_var(const double* m, const double* e, int n) {  // :641-656
    if (n < 5) return NAN;
    double nf = n, s = 0.0;
    for (int i = 0; i < n; i++) s += m[i];
    double mean = s / nf;
    if (std::fabs(mean) < 1e-20) return NAN;
    double ex = 0.0;
    for (int i = 0; i < n; i++) ex += (m[i] - mean) * (m[i] - mean) - e[i] * e[i];
    return ex / (nf * mean * mean);
}

static double cusum(const double* m, int n) {  // :660-682
    if (n < 5) return NAN;
    double nf = n, s = 0.0;
    for (int i = 0; i < n; i++) s += m[i];
    double mean = s / nf, v = 0.0;
    for (int i = 0; i < n; i++) v += (m[i] - mean) * (m[i] - mean);
    double sd = std::sqrt(v / (nf - 1.0));
    if (sd < 1e-20) return NAN;
    double norm = nf * sd, c = 0.0, lo = INFINITY, hi = -INFINITY;
    for (int i = 0; i < n; i++) {
        c += m[i] - mean;
        double x = c / norm;
        lo = fmin_r(lo, x);
        hi = fmax_r(hi, x);
    }
    return hi - lo;
}

static void pre_peak(const double* t, const double* m, const double* e, int n, double tp, double pmag,
                     double* rms_out, double* sig_out) {  // :686-724
    std::vector<double> pm, pe;
    for (int i = 0; i < n; i++)
        if (t[i] < tp) { pm.push_back(m[i]); pe.push_back(e[i]); }
    *rms_out = NAN;
    *sig_out = NAN;
    if (pm.size() < 3) return;
    double nf = pm.size(), s = 0.0;
    for (double v : pm) s += v;
    double mean = s / nf, v2 = 0.0;
    for (double v : pm) v2 += (v - mean) * (v - mean);
    *rms_out = std::sqrt(v2 / (nf - 1.0));
    std::stable_sort(pe.begin(), pe.end(), [](double a, double b) { return a < b; });
    double med = pe[pe.size() / 2];
    *sig_out = med > 1e-10 ? (mean - pmag) / med : NAN;
}

static int local_maxima(const double* p, int n, double thr) {  // :730-757
    if (n < 3) return n == 0 ? 0 : 1;
    int c = 0;
    for (int i = 1; i < n - 1; i++) {
        if (!std::isfinite(p[i]) || !std::isfinite(p[i - 1]) || !std::isfinite(p[i + 1])) continue;
        if (p[i] < p[i - 1] - thr && p[i] < p[i + 1] - thr) c++;
    }
    if (std::isfinite(p[0]) && std::isfinite(p[1]) && p[0] < p[1] - thr) c++;
    if (std::isfinite(p[n - 1]) && std::isfinite(p[n - 2]) && p[n - 1] < p[n - 2] - thr) c++;
    return std::max(c, 1);
}

static double monotonicity(const double* p, int n, int pk) {  // :760-774
    if (pk >= std::max(n - 2, 0)) return NAN;
    int len = n - pk;
    if (len < 3) return NAN;
    int fading = 0;
    for (int i = pk; i + 1 < n; i++)
        if (std::isfinite(p[i + 1]) && std::isfinite(p[i]) && p[i + 1] > p[i]) fading++;
    return (double)fading / (double)(len - 1);
}

// fit_decay_power_law (src/nonparametric.rs:407-472)
static void power_law(const double* tg, const double* p, const double* sd, int n, int pk, double* idx_out,
                      double* chi_out) {
    *idx_out = NAN;
    *chi_out = NAN;
    double tpk = tg[pk];
    std::vector<double> lx, my, ww;
    for (int i = pk + 1; i < n; i++) {
        double dt = tg[i] - tpk;
        if (dt < 5.0 || !std::isfinite(p[i])) continue;
        double w = (std::isfinite(sd[i]) && sd[i] > 1e-10) ? 1.0 / (sd[i] * sd[i]) : 1.0;
        lx.push_back(std::log10(dt + 1.0));
        my.push_back(p[i]);
        ww.push_back(w);
    }
    const int k = (int)lx.size();
    if (k < 3) return;
    double sw = 0, swx = 0, swy = 0, swxx = 0, swxy = 0;
    for (int i = 0; i < k; i++) sw += ww[i];
    for (int i = 0; i < k; i++) swx += ww[i] * lx[i];
    for (int i = 0; i < k; i++) swy += ww[i] * my[i];
    for (int i = 0; i < k; i++) swxx += ww[i] * lx[i] * lx[i];
    for (int i = 0; i < k; i++) swxy += ww[i] * lx[i] * my[i];
    double den = sw * swxx - swx * swx;
    if (std::fabs(den) < 1e-20) return;
    double b = (sw * swxy - swx * swy) / den;
    double a = (swy - b * swx) / sw;
    double chi = 0.0;
    for (int i = 0; i < k; i++) {
        double r = my[i] - (a + b * lx[i]);
        chi += ww[i] * r * r;
    }
    *idx_out = b;
    *chi_out = chi / (double)(std::max(k, 3) - 2);
}

// The grid-only predictive features of compute_predictive_features
// (src/nonparametric.rs:266-375): sigma_f, peak_to_peak, dfdt_max, dfdt_min,
// frac_of_peak, post_var_mean, post_var_max, skewness, kurtosis,
// n_inflections.
static void grid_stats(const double* tg, const double* p, const double* sd, int n, double* o) {
    for (int k = 0; k < 10; k++) o[k] = NAN;
    if (n <= 0) return;
    double s = 0.0;
    for (int i = 0; i < n; i++) s += p[i];
    double mean = s / n, v = 0.0;
    for (int i = 0; i < n; i++) v += (p[i] - mean) * (p[i] - mean);
    o[0] = std::sqrt(v / n);
    double mx = -INFINITY, mn = INFINITY;
    for (int i = 0; i < n; i++) { mx = fmax_r(mx, p[i]); mn = fmin_r(mn, p[i]); }
    o[1] = mx - mn;
    if (n > 1) {
        double dg = tg[1] - tg[0], dmx = -INFINITY, dmn = INFINITY;
        for (int i = 0; i + 1 < n; i++) {
            double d = (p[i + 1] - p[i]) / dg;
            dmx = fmax_r(dmx, d);
            dmn = fmin_r(dmn, d);
        }
        o[2] = dmx;
        o[3] = dmn;
    }
    o[4] = p[n - 1] / mn;
    double sv = 0.0, svx = -INFINITY;
    for (int i = 0; i < n; i++) sv += sd[i] * sd[i];
    for (int i = 0; i < n; i++) svx = fmax_r(svx, sd[i] * sd[i]);
    o[5] = sv / n;
    o[6] = svx;
    if (n > 3) {
        double sdev = std::sqrt(v / n);
        if (sdev > 1e-10) {
            double s3 = 0.0, s4 = 0.0;
            for (int i = 0; i < n; i++) {
                double z = (p[i] - mean) / sdev;
                s3 += z * z * z;
            }
            for (int i = 0; i < n; i++) {
                double z = (p[i] - mean) / sdev;
                double z2 = z * z;
                s4 += z2 * z2;
            }
            o[7] = s3 / n;
            o[8] = s4 / n - 3.0;
        }
    }
    if (n > 2) {
        double dg = tg[1] - tg[0];
        std::vector<double> d2;
        for (int i = 1; i + 1 < n; i++) d2.push_back((p[i + 1] - 2.0 * p[i] + p[i - 1]) / (dg * dg));
        int c = 0;
        for (size_t i = 0; i + 1 < d2.size(); i++) {
            double a = d2[i], b = d2[i + 1];
            if (std::isfinite(a) && std::isfinite(b) && std::fabs(a) > 1e-6 && std::fabs(b) > 1e-6 && a * b < 0.0)
                c++;
        }
        o[9] = c;
    }
}

struct BandOut {
    int status = 0;  // 0 no result, 1 dense, 2 FITC
    double feat[ORC_NP_NFEAT];
};

static void band(const double* t, const double* mg, const double* er, int n, double t_lo, double t_hi, int n_grid,
                 const double* amps, int n_amp, const double* lsf, int n_ls, int dense_max, int max_sub, int n_ind,
                 BandOut& o, double* pred, double* sdv, double* dpred, double* state, int cap) {
    for (int k = 0; k < ORC_NP_NFEAT; k++) o.feat[k] = NAN;
    for (int k = 0; k < n_grid; k++) pred[k] = sdv[k] = dpred[k] = NAN;
    for (int k = 0; k < 2 * cap + 4; k++) state[k] = 0.0;
    o.feat[10] = n;
    const double dur = t_hi - t_lo;
    if (n < 5 || !(dur > 0.0)) return;
    std::vector<double> dt_, dm_, de_;
    for (int i = 0; i < n; i++)
        if (is_detection(mg[i], er[i])) { dt_.push_back(t[i]); dm_.push_back(mg[i]); de_.push_back(er[i]); }
    const int nd = (int)dt_.size();
    o.feat[44] = n - nd;
    if (nd < 5) return;
    // minimum length scale (src/nonparametric.rs:1243-1257)
    std::vector<double> st = dt_;
    std::stable_sort(st.begin(), st.end(), [](double a, double b) { return a < b; });
    std::vector<double> gaps;
    for (int i = 1; i < nd; i++) {
        double d = st[i] - st[i - 1];
        if (std::isfinite(d) && d > 0.0) gaps.push_back(d);
    }
    std::stable_sort(gaps.begin(), gaps.end(), [](double a, double b) { return a < b; });
    const double med = gaps.empty() ? 1.0 : gaps[gaps.size() / 2];
    const double min_ls = fmax_r(med * 2.0, 0.1);
    std::vector<double> nv(nd);
    for (int i = 0; i < nd; i++) nv[i] = fmax_r(de_[i] * de_[i], 1e-6);
    Gp gp, thermal;
    double best = INFINITY, fa = NAN, fl = NAN;
    const bool sparse = nd > dense_max;
    for (int ia = 0; ia < n_amp; ia++)
        for (int il = 0; il < n_ls; il++) {
            double ls = fmax_r(dur / lsf[il], 0.1);
            if (ls < min_ls) continue;
            Gp g = sparse ? sparse_fit(dt_, dm_, nv, amps[ia], ls, std::min(n_ind, nd))
                          : dense_fit(dt_, dm_, nv, amps[ia], ls);
            if (!g.ok) continue;
            double sc = sparse ? approx_nlml(g, dt_, dm_, nv) : train_rms(g, dm_);
            if (std::isfinite(sc) && sc < best) {
                best = sc;
                gp = std::move(g);
                fa = amps[ia];
                fl = ls;
            }
        }
    if (!gp.ok) return;
    if (sparse) {  // dense GP on the subsample for thermal reuse (src/nonparametric.rs:1297-1312)
        std::vector<double> ts, ms, ns;
        if (nd <= max_sub) {
            ts = dt_; ms = dm_;
            for (int i = 0; i < nd; i++) ns.push_back(fmax_r(de_[i] * de_[i], 1e-6));
        } else {
            double step = (double)nd / (double)max_sub;
            for (int i = 0; i < max_sub; i++) {
                int k = std::min((int)std::floor(((double)i + 0.5) * step), nd - 1);
                ts.push_back(dt_[k]); ms.push_back(dm_[k]);
                ns.push_back(fmax_r(de_[k] * de_[k], 1e-6));
            }
        }
        for (int ia = 0; ia < n_amp && !thermal.ok; ia++)
            for (int il = 0; il < n_ls && !thermal.ok; il++)
                thermal = dense_fit(ts, ms, ns, amps[ia], fmax_r(dur / lsf[il], 0.1));
    } else {
        thermal = gp;
    }
    o.status = sparse ? 2 : 1;
    if (thermal.ok && thermal.n <= cap) {
        for (int i = 0; i < thermal.n; i++) { state[i] = thermal.alpha[i]; state[cap + i] = thermal.x[i]; }
        state[2 * cap] = thermal.amp;
        state[2 * cap + 1] = thermal.it;
        state[2 * cap + 2] = thermal.ymean;
        state[2 * cap + 3] = thermal.n;
    }
    // grid predictions (src/nonparametric.rs:1343-1356)
    std::vector<double> tg(n_grid);
    for (int i = 0; i < n_grid; i++) tg[i] = t_lo + (double)i * dur / (double)(n_grid - 1);
    for (int k = 0; k < n_grid; k++) {
        pred[k] = gp.mean(tg[k]);
        sdv[k] = gp.stdv(tg[k]);
        dpred[k] = gp.dmean(tg[k]);
    }
    double rss = 0.0;
    std::vector<double> pobs(nd);
    for (int i = 0; i < nd; i++) pobs[i] = gp.mean(dt_[i]);
    for (int i = 0; i < nd; i++) rss += (dm_[i] - pobs[i]) * (dm_[i] - pobs[i]);
    const double rms = std::sqrt(rss / nd);
    // std rescale (:1367-1395)
    {
        std::vector<double> ss;
        if (nd > 200) {
            double step = (double)nd / 200.0;
            for (int i = 0; i < 200; i++) ss.push_back(dt_[(int)(((double)i + 0.5) * step)]);
        } else {
            ss = dt_;
        }
        double sum = 0.0;
        int cnt = 0;
        for (double q : ss) {
            double s = gp.stdv(q);
            if (std::isfinite(s)) { sum += s; cnt++; }
        }
        if (cnt > 0) {
            double ms = sum / cnt;
            if (ms > 1e-12 && std::isfinite(rms)) {
                double sc = fmin_r(fmax_r(rms / ms, 0.05), 5.0) * 0.6;
                for (int k = 0; k < n_grid; k++) sdv[k] *= sc;
            }
        }
    }
    double chi = 0.0, bv = 0.0, ms = 0.0;
    for (int i = 0; i < nd; i++) ms += dm_[i];
    const double mean_mag = ms / nd;
    for (int i = 0; i < nd; i++) {
        double r = dm_[i] - pobs[i];
        double e2 = de_[i] * de_[i] + 1e-10;
        chi += r * r / e2;
        bv += (dm_[i] - mean_mag) * (dm_[i] - mean_mag) / e2;
    }
    const int pk = argmin_finite(pred, n_grid);
    const double t0 = tg[pk];
    double omin = INFINITY;
    for (double v : dm_) omin = fmin_r(omin, v);
    const double peak_mag = fmax_r(pred[pk], omin - 2.0);
    double t_last = -INFINITY;
    for (int i = 0; i < n; i++) t_last = fmax_r(t_last, t[i]);
    // fine grid (:503-554)
    double fine[11];
    {
        const double a = tg[0], b = tg[n_grid - 1];
        int nf = (int)fmax_r(b - a, 1.0) + 1;
        const double step = nf > 2000 ? (b - a) / 1999.0 : 1.0;
        nf = std::min(nf, 2000);
        std::vector<double> ft(nf), fp(nf);
        for (int i = 0; i < nf; i++) ft[i] = a + (double)i * step;
        for (int i = 0; i < nf; i++) fp[i] = gp.mean(ft[i]);
        curve_metrics(ft.data(), fp.data(), nf, argmin_finite(fp.data(), nf), fine);
    }
    // predictive features (:211-398)
    const double f_m1 = gp.mean(t_last - 1.0), f_0 = gp.mean(t_last), f_p1 = gp.mean(t_last + 1.0);
    const double f_p2 = gp.mean(t_last + 2.0), f_p3 = gp.mean(t_last + 3.0);
    double yw[6];
    for (int i = 0; i < 6; i++) yw[i] = gp.mean(t_last + (double)(i - 2));
    double sl = 0.0;
    int nsl = 0;
    for (int i = 0; i < 5; i++) {
        double d = yw[i + 1] - yw[i];
        if (std::isfinite(d)) { sl += d; nsl++; }
    }
    double snr = -INFINITY;
    for (int i = 0; i < nd; i++) snr = fmax_r(snr, std::fabs(dm_[i]) / de_[i]);
    double gs[10];
    grid_stats(tg.data(), pred, sdv, n_grid, gs);
    double pl_i, pl_c;
    power_law(tg.data(), pred, sdv, n_grid, pk, &pl_i, &pl_c);
    auto at = [&](double off) {
        double v = gp.mean(t0 + off);
        return std::isfinite(v) ? v : NAN;
    };
    double pre_rms, pre_sig;
    pre_peak(t, mg, er, n, t0, peak_mag, &pre_rms, &pre_sig);

    double* f = o.feat;
    f[0] = fine[4];
    f[1] = fine[5];
    f[2] = fine[6];
    f[3] = fine[0];
    f[4] = fine[1];
    f[5] = fine[2];
    f[6] = t0;
    f[7] = peak_mag;
    f[8] = chi / nd;
    f[9] = bv / nd;
    f[11] = fine[3];
    f[12] = fine[7];
    f[13] = fine[8];
    f[14] = fine[9];
    f[15] = fine[10];
    f[16] = nsl > 0 ? sl / nsl : NAN;
    f[17] = (f_p1 - f_0) / 1.0;
    f[18] = (f_p1 - 2.0 * f_0 + f_m1) / (1.0 * 1.0);
    f[19] = f_p1;
    f[20] = f_p2;
    f[21] = t0 - t_last;
    f[22] = (f_p3 - f_p2) / 1.0;
    f[23] = gs[0];
    f[24] = gs[1];
    f[25] = snr;
    f[26] = gs[2];
    f[27] = gs[3];
    f[28] = gs[4];
    f[29] = gs[5];
    f[30] = gs[6];
    f[31] = gs[7];
    f[32] = gs[8];
    f[33] = gs[9];
    f[34] = pl_i;
    f[35] = pl_c;
    f[36] = at(30.0);
    f[37] = at(60.0);
    f[38] = at(90.0);
    f[39] = von_neumann(t, mg, n);
    f[40] = pre_rms;
    f[41] = pre_sig;
    f[42] = monotonicity(pred, n_grid, pk);
    f[43] = local_maxima(pred, n_grid, 0.1);
    f[45] = fa;
    f[46] = fl;
    f[47] = stetson_k(mg, er, n);
    f[48] = excess_var(mg, er, n);
    f[49] = cusum(mg, n);
    for (int k = 0; k < ORC_NP_NFEAT; k++)
        if (!std::isfinite(f[k])) f[k] = NAN;
}

}  // namespace np
}  // namespace orc

extern "C" {

int orc_batch_nonparametric(int64_t n_bands, const int64_t* off, const double* t, const double* mag,
                            const double* err, const double* t_lo, const double* t_hi, int n_grid,
                            const double* amps, int n_amp, const double* ls_factors, int n_ls, int dense_max,
                            int max_sub, int n_inducing, int32_t* status, double* feat, double* pred,
                            double* stdv, double* dpred, double* state, int state_stride, int n_threads) {
    if (n_bands < 0 || n_grid < 2 || n_amp <= 0 || n_ls <= 0 || dense_max < 0 || max_sub < 1 || n_inducing < 1)
        return -1;
    const int cap = (state_stride - 4) / 2;
    if (cap < std::max(dense_max, max_sub)) return -1;
    orc::parallel_for(n_bands, n_threads, [&](int64_t b) {
        orc::np::BandOut o;
        const int64_t a = off[b];
        orc::np::band(t + a, mag + a, err + a, (int)(off[b + 1] - a), t_lo[b], t_hi[b], n_grid, amps, n_amp,
                      ls_factors, n_ls, dense_max, max_sub, n_inducing, o, pred + b * n_grid, stdv + b * n_grid,
                      dpred + b * n_grid, state + b * state_stride, cap);
        status[b] = o.status;
        for (int k = 0; k < ORC_NP_NFEAT; k++) feat[b * ORC_NP_NFEAT + k] = o.feat[k];
    });
    return 0;
}

int orc_np_curve_metrics(const double* t, const double* m, int n, int peak_idx, double* out11) {
    if (n < 0 || peak_idx < 0) return -1;
    orc::np::curve_metrics(t, m, n, peak_idx, out11);
    return 0;
}

int orc_np_raw_features(const double* t, const double* m, const double* e, int n, double t_peak, double peak_mag,
                        double* out6) {
    if (n < 0) return -1;
    out6[0] = orc::np::von_neumann(t, m, n);
    orc::np::pre_peak(t, m, e, n, t_peak, peak_mag, &out6[1], &out6[2]);
    out6[3] = orc::np::stetson_k(m, e, n);
    out6[4] = orc::np::excess_var(m, e, n);
    out6[5] = orc::np::cusum(m, n);
    return 0;
}

int orc_np_grid_features(const double* tg, const double* p, const double* sd, int n, int peak_idx, double* out14) {
    if (n < 1 || peak_idx < 0 || peak_idx >= n) return -1;
    orc::np::grid_stats(tg, p, sd, n, out14);
    orc::np::power_law(tg, p, sd, n, peak_idx, &out14[10], &out14[11]);
    out14[12] = orc::np::monotonicity(p, n, peak_idx);
    out14[13] = orc::np::local_maxima(p, n, 0.1);
    return 0;
}

int orc_sparse_gp(const double* t, const double* y, const double* nv, int n, double amp, double ls, int m,
                  const double* q, int nq, double* mean, double* stdv, double* nlml) {
    std::vector<double> tt(t, t + n), yy(y, y + n), vv(nv, nv + n);
    orc::np::Gp g = orc::np::sparse_fit(tt, yy, vv, amp, ls, m);
    if (!g.ok) return 1;
    for (int k = 0; k < nq; k++) {
        if (mean) mean[k] = g.mean(q[k]);
        if (stdv) stdv[k] = g.stdv(q[k]);
    }
    if (nlml) *nlml = orc::np::approx_nlml(g, tt, yy, vv);
    return 0;
}

}  // extern "C"
