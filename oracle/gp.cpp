// gp.cpp -- oracle restatement of the dense 1D Gaussian-process fits.
//
// TEST INFRASTRUCTURE ONLY (parity checker for the HIP product).
//
//   DenseGP (src/sparse_gp.rs:72-176): RBF kernel a exp(-d^2/(2 l^2)), noise
//   floor 1e-10 on the diagonal, row-major Cholesky-Crout, mean-centred
//   targets, alpha = K^-1 (y - ybar), predictive mean / std (var floored at
//   1e-10), train RMS.
//   batch_gp_fit, GPU path (cuda/gp.cu:124-367, src/gpu_cuda.rs:976-1143):
//   per band, skip n < 3; subsample m = min(n, cap) points at
//   idx_i = floor(i (n-1)/(m-1) + 1/2); y mean over the subsample; every
//   (amp, ls) candidate with ls >= 0.1 scored by the exact negative log
//   marginal likelihood 0.5 y^T alpha + sum ln L_ii + m/2 ln 2 pi (1e99 when
//   the Cholesky fails); the first strict minimum in amp-major order wins;
//   refit, predict mean / std on the query grid and the mean at every
//   observation of the band.
// Extensions used by the HIP kernels (same arithmetic): a train-RMS selection
// criterion (the CPU path's, src/nonparametric.rs:1318-1335), the CPU
// subsample rule of src/gp.rs:5-27, and the analytic time derivative of the
// predictive mean.
#include "lcf_oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace orc {

static const double PI = 3.14159265358979323846;

static inline double rbf(double a, double b, double amp, double i2) {
    double d = a - b;
    return amp * std::exp(-d * d * i2);
}

static bool cholesky(std::vector<double>& a, int n) {
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

static void solve_l(const std::vector<double>& l, const double* b, double* x, int n) {
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int j = 0; j < i; j++) s -= l[i * n + j] * x[j];
        x[i] = s / l[i * n + i];
    }
}

static void solve_lt(const std::vector<double>& l, const double* b, double* x, int n) {
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < n; j++) s -= l[j * n + i] * x[j];
        x[i] = s / l[i * n + i];
    }
}

struct Dense {
    std::vector<double> x, alpha, l;
    int n = 0;
    double amp = 0, i2 = 0, ymean = 0, nlml = 1e99;
    bool ok = false;
};

// DenseGP::fit (src/sparse_gp.rs:88-119) plus the NLML of cuda/gp.cu:91-99.
static Dense dense_fit(const double* t, const double* y, const double* nv, int n, double amp, double i2) {
    Dense g;
    g.n = n;
    g.amp = amp;
    g.i2 = i2;
    if (n == 0) return g;
    g.x.assign(t, t + n);
    double ym = 0.0;
    for (int i = 0; i < n; i++) ym += y[i];
    ym /= n;
    g.ymean = ym;
    std::vector<double> k(n * n);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j <= i; j++) {
            double v = rbf(t[i], t[j], amp, i2);
            k[i * n + j] = v;
            k[j * n + i] = v;
        }
        k[i * n + i] += std::max(nv[i], 1e-10);
    }
    if (!cholesky(k, n)) return g;
    std::vector<double> yc(n), tmp(n);
    for (int i = 0; i < n; i++) yc[i] = y[i] - ym;
    g.alpha.resize(n);
    solve_l(k, yc.data(), tmp.data(), n);
    solve_lt(k, tmp.data(), g.alpha.data(), n);
    double fit = 0.0, ld = 0.0;
    for (int i = 0; i < n; i++) { fit += yc[i] * g.alpha[i]; ld += std::log(k[i * n + i]); }
    double nl = 0.5 * fit + ld + 0.5 * n * std::log(2.0 * PI);
    g.nlml = std::isfinite(nl) ? nl : 1e99;
    g.l = std::move(k);
    g.ok = true;
    return g;
}

static double dense_mean(const Dense& g, double t) {
    double s = 0.0;
    for (int i = 0; i < g.n; i++) s += rbf(t, g.x[i], g.amp, g.i2) * g.alpha[i];
    return s + g.ymean;
}

static double dense_dmean(const Dense& g, double t) {
    double s = 0.0;
    for (int i = 0; i < g.n; i++) s += -2.0 * g.i2 * (t - g.x[i]) * rbf(t, g.x[i], g.amp, g.i2) * g.alpha[i];
    return s;
}

static double dense_std(const Dense& g, double t) {
    std::vector<double> ks(g.n), v(g.n);
    for (int i = 0; i < g.n; i++) ks[i] = rbf(t, g.x[i], g.amp, g.i2);
    solve_l(g.l, ks.data(), v.data(), g.n);
    double vtv = 0.0;
    for (double z : v) vtv += z * z;
    return std::sqrt(std::max(g.amp - vtv, 1e-10));
}

}  // namespace orc

extern "C" {

int orc_dense_gp(const double* t, const double* y, const double* nv, int n, double amp, double ls,
                 const double* q, int nq, double* mean, double* stdv, double* dmean, double* nlml,
                 double* train_rms) {
    if (n <= 0) return -1;
    orc::Dense g = orc::dense_fit(t, y, nv, n, amp, 0.5 / (ls * ls));
    if (!g.ok) return 1;
    for (int k = 0; k < nq; k++) {
        if (mean) mean[k] = orc::dense_mean(g, q[k]);
        if (stdv) stdv[k] = orc::dense_std(g, q[k]);
        if (dmean) dmean[k] = orc::dense_dmean(g, q[k]);
    }
    if (nlml) *nlml = g.nlml;
    if (train_rms) {
        double rss = 0.0;
        for (int i = 0; i < n; i++) {
            double r = orc::dense_mean(g, t[i]) - y[i];
            rss += r * r;
        }
        *train_rms = std::sqrt(rss / n);
    }
    return 0;
}

int orc_batch_gp_fit(int64_t n_bands, const int64_t* off, const double* t, const double* y,
                     const double* nv, const double* q, int nq, const double* amps, int n_amp,
                     const double* lss, int n_ls, int max_sub, int subsample_rule, int criterion,
                     double* state, int state_stride, double* pred, double* stdv, double* dpred,
                     double* pred_obs, int n_threads) {
    if (n_bands < 0 || nq < 0 || n_amp <= 0 || n_ls <= 0 || max_sub < 3 || state_stride < 2 * max_sub + 4)
        return -1;
    orc::parallel_for(n_bands, n_threads, [&](int64_t b) {
        const int64_t a = off[b];
        const int n = (int)(off[b + 1] - a);
        double* st = state + b * state_stride;
        for (int i = 0; i < state_stride; i++) st[i] = 0.0;
        for (int k = 0; k < nq; k++) {
            pred[b * nq + k] = NAN;
            stdv[b * nq + k] = NAN;
            if (dpred) dpred[b * nq + k] = NAN;
        }
        for (int i = 0; i < n; i++) pred_obs[a + i] = NAN;
        if (n < 3) return;
        int m = std::min(n, max_sub);
        std::vector<double> st_t(m), st_y(m), st_v(m);
        for (int i = 0; i < m; i++) {
            int idx;
            if (n <= m) idx = i;
            else if (subsample_rule == ORC_GP_SUBSAMPLE_GPU) {
                double step = (double)(n - 1) / (double)(m - 1);
                idx = std::min((int)(i * step + 0.5), n - 1);
            } else {
                idx = std::min((int)(((double)i + 0.5) * n / m), n - 1);  // src/gp.rs:15-19
            }
            st_t[i] = t[a + idx];
            st_y[i] = y[a + idx];
            st_v[i] = nv[a + idx];
        }
        double best = 1e99;
        int best_c = 0;
        bool any = false;
        for (int c = 0; c < n_amp * n_ls; c++) {
            double amp = amps[c / n_ls], ls = lss[c % n_ls];
            double score = 1e99;
            if (ls >= 0.1) {
                orc::Dense g = orc::dense_fit(st_t.data(), st_y.data(), st_v.data(), m, amp, 0.5 / (ls * ls));
                if (g.ok) {
                    if (criterion == ORC_GP_SELECT_NLML) score = g.nlml;
                    else {
                        double rss = 0.0;
                        for (int i = 0; i < m; i++) {
                            double r = orc::dense_mean(g, st_t[i]) - st_y[i];
                            rss += r * r;
                        }
                        score = std::sqrt(rss / m);
                        if (!std::isfinite(score)) score = 1e99;
                    }
                }
            }
            if (c == 0) any = ls >= 0.1;
            if (score < best) { best = score; best_c = c; any = true; }
        }
        if (!any) return;
        double amp = amps[best_c / n_ls], ls = lss[best_c % n_ls];
        orc::Dense g = orc::dense_fit(st_t.data(), st_y.data(), st_v.data(), m, amp, 0.5 / (ls * ls));
        if (!g.ok) return;
        for (int i = 0; i < m; i++) { st[i] = g.alpha[i]; st[max_sub + i] = st_t[i]; }
        st[2 * max_sub] = amp;
        st[2 * max_sub + 1] = g.i2;
        st[2 * max_sub + 2] = g.ymean;
        st[2 * max_sub + 3] = (double)m;
        for (int k = 0; k < nq; k++) {
            pred[b * nq + k] = orc::dense_mean(g, q[k]);
            stdv[b * nq + k] = orc::dense_std(g, q[k]);
            if (dpred) dpred[b * nq + k] = orc::dense_dmean(g, q[k]);
        }
        for (int i = 0; i < n; i++) pred_obs[a + i] = orc::dense_mean(g, t[a + i]);
    });
    return 0;
}

}  // extern "C"
