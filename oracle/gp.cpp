// gp.cpp -- oracle restatement of the dense Gaussian-process fits.
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
//   batch_gp_2d (cuda/gp2d.cu:123-359): the same with the anisotropic kernel
//   a exp(-dt^2/(2 lt^2) - dw^2/(2 lw^2)) over (time, log10 wavelength),
//   candidates amp-major then lt then lw, valid when lt >= 0.1 and lw >= 0.001.
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

struct Kern {
    double amp, it, iw;
    double operator()(double t1, double w1, double t2, double w2) const {
        double dt = t1 - t2, dw = w1 - w2;
        return amp * std::exp(-dt * dt * it - dw * dw * iw);
    }
};

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
    std::vector<double> x, w, alpha, l;
    int n = 0;
    Kern k{0, 0, 0};
    double ymean = 0, nlml = 1e99;
    bool ok = false;
};

static Dense dense_fit(const double* t, const double* w, const double* y, const double* nv, int n, Kern k) {
    Dense g;
    g.n = n;
    g.k = k;
    if (n == 0) return g;
    g.x.assign(t, t + n);
    g.w.assign(n, 0.0);
    if (w) g.w.assign(w, w + n);
    double ym = 0.0;
    for (int i = 0; i < n; i++) ym += y[i];
    ym /= n;
    g.ymean = ym;
    std::vector<double> K(n * n);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j <= i; j++) {
            double v = k(g.x[i], g.w[i], g.x[j], g.w[j]);
            K[i * n + j] = v;
            K[j * n + i] = v;
        }
        K[i * n + i] += std::max(nv[i], 1e-10);
    }
    if (!cholesky(K, n)) return g;
    std::vector<double> yc(n), tmp(n);
    for (int i = 0; i < n; i++) yc[i] = y[i] - ym;
    g.alpha.resize(n);
    solve_l(K, yc.data(), tmp.data(), n);
    solve_lt(K, tmp.data(), g.alpha.data(), n);
    double fit = 0.0, ld = 0.0;
    for (int i = 0; i < n; i++) { fit += yc[i] * g.alpha[i]; ld += std::log(K[i * n + i]); }
    double nl = 0.5 * fit + ld + 0.5 * n * std::log(2.0 * PI);
    g.nlml = std::isfinite(nl) ? nl : 1e99;
    g.l = std::move(K);
    g.ok = true;
    return g;
}

static double dense_mean(const Dense& g, double t, double w) {
    double s = 0.0;
    for (int i = 0; i < g.n; i++) s += g.k(t, w, g.x[i], g.w[i]) * g.alpha[i];
    return s + g.ymean;
}

static double dense_dmean(const Dense& g, double t, double w) {
    double s = 0.0;
    for (int i = 0; i < g.n; i++) s += -2.0 * g.k.it * (t - g.x[i]) * g.k(t, w, g.x[i], g.w[i]) * g.alpha[i];
    return s;
}

static double dense_std(const Dense& g, double t, double w) {
    std::vector<double> ks(g.n), v(g.n);
    for (int i = 0; i < g.n; i++) ks[i] = g.k(t, w, g.x[i], g.w[i]);
    solve_l(g.l, ks.data(), v.data(), g.n);
    double vtv = 0.0;
    for (double z : v) vtv += z * z;
    return std::sqrt(std::max(g.k.amp - vtv, 1e-10));
}

// Shared batch driver for the 1D (w == null) and 2D fits.
static int batch_fit(int64_t n_bands, const int64_t* off, const double* t, const double* w, const double* y,
                     const double* nv, const double* q, const double* qw, int nq, const double* amps, int n_amp,
                     const double* lss, int n_ls, const double* lws, int n_lw, int max_sub, int rule,
                     int criterion, double* state, int stride, double* pred, double* stdv, double* dpred,
                     double* pred_obs, int n_threads) {
    const bool two_d = w != nullptr;
    const int nw = two_d ? n_lw : 1;
    if (n_bands < 0 || nq < 0 || n_amp <= 0 || n_ls <= 0 || nw <= 0 || max_sub < 3 ||
        stride < (two_d ? 3 * max_sub + 5 : 2 * max_sub + 4))
        return -1;
    parallel_for(n_bands, n_threads, [&](int64_t b) {
        const int64_t a = off[b];
        const int n = (int)(off[b + 1] - a);
        double* st = state + b * stride;
        for (int i = 0; i < stride; i++) st[i] = 0.0;
        for (int k = 0; k < nq; k++) {
            pred[b * nq + k] = NAN;
            stdv[b * nq + k] = NAN;
            if (dpred) dpred[b * nq + k] = NAN;
        }
        for (int i = 0; i < n; i++) pred_obs[a + i] = NAN;
        if (n < 3) return;
        int m = std::min(n, max_sub);
        std::vector<double> st_t(m), st_w(m, 0.0), st_y(m), st_v(m);
        for (int i = 0; i < m; i++) {
            int idx;
            if (n <= m) idx = i;
            else if (rule == ORC_GP_SUBSAMPLE_GPU) {
                double step = (double)(n - 1) / (double)(m - 1);
                idx = std::min((int)(i * step + 0.5), n - 1);
            } else {
                idx = std::min((int)(((double)i + 0.5) * ((double)n / (double)m)), n - 1);  // src/gp.rs:15-19
            }
            st_t[i] = t[a + idx];
            if (two_d) st_w[i] = w[a + idx];
            st_y[i] = y[a + idx];
            st_v[i] = nv[a + idx];
        }
        auto cand = [&](int c, bool& valid) {
            int ia = c / (n_ls * nw), il = (c / nw) % n_ls, iw = c % nw;
            Kern k{amps[ia], 0.5 / (lss[il] * lss[il]), 0.0};
            valid = lss[il] >= 0.1;
            if (two_d) {
                k.iw = 0.5 / (lws[iw] * lws[iw]);
                valid = valid && lws[iw] >= 0.001;
            }
            return k;
        };
        double best = 1e99;
        int best_c = 0;
        bool any = false;
        for (int c = 0; c < n_amp * n_ls * nw; c++) {
            bool valid;
            Kern k = cand(c, valid);
            double score = 1e99;
            if (valid) {
                Dense g = dense_fit(st_t.data(), st_w.data(), st_y.data(), st_v.data(), m, k);
                if (g.ok) {
                    if (criterion == ORC_GP_SELECT_NLML) score = g.nlml;
                    else {
                        double rss = 0.0;
                        for (int i = 0; i < m; i++) {
                            double r = dense_mean(g, st_t[i], st_w[i]) - st_y[i];
                            rss += r * r;
                        }
                        score = std::sqrt(rss / m);
                        if (!std::isfinite(score)) score = 1e99;
                    }
                }
            }
            if (c == 0) any = valid;
            if (score < best) { best = score; best_c = c; any = true; }
        }
        if (!any) return;
        // the CPU dense path keeps no fit when no candidate has a finite score
        // (src/nonparametric.rs:1318-1339)
        if (criterion == ORC_GP_SELECT_TRAIN_RMS && !(best < 1e99)) return;
        bool valid;
        Kern k = cand(best_c, valid);
        Dense g = dense_fit(st_t.data(), st_w.data(), st_y.data(), st_v.data(), m, k);
        if (!g.ok) return;
        for (int i = 0; i < m; i++) {
            st[i] = g.alpha[i];
            st[max_sub + i] = st_t[i];
            if (two_d) st[2 * max_sub + i] = st_w[i];
        }
        int o = two_d ? 3 * max_sub : 2 * max_sub;
        st[o++] = k.amp;
        st[o++] = k.it;
        if (two_d) st[o++] = k.iw;
        st[o++] = g.ymean;
        st[o] = (double)m;
        for (int kq = 0; kq < nq; kq++) {
            double wq = two_d ? qw[kq] : 0.0;
            pred[b * nq + kq] = dense_mean(g, q[kq], wq);
            stdv[b * nq + kq] = dense_std(g, q[kq], wq);
            if (dpred) dpred[b * nq + kq] = dense_dmean(g, q[kq], wq);
        }
        for (int i = 0; i < n; i++) pred_obs[a + i] = dense_mean(g, t[a + i], two_d ? w[a + i] : 0.0);
    });
    return 0;
}

}  // namespace orc

extern "C" {

int orc_dense_gp(const double* t, const double* y, const double* nv, int n, double amp, double ls,
                 const double* q, int nq, double* mean, double* stdv, double* dmean, double* nlml,
                 double* train_rms) {
    if (n <= 0) return -1;
    orc::Dense g = orc::dense_fit(t, nullptr, y, nv, n, orc::Kern{amp, 0.5 / (ls * ls), 0.0});
    if (!g.ok) return 1;
    for (int k = 0; k < nq; k++) {
        if (mean) mean[k] = orc::dense_mean(g, q[k], 0.0);
        if (stdv) stdv[k] = orc::dense_std(g, q[k], 0.0);
        if (dmean) dmean[k] = orc::dense_dmean(g, q[k], 0.0);
    }
    if (nlml) *nlml = g.nlml;
    if (train_rms) {
        double rss = 0.0;
        for (int i = 0; i < n; i++) {
            double r = orc::dense_mean(g, t[i], 0.0) - y[i];
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
    return orc::batch_fit(n_bands, off, t, nullptr, y, nv, q, nullptr, nq, amps, n_amp, lss, n_ls, nullptr, 0,
                          max_sub, subsample_rule, criterion, state, state_stride, pred, stdv, dpred, pred_obs,
                          n_threads);
}

int orc_batch_gp2d_fit(int64_t n_src, const int64_t* off, const double* t, const double* w, const double* y,
                       const double* nv, const double* qt, const double* qw, int nq, const double* amps,
                       int n_amp, const double* lst, int n_lst, const double* lsw, int n_lsw, int max_sub,
                       int subsample_rule, int criterion, double* state, int state_stride, double* pred,
                       double* stdv, double* dpred, double* pred_obs, int n_threads) {
    if (!w || !qw) return -1;
    return orc::batch_fit(n_src, off, t, w, y, nv, qt, qw, nq, amps, n_amp, lst, n_lst, lsw, n_lsw, max_sub,
                          subsample_rule, criterion, state, state_stride, pred, stdv, dpred, pred_obs, n_threads);
}


// fit_gp_2d (src/gp2d.rs:336-439) on prepared training arrays: subsample of
// max_n = 200 points at floor(i n / 200); per-source grid amp {0.3, 1, 3, 10}
// x lt {dur/4, dur/8, dur/16} x lw {2 wr, wr, wr/2} (dur = max(t range, 1),
// wr = max(log10-wavelength range, 0.01) of the subsample), first strict
// minimum of the train RMS; then predict_with_std at the source's queries
// (src/gp2d.rs:181-209).  hyp[3] = amp, lt, lw (NaN and n_train 0 when no
// candidate factorises).
int orc_batch_fit_gp_2d(int64_t n_src, const int64_t* off, const double* t, const double* w, const double* y,
                        const double* nv, const double* qt, const double* qw, int nq, double* hyp,
                        double* train_rms, int32_t* n_train, double* pred, double* stdv, int n_threads) {
    orc::parallel_for(n_src, n_threads, [&](int64_t s) {
        const int64_t a = off[s];
        const int n = (int)(off[s + 1] - a);
        hyp[s * 3] = hyp[s * 3 + 1] = hyp[s * 3 + 2] = NAN;
        train_rms[s] = NAN;
        n_train[s] = 0;
        for (int k = 0; k < nq; k++) pred[s * nq + k] = stdv[s * nq + k] = NAN;
        if (n <= 0) return;
        const int max_n = 200;
        std::vector<double> st, sw, sy, sv;
        if (n <= max_n) {
            st.assign(t + a, t + a + n);
            sw.assign(w + a, w + a + n);
            sy.assign(y + a, y + a + n);
            sv.assign(nv + a, nv + a + n);
        } else {
            const double step = (double)n / (double)max_n;
            for (int i = 0; i < max_n; i++) {
                const int idx = (int)((double)i * step);
                st.push_back(t[a + idx]);
                sw.push_back(w[a + idx]);
                sy.push_back(y[a + idx]);
                sv.push_back(nv[a + idx]);
            }
        }
        const int m = (int)st.size();
        double t0 = INFINITY, t1 = -INFINITY, w0 = INFINITY, w1 = -INFINITY;
        for (int i = 0; i < m; i++) {
            t0 = std::fmin(t0, st[i]);
            t1 = std::fmax(t1, st[i]);
            w0 = std::fmin(w0, sw[i]);
            w1 = std::fmax(w1, sw[i]);
        }
        const double dur = std::fmax(t1 - t0, 1.0), wr = std::fmax(w1 - w0, 0.01);
        const double amps[4] = {0.3, 1.0, 3.0, 10.0};
        const double lts[3] = {dur / 4.0, dur / 8.0, dur / 16.0};
        const double lws[3] = {wr * 2.0, wr, wr / 2.0};
        double best = INFINITY, bh[3] = {0, 0, 0};
        bool found = false;
        orc::Dense bg;
        for (double amp : amps)
            for (double lt : lts)
                for (double lw : lws) {
                    orc::Kern k{amp, 0.5 / (lt * lt), 0.5 / (lw * lw)};
                    orc::Dense g = orc::dense_fit(st.data(), sw.data(), sy.data(), sv.data(), m, k);
                    if (!g.ok) continue;
                    double rss = 0.0;
                    for (int i = 0; i < m; i++) {
                        const double r = orc::dense_mean(g, st[i], sw[i]) - sy[i];
                        rss += r * r;
                    }
                    const double rms = std::sqrt(rss / (double)m);
                    if (rms < best) {
                        best = rms;
                        bh[0] = amp;
                        bh[1] = lt;
                        bh[2] = lw;
                        bg = g;
                        found = true;
                    }
                }
        if (!found) return;
        for (int k = 0; k < 3; k++) hyp[s * 3 + k] = bh[k];
        train_rms[s] = best;
        n_train[s] = m;
        for (int k = 0; k < nq; k++) {
            pred[s * nq + k] = orc::dense_mean(bg, qt[s * nq + k], qw[s * nq + k]);
            stdv[s * nq + k] = orc::dense_std(bg, qt[s * nq + k], qw[s * nq + k]);
        }
    });
    return 0;
}

}  // extern "C"
