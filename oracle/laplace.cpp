// laplace.cpp -- oracle restatement of the Laplace posterior approximation.
//
// TEST INFRASTRUCTURE ONLY (parity checker for the HIP product).
//
// laplace_fit (src/parametric.rs:2623-2693): Hessian of the un-normalised
// negative log posterior (neg_log_posterior, :2464-2511: Gaussian / censored
// likelihood with the 1/2 and ln 2 pi terms, prior_params priors, MetzgerKN
// rescaled to its peak) over the shape parameters with sigma_x held at its MAP
// value, by the 4-point central difference with h_i = 1e-4 max(|theta_i|, 1);
// cyclic-max Jacobi eigendecomposition (symmetric_eigen, :2517-2587);
// eigenvalues floored at 1e-6; log sigma_i = 1/2 ln (H^-1)_ii
// (log_sigma_from_hessian, :2590-2613); sigma_x's slot is ln(prior width);
// elbo = -f(MAP).
// Variant ORC_LAPLACE_GRAD_FD ("forward over adjoint"): H_ij from central
// differences of the analytic gradient, symmetrised.
#include "lcf_oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace orc {

static const double PI = 3.14159265358979323846;

double neg_log_posterior(int m, const double* p, const double* t, const double* f, const double* v,
                         const int32_t* up, const double* uf, int64_t n) {
    int np = n_params(m), se = np - 1;
    double se2 = std::exp(p[se]);
    se2 *= se2;
    std::vector<double> pred(std::max<int64_t>(n, 1));
    eval_times(m, p, t, n, pred.data());
    if (m == 7) {
        double mx = -INFINITY;
        for (int64_t i = 0; i < n; i++)
            if (!up[i]) mx = std::max(mx, pred[i]);
        if (mx > 1e-10 && std::isfinite(mx)) {
            double sc = std::min(std::max(1.0 / mx, 0.1), 10.0);
            for (int64_t i = 0; i < n; i++) pred[i] *= sc;
        }
    }
    double nll = 0.0;
    for (int64_t i = 0; i < n; i++) {
        if (!std::isfinite(pred[i])) return 1e99;
        double tv = v[i] + se2;
        if (up[i]) nll -= log_normal_cdf((uf[i] - pred[i]) / std::sqrt(tv));
        else {
            double d = pred[i] - f[i];
            nll += 0.5 * (d * d / tv + std::log(2.0 * PI * tv));
        }
    }
    double c[7], w[7];
    prior_params(m, c, w);
    double lp = 0.0;
    for (int j = 0; j < np; j++) lp += 0.5 * (p[j] - c[j]) * (p[j] - c[j]) / (w[j] * w[j]);
    return nll + lp;
}

// Analytic gradient of neg_log_posterior (closed-form models only).
static void nlp_grad(int m, const double* p, const double* t, const double* f, const double* v,
                     const int32_t* up, const double* uf, int64_t n, double* g) {
    int np = n_params(m), se = np - 1;
    double se2 = std::exp(p[se]);
    se2 *= se2;
    double mg[7];
    for (int j = 0; j < np; j++) g[j] = 0.0;
    for (int64_t i = 0; i < n; i++) {
        double pred = eval_scalar(m, p, t[i]);
        if (!std::isfinite(pred)) continue;
        eval_grad(m, p, t[i], mg);
        double tv = v[i] + se2, st = std::sqrt(tv);
        if (up[i]) {
            double z = (uf[i] - pred) / st;
            double phi = std::exp(-0.5 * z * z) / std::sqrt(2.0 * PI);
            double cdf = std::max(0.5 * (1.0 + erf_approx(z * 0.70710678118654752440)), 1e-300);
            double r = phi / cdf;
            for (int j = 0; j < np; j++)
                if (j != se) g[j] += r / st * mg[j];
            g[se] += r * (uf[i] - pred) * se2 / (st * tv);
        } else {
            double d = pred - f[i];
            for (int j = 0; j < np; j++)
                if (j != se) g[j] += d / tv * mg[j];
            g[se] += (1.0 / tv - d * d / (tv * tv)) * se2;
        }
    }
    double c[7], w[7];
    prior_params(m, c, w);
    for (int j = 0; j < np; j++) g[j] += (p[j] - c[j]) / (w[j] * w[j]);
}

static void jacobi_eigen(const double* mat, int n, double* evals, double* evecs) {
    std::vector<double> a(mat, mat + n * n);
    for (int i = 0; i < n * n; i++) evecs[i] = 0.0;
    for (int i = 0; i < n; i++) evecs[i * n + i] = 1.0;
    const int max_iter = 100 * n * n;
    for (int it = 0; it < max_iter; it++) {
        double mv = 0.0;
        int p = 0, q = 1;
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++)
                if (std::fabs(a[i * n + j]) > mv) { mv = std::fabs(a[i * n + j]); p = i; q = j; }
        if (mv < 1e-15) break;
        double app = a[p * n + p], aqq = a[q * n + q], apq = a[p * n + q];
        double th = std::fabs(app - aqq) < 1e-30 ? PI / 4 : 0.5 * std::atan(2.0 * apq / (app - aqq));
        double c = std::cos(th), s = std::sin(th);
        for (int i = 0; i < n; i++) {
            double aip = a[i * n + p], aiq = a[i * n + q];
            a[i * n + p] = c * aip + s * aiq;
            a[i * n + q] = -s * aip + c * aiq;
        }
        for (int j = 0; j < n; j++) {
            double apj = a[p * n + j], aqj = a[q * n + j];
            a[p * n + j] = c * apj + s * aqj;
            a[q * n + j] = -s * apj + c * aqj;
        }
        a[p * n + p] = c * c * app + 2.0 * c * s * apq + s * s * aqq;
        a[q * n + q] = s * s * app - 2.0 * c * s * apq + c * c * aqq;
        a[p * n + q] = 0.0;
        a[q * n + p] = 0.0;
        for (int i = 0; i < n; i++) {
            double vip = evecs[i * n + p], viq = evecs[i * n + q];
            evecs[i * n + p] = c * vip + s * viq;
            evecs[i * n + q] = -s * vip + c * viq;
        }
    }
    for (int i = 0; i < n; i++) evals[i] = a[i * n + i];
}

static void log_sigma_from_hessian(const double* h, int n, double* out) {
    double ev[7], vec[49];
    jacobi_eigen(h, n, ev, vec);
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int k = 0; k < n; k++) s += vec[i * n + k] * vec[i * n + k] / std::max(ev[k], 1e-6);
        out[i] = 0.5 * std::log(std::max(s, 1e-30));
    }
}

}  // namespace orc

extern "C" {

double orc_neg_log_posterior(int m, const double* p, const double* t, const double* f, const double* v,
                             const int32_t* up, const double* uf, int64_t n) {
    return orc::neg_log_posterior(m, p, t, f, v, up, uf, n);
}

int orc_symmetric_eigen(const double* mat, int n, double* evals, double* evecs) {
    if (n <= 0 || n > 7) return -1;
    orc::jacobi_eigen(mat, n, evals, evecs);
    return 0;
}

int orc_laplace_fit(int m, const double* map, const double* t, const double* f, const double* v,
                    const int32_t* up, const double* uf, int64_t n, int mode, double* log_sigma,
                    double* elbo, double* hessian) {
    const int np = orc::n_params(m);
    if (np < 0 || (mode == ORC_LAPLACE_GRAD_FD && m == 7)) return -1;
    const int se = np - 1, ns = np - 1;
    double f0 = orc::neg_log_posterior(m, map, t, f, v, up, uf, n);
    double H[36];
    double pp[7], pm[7], mp[7], mm[7];
    if (mode == ORC_LAPLACE_FD4) {
        for (int i = 0; i < ns; i++) {
            double hi = 1e-4 * std::max(std::fabs(map[i]), 1.0);
            for (int j = i; j < ns; j++) {
                double hj = 1e-4 * std::max(std::fabs(map[j]), 1.0);
                for (int k = 0; k < np; k++) pp[k] = pm[k] = mp[k] = mm[k] = map[k];
                pp[i] += hi; pp[j] += hj;
                pm[i] += hi; pm[j] -= hj;
                mp[i] -= hi; mp[j] += hj;
                mm[i] -= hi; mm[j] -= hj;
                double hij = (orc::neg_log_posterior(m, pp, t, f, v, up, uf, n) -
                              orc::neg_log_posterior(m, pm, t, f, v, up, uf, n) -
                              orc::neg_log_posterior(m, mp, t, f, v, up, uf, n) +
                              orc::neg_log_posterior(m, mm, t, f, v, up, uf, n)) / (4.0 * hi * hj);
                H[i * ns + j] = hij;
                H[j * ns + i] = hij;
            }
        }
    } else {
        double gp[7], gm[7], G[36];
        for (int i = 0; i < ns; i++) {
            double hi = 1e-4 * std::max(std::fabs(map[i]), 1.0);
            for (int k = 0; k < np; k++) pp[k] = mm[k] = map[k];
            pp[i] += hi;
            mm[i] -= hi;
            orc::nlp_grad(m, pp, t, f, v, up, uf, n, gp);
            orc::nlp_grad(m, mm, t, f, v, up, uf, n, gm);
            for (int j = 0; j < ns; j++) G[i * ns + j] = (gp[j] - gm[j]) / (2.0 * hi);
        }
        for (int i = 0; i < ns; i++)
            for (int j = 0; j < ns; j++) H[i * ns + j] = 0.5 * (G[i * ns + j] + G[j * ns + i]);
    }
    (void)se;
    double ls[6];
    orc::log_sigma_from_hessian(H, ns, ls);
    double c[7], w[7];
    orc::prior_params(m, c, w);
    for (int i = 0; i < ns; i++) log_sigma[i] = ls[i];
    log_sigma[se] = std::log(w[se]);
    *elbo = -f0;
    if (hessian)
        for (int i = 0; i < ns * ns; i++) hessian[i] = H[i];
    return 0;
}

}  // extern "C"
