// finalize.cpp -- oracle restatement of the host post-processing that turns
// GPU PSO + SVI/Laplace fits into parametric band results.
//
// TEST INFRASTRUCTURE ONLY (parity checker for the HIP product).
//
//   * profile_t0_refine (src/parametric.rs:2788-3008): 25-point t0 grid on
//     [t_first - 30, t_peak + 5]; at each point the amplitude (and, for Bazin
//     and TDE, the baseline) is refitted by weighted least squares on the
//     detections, MetzgerKN predictions are rescaled to their peak, and the
//     Gaussian / censored log-likelihood is evaluated; the best t0 replaces
//     mu[t0], and the Delta LL = 0.5 width of the un-refitted profile, walked
//     outwards from the best point until the first grid point below the
//     threshold, gives log sigma[t0] (half-width, at least 0.1);
//   * profile_refine_all (:2705-2786): for every parameter except sigma_x, a
//     15-point scan over the CPU PSO bounds with the others at mu; sigma =
//     max(exp(log sigma), half-width of the points with LL >= LL(mu) - 0.5
//     (at least 1e-6), population-prior width);
//   * the magnitude-space chi2 of finalize_parametric_with_gpu_svi
//     (:3366-3393): SVI predictions unless fewer than n/2 are positive (then
//     the PSO parameters), residual mag_obs - (-2.5 log10(pred * peak) + 23.9)
//     over sigma_m = 1.0857 flux_err / flux, mean over the terms with positive
//     predicted and observed flux (ORC_FINALIZE_SVI_CHI2: from mu only, as
//     fit_parametric_multiband does, :4069-4085).
// Every sum runs over the observations in order, as the reference does.
#include "lcf_oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace orc {

namespace {

const double PI = 3.14159265358979323846;
// SviModel::t0_idx (src/parametric.rs:152-163), by model id
const int T0_IDX[8] = {2, 3, 2, 1, 1, 1, 1, 3};

struct BandView {
    const double* t;
    const double* f;
    const double* v;
    const int32_t* up;
    const double* uf;
    const double* fe;
    const double* mag;
    double peak;
    int64_t n;
};

// Gaussian / censored log-likelihood skipping non-finite predictions
// (src/parametric.rs:2717-2726).
double band_ll(const BandView& b, const double* pred, double log_sx) {
    double sx = std::exp(log_sx);
    double sx2 = sx * sx;
    double ll = 0.0;
    for (int64_t i = 0; i < b.n; i++) {
        if (!std::isfinite(pred[i])) continue;
        double tv = b.v[i] + sx2;
        if (b.up[i]) {
            ll += log_normal_cdf((b.uf[i] - pred[i]) / std::sqrt(tv));
        } else {
            double r = b.f[i] - pred[i];
            ll += -0.5 * (r * r / tv + std::log(2.0 * PI * tv));
        }
    }
    return ll;
}

double ll_at(int m, const BandView& b, const double* p, std::vector<double>& buf) {
    eval_times(m, p, b.t, b.n, buf.data());
    return band_ll(b, buf.data(), p[n_params(m) - 1]);
}

void profile_t0_refine(int m, double* mu, double* ls, const BandView& b) {
    const int np = n_params(m), se = np - 1, t0i = T0_IDX[m];
    const int64_t n = b.n;
    double t_first = INFINITY;
    for (int64_t i = 0; i < n; i++) t_first = std::fmin(t_first, b.t[i]);
    int64_t pk = 0;
    double pv = -INFINITY;
    for (int64_t i = 0; i < n; i++)
        if (b.f[i] > pv) { pv = b.f[i]; pk = i; }
    const double lo = t_first - 30.0, hi = b.t[pk] + 5.0;
    if (lo >= hi) return;
    const int NG = 25;
    auto grid = [&](int g) { return lo + (hi - lo) * (double)g / (double)(NG - 1); };
    double sx = std::exp(mu[se]);
    const double sx2 = sx * sx;
    const bool has_b = (m == 0 || m == 2);
    std::vector<double> sh(std::max<int64_t>(n, 1)), pr(std::max<int64_t>(n, 1));
    double best_t0 = mu[t0i], best_ll = -INFINITY;
    double p[7];
    for (int g = 0; g < NG; g++) {
        const double t0 = grid(g);
        for (int j = 0; j < np; j++) p[j] = mu[j];
        p[t0i] = t0;
        const double sa = p[0];
        p[0] = 0.0;
        double sb = 0.0;
        if (has_b) { sb = p[1]; p[1] = 0.0; }
        eval_times(m, p, b.t, n, sh.data());
        if (has_b) {
            double sw = 0, sy = 0, sf = 0, syf = 0, sff = 0;
            for (int64_t i = 0; i < n; i++) {
                if (b.up[i] || !std::isfinite(sh[i])) continue;
                double w = 1.0 / (b.v[i] + sx2);
                sw += w;
                sy += w * b.f[i];
                sf += w * sh[i];
                syf += w * b.f[i] * sh[i];
                sff += w * sh[i] * sh[i];
            }
            double det = sw * sff - sf * sf;
            p[0] = sa;
            p[1] = sb;
            if (std::fabs(det) > 1e-20) {
                double a = (sw * syf - sf * sy) / det;
                double bo = (sff * sy - sf * syf) / det;
                if (a > 1e-10) { p[0] = std::log(a); p[1] = bo; }
            }
        } else {
            double num = 0, den = 0;
            for (int64_t i = 0; i < n; i++) {
                if (b.up[i] || !std::isfinite(sh[i])) continue;
                double w = 1.0 / (b.v[i] + sx2);
                num += w * b.f[i] * sh[i];
                den += w * sh[i] * sh[i];
            }
            p[0] = (den > 1e-20 && num / den > 1e-10) ? std::log(num / den) : sa;
        }
        eval_times(m, p, b.t, n, pr.data());
        if (m == 7) {
            double mx = -INFINITY;
            for (int64_t i = 0; i < n; i++)
                if (!b.up[i]) mx = std::fmax(mx, pr[i]);
            if (mx > 1e-10 && std::isfinite(mx)) {
                double sc = std::min(std::max(1.0 / mx, 0.1), 10.0);
                for (int64_t i = 0; i < n; i++) pr[i] *= sc;
            }
        }
        double ll = band_ll(b, pr.data(), p[se]);
        if (ll > best_ll) { best_ll = ll; best_t0 = t0; }
    }
    const double thr = best_ll - 0.5;
    double lo1 = best_t0, hi1 = best_t0;
    for (int g = NG - 1; g >= 0; g--) {
        const double t0 = grid(g);
        if (t0 >= best_t0) continue;
        for (int j = 0; j < np; j++) p[j] = mu[j];
        p[t0i] = t0;
        if (ll_at(m, b, p, pr) >= thr) lo1 = t0;
        else break;
    }
    for (int g = 0; g < NG; g++) {
        const double t0 = grid(g);
        if (t0 <= best_t0) continue;
        for (int j = 0; j < np; j++) p[j] = mu[j];
        p[t0i] = t0;
        if (ll_at(m, b, p, pr) >= thr) hi1 = t0;
        else break;
    }
    mu[t0i] = best_t0;
    ls[t0i] = std::log(std::fmax((hi1 - lo1) / 2.0, 0.1));
}

void profile_refine_all(int m, const double* mu, double* ls, const BandView& b) {
    const int np = n_params(m), se = np - 1;
    std::vector<double> pr(std::max<int64_t>(b.n, 1));
    double p[7];
    for (int j = 0; j < np; j++) p[j] = mu[j];
    const double thr = ll_at(m, b, p, pr) - 0.5;
    double lo[7], hi[7], pc[7], pw[7];
    pso_bounds_cpu(m, lo, hi);
    population_priors(m, pc, pw);
    const int NG = 15;
    for (int j = 0; j < np; j++) {
        if (j == se) continue;
        const double svi_sigma = std::exp(ls[j]);
        if (hi[j] - lo[j] < 1e-10) continue;
        double lo1 = mu[j], hi1 = mu[j];
        for (int g = 0; g < NG; g++) {
            const double val = lo[j] + (hi[j] - lo[j]) * (double)g / (double)(NG - 1);
            for (int k = 0; k < np; k++) p[k] = mu[k];
            p[j] = val;
            if (ll_at(m, b, p, pr) >= thr) {
                lo1 = std::fmin(lo1, val);
                hi1 = std::fmax(hi1, val);
            }
        }
        const double prof = std::fmax((hi1 - lo1) / 2.0, 1e-6);
        ls[j] = std::log(std::fmax(std::fmax(svi_sigma, prof), pw[j]));
    }
}

double mag_chi2(int m, const double* mu, const double* pso, const BandView& b, bool svi_only) {
    const int64_t n = b.n;
    std::vector<double> sv(std::max<int64_t>(n, 1)), ps(std::max<int64_t>(n, 1));
    eval_times(m, mu, b.t, n, sv.data());
    int64_t npos = 0;
    for (int64_t i = 0; i < n; i++)
        if (sv[i] * b.peak > 0.0) npos++;
    const double* mp = sv.data();
    if (!svi_only && npos < n / 2) {
        eval_times(m, pso, b.t, n, ps.data());
        mp = ps.data();
    }
    double s = 0.0;
    int64_t cnt = 0;
    for (int64_t i = 0; i < n; i++) {
        double pf = mp[i] * b.peak;
        if (pf > 0.0 && b.f[i] * b.peak > 0.0) {
            double mpred = -2.5 * std::log10(pf) + 23.9;
            double me = 1.0857 * b.fe[i] / b.f[i];
            if (me > 0.0) {
                double r = b.mag[i] - mpred;
                s += r * r / (me * me);
                cnt++;
            }
        }
    }
    return cnt > 0 ? s / (double)cnt : NAN;
}

}  // namespace

}  // namespace orc

extern "C" int orc_finalize(int64_t n_bands, const int64_t* off, const double* t, const double* f,
                            const double* v, const int32_t* up, const double* uf, const double* fe,
                            const double* mag_obs, const double* peak, const int32_t* model_ids,
                            const double* pso, const double* mu, const double* log_sigma, int flags,
                            double* out_mu, double* out_log_sigma, double* out_mag_chi2, int n_threads) {
    for (int64_t k = 0; k < n_bands; k++)
        if (orc::n_params(model_ids[k]) < 0) return -1;
    orc::parallel_for(n_bands, n_threads, [&](int64_t k) {
        const int m = model_ids[k];
        const int64_t a = off[k];
        orc::BandView b{t + a, f + a, v + a, up + a, uf + a, fe + a, mag_obs + a, peak[k], off[k + 1] - a};
        double* om = out_mu + k * 7;
        double* ol = out_log_sigma + k * 7;
        for (int j = 0; j < 7; j++) {
            om[j] = mu[k * 7 + j];
            ol[j] = log_sigma[k * 7 + j];
        }
        out_mag_chi2[k] = NAN;
        if (b.n == 0) return;
        if (flags & ORC_FINALIZE_T0) orc::profile_t0_refine(m, om, ol, b);
        if (flags & ORC_FINALIZE_ALL) orc::profile_refine_all(m, om, ol, b);
        if (flags & ORC_FINALIZE_MAG_CHI2)
            out_mag_chi2[k] = orc::mag_chi2(m, om, pso + k * 7, b, (flags & ORC_FINALIZE_SVI_CHI2) != 0);
    });
    return 0;
}
