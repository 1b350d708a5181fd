// thermal.cpp -- oracle restatement of the blackbody colour-temperature fit.
//
// TEST INFRASTRUCTURE ONLY (parity checker for the HIP product).
//
// fit_thermal (src/thermal.rs:222-418) per source, after the host has chosen
// the reference band (bluest of u > g > r with >= 5 points) and listed the
// other bands' observations with their effective wavelengths:
//   * reference-band GP: the prefit dense GP when given, else fit_ref_band_gp
//     (src/thermal.rs:167-212): subsample (all points up to 30, else 25, by
//     src/gp.rs:5-27), amps {0.1, 0.3} x l = max(duration / {6, 12, 24}, 0.1),
//     train RMS at the subsample, fits whose minimum prediction over the whole
//     band is more than 6 mag from the brightest subsampled point rejected;
//   * colour observations: ref GP mean - mag at every observation of the other
//     bands whose GP mean is finite, error sqrt(err^2 + 0.02^2);
//   * cost (src/thermal.rs:112-141): mean of (colour - bb_color_mag)^2 /
//     (err^2 + 1e-10), 1e10 for a non-finite model colour; bb_color_mag
//     (:75-92) with the Wien branch above an exponent of 500; temperature
//     10^(log T0 + rate t);
//   * two particle-swarm runs of 10 particles x 30 iterations in the box
//     [3, 6] x [-0.05, 0.01]: the particle swarm of argmin 0.11
//     (src/thermal.rs:336-339, argmin's ParticleSwarm: inertia 1/(2 ln 2),
//     cognitive = social = 1/2 + ln 2, positions uniform in the box, velocities
//     uniform in [-range, range], every particle moved, clamped to the box,
//     then all evaluated; personal / global bests on strict improvement in
//     particle order); best of the two runs, errors = population std of the
//     two runs' best positions; log T at the last colour time clamped to
//     [2, 7].
// Random numbers: Philox4x32-10 keyed by the seed, counter (0,
// 64 run + particle, iteration or 0xFFFFFFFF, 0x54480000) for every source, in place of
// argmin's entropy-seeded generator (not reproducible even by the reference:
// parity of the swarm's path is unpinned).
#include "lcf_oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace orc {
namespace th {

static const double HC_OVER_K = 1.4388e8;
static const double LOG10_E = 0.434294481903251827651128918916605082;

static inline double u24(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

// bb_color_mag (src/thermal.rs:75-92)
static double bb_color_mag(double lref, double l, double temp) {
    double x_ref = HC_OVER_K / (lref * temp);
    double x = HC_OVER_K / (l * temp);
    if (x > 500.0 || x_ref > 500.0) return -2.5 * (3.0 * std::log10(l / lref) + (x - x_ref) * LOG10_E);
    double q = l / lref;
    double ratio = q * q * q * (std::exp(x) - 1.0) / (std::exp(x_ref) - 1.0);
    if (ratio <= 0.0 || !std::isfinite(ratio)) return NAN;
    return -2.5 * std::log10(ratio);
}

struct Obs {
    double t, c, e, l;
};

// ThermalCost::cost (src/thermal.rs:116-140)
static double cost(const std::vector<Obs>& o, double lref, double lt0, double rate) {
    double chi2 = 0.0;
    for (const Obs& b : o) {
        double temp = std::pow(10.0, lt0 + rate * b.t);
        double mc = bb_color_mag(lref, b.l, temp);
        if (!std::isfinite(mc)) return 1e10;
        double r = b.c - mc;
        chi2 += r * r / (b.e * b.e + 1e-10);
    }
    return chi2 / (double)std::max<size_t>(o.size(), 1);
}

// One particle-swarm run; returns the best position and cost.
static void swarm(const std::vector<Obs>& o, double lref, int run, const uint32_t key[2], double* best_p,
                  double* best_c) {
    const int P = 10, IT = 30;
    const double lo[2] = {3.0, -0.05}, hi[2] = {6.0, 0.01};
    const double w = 1.0 / (2.0 * std::log(2.0)), cc = 0.5 + std::log(2.0), cs = 0.5 + std::log(2.0);
    double x[P][2], v[P][2], pb[P][2], pc[P], c[P];
    for (int p = 0; p < P; p++) {
        uint32_t ctr[4] = {0u, (uint32_t)(64 * run + p), 0xFFFFFFFFu, 0x54480000u}, r[4];
        orc_philox4x32_10(ctr, key, r);
        for (int d = 0; d < 2; d++) {
            double rg = hi[d] - lo[d];
            x[p][d] = lo[d] + u24(r[d]) * rg;
            v[p][d] = -rg + u24(r[2 + d]) * (2.0 * rg);
            pb[p][d] = x[p][d];
        }
        c[p] = pc[p] = cost(o, lref, x[p][0], x[p][1]);
    }
    int g = 0;
    for (int p = 1; p < P; p++)
        if (c[p] < c[g]) g = p;
    double gp[2] = {x[g][0], x[g][1]}, gc = c[g];
    for (int it = 0; it < IT; it++) {
        for (int p = 0; p < P; p++) {
            uint32_t ctr[4] = {0u, (uint32_t)(64 * run + p), (uint32_t)it, 0x54480000u}, r[4];
            orc_philox4x32_10(ctr, key, r);
            for (int d = 0; d < 2; d++) {
                double mom = v[p][d] * w;
                double po = (pb[p][d] - x[p][d]) * u24(r[d]) * cc;
                double pg = (gp[d] - x[p][d]) * u24(r[2 + d]) * cs;
                v[p][d] = mom + po + pg;
                double nx = x[p][d] + v[p][d];
                x[p][d] = std::min(std::max(nx, lo[d]), hi[d]);
            }
        }
        for (int p = 0; p < P; p++) c[p] = cost(o, lref, x[p][0], x[p][1]);
        for (int p = 0; p < P; p++)
            if (c[p] < pc[p]) {
                pc[p] = c[p];
                pb[p][0] = x[p][0];
                pb[p][1] = x[p][1];
                if (c[p] < gc) {
                    gc = c[p];
                    gp[0] = x[p][0];
                    gp[1] = x[p][1];
                }
            }
    }
    best_p[0] = gp[0];
    best_p[1] = gp[1];
    *best_c = gc;
}

static inline double rbf(double a, double b, double amp, double it) {
    double d = a - b;
    return amp * std::exp(-d * d * it);
}

// Dense GP mean from a state row (alpha[cap], x[cap], amp, 1/(2 l^2), mean, m).
struct GpRow {
    const double* a;
    const double* x;
    double amp, it, ym;
    int m;
    double mean(double t) const {
        double s = 0.0;
        for (int i = 0; i < m; i++) s += rbf(t, x[i], amp, it) * a[i];
        return s + ym;
    }
};

// fit_ref_band_gp (src/thermal.rs:167-212) into a state row; false if no fit.
static bool fit_ref_gp(const double* t, const double* m, const double* e, int n, double dur, double* st, int cap) {
    const int ms = n <= 30 ? n : 25;
    std::vector<double> ts, ys, ns;
    if (n <= ms) {
        ts.assign(t, t + n);
        ys.assign(m, m + n);
        for (int i = 0; i < n; i++) ns.push_back(std::max(e[i] * e[i], 1e-6));
    } else {
        double step = (double)n / (double)ms;
        for (int i = 0; i < ms; i++) {
            int k = std::min((int)std::floor(((double)i + 0.5) * step), n - 1);
            ts.push_back(t[k]);
            ys.push_back(m[k]);
            ns.push_back(std::max(e[k] * e[k], 1e-6));
        }
    }
    const int k = (int)ts.size();
    if (k == 0 || k > cap) return false;
    const double amps[2] = {0.1, 0.3}, fac[3] = {6.0, 12.0, 24.0};
    double best = INFINITY;
    bool have = false;
    double ym = 0.0;
    for (double v : ys) ym += v;
    ym /= k;
    double omin = INFINITY;
    for (double v : ys) omin = std::isnan(v) ? omin : std::min(omin, v);
    for (double amp : amps)
        for (double f : fac) {
            double ls = std::max(dur / f, 0.1);
            double it = 0.5 / (ls * ls);
            std::vector<double> K(k * k);
            for (int i = 0; i < k; i++) {
                for (int j = 0; j <= i; j++) K[i * k + j] = K[j * k + i] = rbf(ts[i], ts[j], amp, it);
                K[i * k + i] += std::max(ns[i], 1e-10);
            }
            // Cholesky-Crout + solves (src/sparse_gp.rs:27-65)
            bool ok = true;
            for (int j = 0; j < k && ok; j++) {
                double s = K[j * k + j];
                for (int q = 0; q < j; q++) s -= K[j * k + q] * K[j * k + q];
                if (s <= 0.0) { ok = false; break; }
                K[j * k + j] = std::sqrt(s);
                for (int i = j + 1; i < k; i++) {
                    double si = K[i * k + j];
                    for (int q = 0; q < j; q++) si -= K[i * k + q] * K[j * k + q];
                    K[i * k + j] = si / K[j * k + j];
                }
            }
            if (!ok) continue;
            std::vector<double> z(k), al(k);
            for (int i = 0; i < k; i++) {
                double s = ys[i] - ym;
                for (int j = 0; j < i; j++) s -= K[i * k + j] * z[j];
                z[i] = s / K[i * k + i];
            }
            for (int i = k - 1; i >= 0; i--) {
                double s = z[i];
                for (int j = i + 1; j < k; j++) s -= K[j * k + i] * al[j];
                al[i] = s / K[i * k + i];
            }
            GpRow g{al.data(), ts.data(), amp, it, ym, k};
            double rss = 0.0;
            for (int i = 0; i < k; i++) {
                double r = ys[i] - g.mean(ts[i]);
                rss += r * r;
            }
            double rms = std::sqrt(rss / (double)k);
            double pmin = INFINITY;
            for (int i = 0; i < n; i++) {
                double p = g.mean(t[i]);
                if (!std::isnan(p)) pmin = std::min(pmin, p);
            }
            if (std::isfinite(pmin) && std::fabs(pmin - omin) > 6.0) continue;
            if (std::isfinite(rms) && rms < best) {
                best = rms;
                have = true;
                for (int i = 0; i < k; i++) { st[i] = al[i]; st[cap + i] = ts[i]; }
                st[2 * cap] = amp;
                st[2 * cap + 1] = it;
                st[2 * cap + 2] = ym;
                st[2 * cap + 3] = k;
            }
        }
    return have;
}

}  // namespace th
}  // namespace orc

extern "C" {

double orc_bb_color_mag(double lambda_ref, double lambda, double temp) {
    return orc::th::bb_color_mag(lambda_ref, lambda, temp);
}

int orc_batch_thermal(int64_t n_src, const int64_t* ref_off, const double* ref_t, const double* ref_m,
                      const double* ref_e, const int64_t* col_off, const double* col_t, const double* col_m,
                      const double* col_e, const double* col_lambda, const double* lambda_ref,
                      const double* duration, double* gp_state, int state_stride, uint64_t seed, int32_t* status,
                      double* out, int32_t* col_used, int n_threads) {
    if (n_src < 0 || state_stride < 8) return -1;
    const int cap = (state_stride - 4) / 2;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc::parallel_for(n_src, n_threads, [&](int64_t s) {
        double* o = out + s * 8;
        for (int k = 0; k < 8; k++) o[k] = NAN;
        status[s] = 0;
        double* st = gp_state + s * state_stride;
        const int64_t c0 = col_off[s], c1 = col_off[s + 1];
        for (int64_t i = c0; i < c1; i++) col_used[i] = 0;
        if (!(duration[s] > 0.0)) return;
        if (!(st[2 * cap + 3] > 0.0)) {
            const int64_t r0 = ref_off[s];
            for (int k = 0; k < state_stride; k++) st[k] = 0.0;
            if (!orc::th::fit_ref_gp(ref_t + r0, ref_m + r0, ref_e + r0, (int)(ref_off[s + 1] - r0), duration[s],
                                     st, cap))
                return;
        }
        orc::th::GpRow g{st, st + cap, st[2 * cap], st[2 * cap + 1], st[2 * cap + 2], (int)st[2 * cap + 3]};
        std::vector<orc::th::Obs> obs;
        for (int64_t i = c0; i < c1; i++) {
            double rm = g.mean(col_t[i]);
            if (!std::isfinite(rm)) continue;
            obs.push_back({col_t[i], rm - col_m[i], std::sqrt(col_e[i] * col_e[i] + 0.02 * 0.02), col_lambda[i]});
            col_used[i] = 1;
        }
        o[7] = (double)obs.size();
        status[s] = 1;
        if (obs.size() < 3) return;
        double bp[2][2], bc[2];
        for (int run = 0; run < 2; run++)
            orc::th::swarm(obs, lambda_ref[s], run, key, bp[run], &bc[run]);
        const int b = bc[1] < bc[0] ? 1 : 0;
        double m0 = (bp[0][0] + bp[1][0]) / 2.0, m1 = (bp[0][1] + bp[1][1]) / 2.0;
        double v0 = ((bp[0][0] - m0) * (bp[0][0] - m0) + (bp[1][0] - m0) * (bp[1][0] - m0)) / 2.0;
        double v1 = ((bp[0][1] - m1) * (bp[0][1] - m1) + (bp[1][1] - m1) * (bp[1][1] - m1)) / 2.0;
        double tl = -INFINITY;
        for (const auto& q : obs) tl = std::max(tl, q.t);
        double lt = bp[b][0] + bp[b][1] * tl;
        o[0] = bp[b][0];                                                     // log_temp_peak
        o[1] = std::isfinite(lt) ? std::min(std::max(lt, 2.0), 7.0) : NAN;  // log_temp_latest
        o[2] = bp[b][1];                                                     // cooling_rate
        o[3] = std::sqrt(v0);                                                // log_temp_peak_err
        o[4] = std::sqrt(v1);                                                // cooling_rate_err
        o[5] = std::isfinite(bc[b]) ? bc[b] : NAN;                           // chi2
        o[6] = bc[b];
        status[s] = 2;
    });
    return 0;
}

}  // extern "C"

// ---- fit_bb_temperature (src/gp2d.rs:264-324) --------------------------------
namespace orc {
namespace bb {

static double planck_relative_mag(double lam, double temp) {  // src/gp2d.rs:264-271
    const double HC = 1.4387769e8;
    double x = HC / (lam * temp);
    double log_b = -5.0 * std::log(lam) - std::log(std::exp(x) - 1.0);
    return -2.5 * log_b / 2.302585092994045684017991454684364208;
}

}  // namespace bb
}  // namespace orc

extern "C" int orc_fit_bb_temperature(const double* lam, const double* mag, const double* err, int n,
                                      double* out3) {
    if (n < 2) return 1;
    double best_chi2 = INFINITY, best_lt = 4.0, best_off = 0.0;
    const double nf = n;
    for (int i = 0; i < 200; i++) {
        double lt = 3.0 + (double)i * 2.5 / (200.0 - 1.0);
        double temp = std::pow(10.0, lt);
        std::vector<double> model(n);
        for (int j = 0; j < n; j++) model[j] = orc::bb::planck_relative_mag(lam[j], temp);
        double sw = 0.0, swd = 0.0;
        for (int j = 0; j < n; j++) {
            double w = 1.0 / std::max(err[j] * err[j], 1e-10);
            sw += w;
            swd += w * (mag[j] - model[j]);
        }
        double off = swd / sw;
        double chi = 0.0;
        for (int j = 0; j < n; j++) {
            double r = mag[j] - model[j] - off;
            chi += r * r / std::max(err[j] * err[j], 1e-10);
        }
        chi = chi / std::max(nf - 2.0, 1.0);
        if (chi < best_chi2) { best_chi2 = chi; best_lt = lt; best_off = off; }
    }
    out3[0] = best_lt;
    out3[1] = best_chi2;
    out3[2] = best_off;
    return 0;
}
