// fit_parametric.cpp -- oracle restatement of the reference's CPU
// parametric pipeline: the job building and decisions around the swarm fits.
//
// TEST INFRASTRUCTURE ONLY (parity checker for the HIP product).
//
// Per source (flux bands in the caller's order, which stands for the
// reference's HashMap iteration order):
//   fit_parametric (src/parametric.rs:3742-3761): the joint multi-band Villar
//     fit when >= 2 bands have values and not fit_all_models, the per-band
//     path otherwise or when the joint fit returns nothing;
//   band normalisation (:3581-3632): bands with no values or a peak flux
//     <= 0 dropped, flux / err / peak, obs_var = err^2 + 1e-10, upper limit
//     = err > 0 && f / err < 3 at 3 err / peak, magnitudes -2.5 log10 f +
//     23.9; bands sorted by size, largest first (stable);
//   fit_parametric_model (:3563-3739): per band pso_model_select, or the
//     forced model; profile_t0_search on the first band for SVI; the
//     posterior (SVI or Laplace), profile_t0_refine + profile_refine_all +
//     magnitude chi2; fit_multi_bazin on the first band;
//   pso_fit_single_model (:1864-1999): CPU box with t0 narrowed to
//     [t_first - 30, t_peak + 10] when that is non-empty, restarts seed 42,
//     137 (seeded at t_peak - 5) and 271 (seeded at t_first - 15) with the
//     other coordinates at the population-prior centres, best in restart
//     order, then lbfgs_refine (not for MetzgerKN);
//   pso_model_select (:2001-2066): Bazin first; unless fit_all_models and
//     while Bazin's cost is >= 2, Arnett, Tde, Afterglow, Villar, Magnetar,
//     ShockCooling, MetzgerKN; the lowest BIC = 2 cost n + k ln n wins
//     (strict, in that order);
//   profile_t0_search (:1767-1849): fixed-t0 swarms (30 x 60, stall 12,
//     seed 42, no prior) at 5 t0 on [max(pilot - 5, t_first - 10),
//     pilot + 5], then 3 within 0.5 d of the best, kept only if cheaper;
//   fit_parametric_multiband (:3813-4131): reference band + offsets (CPU
//     Villar box, offsets +-3 sigma of the relative priors), t0 narrowed on
//     the reference band, restarts 42 / 137 (seeded: prior centres, t0 at
//     t_peak - 5, offsets at the relative means) / 271 of 200 x 1500, stall
//     60; per band the posterior, profile refinement and the magnitude chi2
//     of the posterior mean; fit_multi_bazin on the first band;
//   fit_multi_bazin (:1008-1157): K = 1..4 with restarts 42, 137 (from the
//     K - 1 solution plus a component at the largest detection residual) and
//     271, BIC choice and the +2 stop.
// The swarms and polishes are the oracle's cpu_pso jobs (Philox streams in
// place of SmallRng, every job on stream 0 as the reference reseeds each
// call; backtracking Armijo in place of argmin's More-Thuente).  The SVI
// posterior is the oracle's batched SVI with the CPU fit's settings (1000
// steps, 4 samples, lr 0.01, min 200 / stall 50, + ln 4, the same noise
// stream for every band); the Laplace posterior is laplace_fit.  Deviations
// kept as the device keeps them: more than 9 bands skip the joint fit, and
// the joint t0 narrowing is skipped when it would empty the box.
#include "lcf_oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace orc {

namespace {

const int T0_IDX[8] = {2, 3, 2, 1, 1, 1, 1, 3};
const double LN10 = 2.30258509299404568402;
const double REL_MEAN[7] = {-0.015 * LN10, -0.001, -0.095 * LN10, -1.181, -0.195 * LN10, -0.323 * LN10,
                            -0.025 * LN10};
const double REL_SIG[7] = {0.08 * LN10, 0.008, 0.25 * LN10, 1.5, 0.35 * LN10, 0.25 * LN10, 0.30 * LN10};
// pso_model_select's order after Bazin (src/parametric.rs:2002-2011)
const int SELECT_REST[7] = {3, 2, 6, 1, 4, 5, 7};
const uint64_t SEEDS[3] = {42, 137, 271};

struct FitBand {
    int input;  // index of the input band within the source
    std::vector<double> t, f, fe, v, uf, mag;
    std::vector<int32_t> up;
    double peak = 0.0;
    int n() const { return (int)t.size(); }
};

// The source's normalised bands as one CSR (the jobs' band index = position).
struct Batch {
    std::vector<int64_t> off{0};
    std::vector<double> t, f, fe, v, uf, mag;
    std::vector<int32_t> up;
    std::vector<double> peak;
    void add(const FitBand& b) {
        t.insert(t.end(), b.t.begin(), b.t.end());
        f.insert(f.end(), b.f.begin(), b.f.end());
        fe.insert(fe.end(), b.fe.begin(), b.fe.end());
        v.insert(v.end(), b.v.begin(), b.v.end());
        uf.insert(uf.end(), b.uf.begin(), b.uf.end());
        mag.insert(mag.end(), b.mag.begin(), b.mag.end());
        up.insert(up.end(), b.up.begin(), b.up.end());
        peak.push_back(b.peak);
        off.push_back((int64_t)t.size());
    }
};

bool band_fit_data(const double* t, const double* f, const double* e, int n, FitBand& o) {
    if (n == 0) return false;
    double peak = -1.7976931348623157e308;  // f64::MIN, NaN ignored as f64::max does
    for (int i = 0; i < n; i++)
        if (!std::isnan(f[i])) peak = std::max(peak, f[i]);
    if (!(peak > 0.0)) return false;
    o.peak = peak;
    for (int i = 0; i < n; i++) {
        o.t.push_back(t[i]);
        o.f.push_back(f[i] / peak);
        o.fe.push_back(e[i] / peak);
        o.v.push_back((e[i] / peak) * (e[i] / peak) + 1e-10);
        o.up.push_back(e[i] > 0.0 && f[i] / e[i] < 3.0 ? 1 : 0);
        o.uf.push_back(3.0 * e[i] / peak);
        o.mag.push_back(-2.5 * std::log10(f[i]) + 23.9);
    }
    return true;
}

orc_pso_job make_job(int kind, int model, int band, int dim, const double* lo, const double* hi, int particles,
                     int iters, int stall, int prior, uint64_t seed) {
    orc_pso_job j;
    std::memset(&j, 0, sizeof j);
    j.kind = kind;
    j.model = model;
    j.band = band;
    j.n_bands = 1;
    j.dim = dim;
    j.n_particles = particles;
    j.max_iters = iters;
    j.stall_iters = stall;
    j.use_prior = prior;
    j.seed = seed;
    for (int d = 0; d < dim; d++) {
        j.lo[d] = lo[d];
        j.hi[d] = hi[d];
    }
    return j;
}

struct JobOut {
    std::vector<double> x;
    double cost;
};

JobOut run_job(const Batch& B, const orc_pso_job& j) {
    double x[64], c;
    int32_t it;
    orc_cpu_pso(B.off.data(), B.t.data(), B.f.data(), B.v.data(), B.up.data(), B.uf.data(), (int64_t)B.t.size(), &j,
                1, x, &c, &it, 1);
    return {std::vector<double>(x, x + j.dim), c};
}

double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Time of the largest detected flux (times[0] when none) and the earliest time.
void peak_and_first(const FitBand& b, double& peak_t, double& t_first) {
    peak_t = b.t[0];
    double pf = -INFINITY;
    t_first = INFINITY;
    for (int i = 0; i < b.n(); i++) {
        if (!b.up[i] && b.f[i] > pf) {
            pf = b.f[i];
            peak_t = b.t[i];
        }
        t_first = std::min(t_first, b.t[i]);
    }
}

struct Fit {
    std::vector<double> p;
    double cost = INFINITY;
};

// pso_fit_single_model (src/parametric.rs:1864-1999) with lbfgs_refine.
Fit single_model(const Batch& B, const FitBand& b, int bi, int m) {
    const int np = n_params(m), t0i = T0_IDX[m];
    double lo[7], hi[7], pc[7], pw[7];
    pso_bounds_cpu(m, lo, hi);
    population_priors(m, pc, pw);
    double peak_t, t_first;
    peak_and_first(b, peak_t, t_first);
    const double t0_lo = std::max(t_first - 30.0, lo[t0i]), t0_hi = std::min(peak_t + 10.0, hi[t0i]);
    if (t0_lo < t0_hi) {
        lo[t0i] = t0_lo;
        hi[t0i] = t0_hi;
    }
    double at_peak[7], early[7];
    for (int d = 0; d < np; d++) at_peak[d] = early[d] = clampd(pc[d], lo[d], hi[d]);
    at_peak[t0i] = clampd(peak_t - 5.0, lo[t0i], hi[t0i]);
    early[t0i] = clampd(t_first - 15.0, lo[t0i], hi[t0i]);
    Fit best;
    for (int r = 0; r < 3; r++) {
        orc_pso_job j = make_job(ORC_CPSO_SWARM_BAND, m, bi, np, lo, hi, 30, 60, 12, 1, SEEDS[r]);
        if (r > 0) {
            j.seeded = 1;
            for (int d = 0; d < np; d++) j.x0[d] = r == 1 ? at_peak[d] : early[d];
        }
        JobOut o = run_job(B, j);
        if (o.cost < best.cost) {
            best.cost = o.cost;
            best.p = o.x;
        }
    }
    if (!best.p.empty() && m != 7) {
        orc_pso_job j = make_job(ORC_CPSO_POLISH, m, bi, np, lo, hi, 0, 50, 0, 1, 0);
        for (int d = 0; d < np; d++) j.x0[d] = best.p[(size_t)d];
        j.start_cost = best.cost;
        JobOut o = run_job(B, j);
        best.p = o.x;
        best.cost = o.cost;
    }
    return best;
}

struct Selection {
    int model = 0;
    Fit fit;
    double chi2[8];    // per_model_chi2 value (NaN = None)
    int has[8];        // key present
    Fit params[8];     // per_model_params (fit_all_models)
};

// pso_model_select (src/parametric.rs:2001-2066).
Selection model_select(const Batch& B, const FitBand& b, int bi, bool fit_all) {
    Selection s;
    for (int k = 0; k < 8; k++) {
        s.chi2[k] = NAN;
        s.has[k] = 0;
    }
    const double n = (double)std::max(b.n(), 1), ln_n = std::log(n);
    auto bic = [&](double cost, int m) { return 2.0 * cost * n + (double)n_params(m) * ln_n; };
    auto fin = [](double v) { return std::isfinite(v) ? v : NAN; };
    Fit bz = single_model(B, b, bi, 0);
    s.chi2[0] = fin(bz.cost);
    s.has[0] = 1;
    if (fit_all) s.params[0] = bz;
    s.model = 0;
    s.fit = bz;
    double best_bic = bic(bz.cost, 0);
    if (!fit_all && bz.cost < 2.0) {
        for (int k = 0; k < 7; k++) s.has[SELECT_REST[k]] = 1;  // None entries
    } else {
        for (int k = 0; k < 7; k++) {
            const int m = SELECT_REST[k];
            Fit f = single_model(B, b, bi, m);
            s.chi2[m] = fin(f.cost);
            s.has[m] = 1;
            if (fit_all) s.params[m] = f;
            const double mb = bic(f.cost, m);
            if (mb < best_bic) {
                best_bic = mb;
                s.model = m;
                s.fit = f;
            }
        }
    }
    return s;
}

std::vector<double> linspace(double a, double b, int n) {
    if (n <= 1) return {a};
    std::vector<double> v((size_t)n);
    const double step = (b - a) / (double)(n - 1);
    for (int i = 0; i < n; i++) v[(size_t)i] = a + (double)i * step;
    return v;
}

// profile_t0_search (src/parametric.rs:1767-1849).
Fit profile_t0_search(const Batch& B, const FitBand& b, int bi, int m, const Fit& pilot) {
    const int np = n_params(m), t0i = T0_IDX[m];
    const double pilot_t0 = pilot.p[(size_t)t0i];
    double t_first = INFINITY;
    for (double x : b.t) t_first = std::min(t_first, x);
    const double t0_min = std::max(pilot_t0 - 5.0, t_first - 10.0), t0_max = pilot_t0 + 5.0;
    double lo7[7], hi7[7], lo[7], hi[7];
    pso_bounds_cpu(m, lo7, hi7);
    for (int d = 0, k = 0; d < np; d++)
        if (d != t0i) {
            lo[k] = lo7[d];
            hi[k] = hi7[d];
            k++;
        }
    auto run_at = [&](double t0) {
        orc_pso_job j = make_job(ORC_CPSO_SWARM_BAND, m, bi, np - 1, lo, hi, 30, 60, 12, 0, 42);
        j.fixed_t0 = 1;
        j.t0_value = t0;
        JobOut o = run_job(B, j);
        Fit f;
        f.cost = o.cost;
        for (int d = 0, k = 0; d < np; d++) f.p.push_back(d == t0i ? t0 : o.x[(size_t)k++]);
        return f;
    };
    Fit best = pilot;
    double best_t0 = pilot_t0;
    for (double t0 : linspace(t0_min, t0_max, 5)) {
        Fit f = run_at(t0);
        if (f.cost < best.cost && !f.p.empty()) {
            best = f;
            best_t0 = t0;
        }
    }
    for (double t0 : linspace(std::max(best_t0 - 0.5, t0_min), std::min(best_t0 + 0.5, t0_max), 3)) {
        Fit f = run_at(t0);
        if (f.cost < best.cost && !f.p.empty()) best = f;
    }
    return best;
}

double mb_eval(const std::vector<double>& p, int k, double t) {
    double s = 0.0;
    for (int c = 0; c < k; c++) {
        const double a = std::exp(p[(size_t)(4 * c)]), tr = std::exp(p[(size_t)(4 * c + 2)]),
                     tf = std::exp(p[(size_t)(4 * c + 3)]);
        const double dt = t - p[(size_t)(4 * c + 1)];
        s += a * std::exp(-dt / tf) * (1.0 / (1.0 + std::exp(-dt / tr)));
    }
    return s + p[(size_t)(4 * k)];
}

struct MbResult {
    bool has = false;
    int best_k = 1;
    std::vector<double> params;
    double cost = NAN, bic = NAN;
    double pk_cost[4], pk_bic[4];
};

// fit_multi_bazin (src/parametric.rs:1008-1157).
MbResult fit_multi_bazin(const Batch& B, const FitBand& b, int bi) {
    MbResult r;
    for (int k = 0; k < 4; k++) r.pk_cost[k] = r.pk_bic[k] = NAN;
    const int n = b.n();
    if (n < 5) return r;
    double t_min = INFINITY, t_max = -INFINITY;
    for (double x : b.t) {
        t_min = std::min(t_min, x);
        t_max = std::max(t_max, x);
    }
    const double nn = (double)n, ln_n = std::log(nn);
    double best_cost = INFINITY, best_bic = INFINITY;
    std::vector<double> best_params, prev;
    for (int k = 1; k <= 4; k++) {
        const int np = 4 * k + 2;
        std::vector<double> lo, hi;
        for (int c = 0; c < k; c++) {
            lo.insert(lo.end(), {-3.0, t_min - 30.0, -2.0, -2.0});
            hi.insert(hi.end(), {3.0, t_max + 30.0, 5.0, 6.0});
        }
        lo.insert(lo.end(), {-0.3, -5.0});
        hi.insert(hi.end(), {0.3, 0.0});
        std::vector<double> init;
        if (k > 1 && (int)prev.size() == 4 * (k - 1) + 2) {
            const int nc = 4 * (k - 1);
            init.assign(prev.begin(), prev.begin() + nc);
            int peak_idx = 0;
            double peak_val = -INFINITY;
            for (int i = 0; i < n; i++) {
                const double res = b.f[(size_t)i] - mb_eval(prev, k - 1, b.t[(size_t)i]);
                if (!b.up[(size_t)i] && res > peak_val) {
                    peak_val = res;
                    peak_idx = i;
                }
            }
            init.insert(init.end(), {std::log(std::max(peak_val, 1e-10)), b.t[(size_t)peak_idx], 1.0, 1.0,
                                     prev[(size_t)nc], prev[(size_t)nc + 1]});
            for (int i = 0; i < np; i++) init[(size_t)i] = clampd(init[(size_t)i], lo[(size_t)i], hi[(size_t)i]);
        }
        auto job = [&](uint64_t seed) {
            return make_job(ORC_CPSO_SWARM_MULTIBAZIN, k, bi, np, lo.data(), hi.data(), 30, 60, 12, 0, seed);
        };
        JobOut o1 = run_job(B, job(42));
        if (!init.empty()) {
            orc_pso_job j = job(137);
            j.seeded = 1;
            for (int d = 0; d < np; d++) j.x0[d] = init[(size_t)d];
            JobOut o2 = run_job(B, j);
            if (o2.cost < o1.cost) o1 = o2;
        }
        JobOut o3 = run_job(B, job(271));
        if (o3.cost < o1.cost) o1 = o3;
        const double kb = 2.0 * o1.cost * nn + (double)np * ln_n;
        r.pk_cost[k - 1] = o1.cost;
        r.pk_bic[k - 1] = kb;
        if (kb < best_bic) {
            best_bic = kb;
            best_cost = o1.cost;
            r.best_k = k;
            best_params = o1.x;
        }
        prev = o1.x;
        if (k > 1 && kb > r.pk_bic[k - 2] + 2.0) break;
    }
    r.has = !best_params.empty();
    r.params = best_params;
    r.cost = best_cost;
    r.bic = best_bic;
    return r;
}

// Output record (ORC_FP_RECORD doubles, indexed by input band).
enum {
    R_RANK = 0, R_MODEL = 1, R_PSO = 2, R_PSO_CHI2 = 9, R_MU = 10, R_LS = 17, R_ELBO = 24, R_NOBS = 25,
    R_MAG_CHI2 = 26, R_CHI2 = 27, R_HAS = 35, R_ROUTE = 43, R_MB_HAS = 44, R_MB_K = 45, R_MB_P = 46,
    R_MB_COST = 64, R_MB_BIC = 65, R_MB_PKC = 66, R_MB_PKB = 70, R_PARAMS = 74, R_END = 130
};
static_assert(R_END == ORC_FP_RECORD, "record layout");

struct Posterior {
    double mu[7], ls[7], elbo;
};

Posterior posterior(const Batch& B, int bi, int m, const std::vector<double>& p, int method) {
    Posterior q;
    double p7[7] = {0}, pc[7] = {0}, pw[7] = {0};
    for (size_t d = 0; d < p.size(); d++) p7[d] = p[d];
    const int64_t a = B.off[(size_t)bi], n = B.off[(size_t)bi + 1] - a;
    if (method == 0) {
        orc_svi_prior_for_model(m, p7, pc, pw);
        const int64_t off[2] = {0, n};
        int32_t mid = m, steps;
        orc_batch_svi(1, off, B.t.data() + a, B.f.data() + a, B.v.data() + a, B.up.data() + a, B.uf.data() + a,
                      &mid, p7, pc, pw, 1000, 4, 0.01, ORC_SVI_RNG_PHILOX, 42, 0,
                      ORC_SVI_CPU_STOPPING | ORC_SVI_INFLATE_SIGMA | ORC_SVI_SHARED_STREAM, q.mu, q.ls, &q.elbo,
                      &steps, nullptr, 1);
    } else {
        for (int d = 0; d < 7; d++) q.mu[d] = p7[d];
        orc_laplace_fit(m, p7, B.t.data() + a, B.f.data() + a, B.v.data() + a, B.up.data() + a, B.uf.data() + a, n,
                        ORC_LAPLACE_FD4, q.ls, &q.elbo, nullptr);
    }
    return q;
}

// profile_t0_refine + profile_refine_all + magnitude chi2 of one band.
void finalize(const Batch& B, int bi, int m, const std::vector<double>& pso, Posterior& q, int flags,
              double& mag_chi2) {
    const int64_t a = B.off[(size_t)bi], n = B.off[(size_t)bi + 1] - a;
    const int64_t off[2] = {0, n};
    double p7[7] = {0}, mu[7], ls[7];
    for (size_t d = 0; d < pso.size(); d++) p7[d] = pso[d];
    int32_t mid = m;
    orc_finalize(1, off, B.t.data() + a, B.f.data() + a, B.v.data() + a, B.up.data() + a, B.uf.data() + a,
                 B.fe.data() + a, B.mag.data() + a, &B.peak[(size_t)bi], &mid, p7, q.mu, q.ls, flags, mu, ls,
                 &mag_chi2, 1);
    for (int d = 0; d < 7; d++) {
        q.mu[d] = mu[d];
        q.ls[d] = ls[d];
    }
}

void put(double* rec, int model, const std::vector<double>& pso, double chi2, const Posterior& q, int n,
         double mag_chi2, int route) {
    rec[R_MODEL] = model;
    for (int d = 0; d < 7; d++) rec[R_PSO + d] = d < (int)pso.size() ? pso[(size_t)d] : 0.0;
    rec[R_PSO_CHI2] = std::isfinite(chi2) ? chi2 : NAN;
    for (int d = 0; d < 7; d++) {
        rec[R_MU + d] = q.mu[d];
        rec[R_LS + d] = q.ls[d];
    }
    rec[R_ELBO] = std::isfinite(q.elbo) ? q.elbo : NAN;
    rec[R_NOBS] = n;
    rec[R_MAG_CHI2] = std::isfinite(mag_chi2) ? mag_chi2 : NAN;
    rec[R_ROUTE] = route;
}

void put_mb(double* rec, const MbResult& r) {
    rec[R_MB_HAS] = r.has ? 1.0 : 0.0;
    if (!r.has) return;
    rec[R_MB_K] = r.best_k;
    for (int d = 0; d < 18; d++) rec[R_MB_P + d] = d < (int)r.params.size() ? r.params[(size_t)d] : 0.0;
    rec[R_MB_COST] = r.cost;
    rec[R_MB_BIC] = r.bic;
    for (int k = 0; k < 4; k++) {
        rec[R_MB_PKC + k] = r.pk_cost[k];
        rec[R_MB_PKB + k] = r.pk_bic[k];
    }
}

std::vector<FitBand> prepare(const int64_t* band_off, int64_t b0, int64_t b1, const double* t, const double* f,
                             const double* e) {
    std::vector<FitBand> out;
    for (int64_t b = b0; b < b1; b++) {
        FitBand fb;
        fb.input = (int)(b - b0);
        const int64_t a = band_off[b];
        if (band_fit_data(t + a, f + a, e + a, (int)(band_off[b + 1] - a), fb)) out.push_back(std::move(fb));
    }
    std::stable_sort(out.begin(), out.end(), [](const FitBand& x, const FitBand& y) { return x.n() > y.n(); });
    return out;
}

// fit_parametric_multiband (src/parametric.rs:3813-4131); false when it
// yields nothing.
bool multiband(const std::vector<FitBand>& bands, int method, double* recs) {
    const int nb = (int)bands.size();
    if (nb == 0 || nb > 9) return false;
    Batch B;
    for (const FitBand& b : bands) B.add(b);
    const int D = 7 * nb;
    double lo[64], hi[64], pc[7], pw[7];
    pso_bounds_cpu(1, lo, hi);
    for (int b = 1; b < nb; b++)
        for (int j = 0; j < 7; j++) {
            lo[7 * b + j] = -3.0 * REL_SIG[j];
            hi[7 * b + j] = 3.0 * REL_SIG[j];
        }
    double peak_t, t_first;
    peak_and_first(bands[0], peak_t, t_first);
    const double t0_lo = std::max(t_first - 30.0, lo[3]), t0_hi = std::min(peak_t + 10.0, hi[3]);
    if (t0_lo < t0_hi) {
        lo[3] = t0_lo;
        hi[3] = t0_hi;
    }
    population_priors(1, pc, pw);
    double x0[64];
    for (int j = 0; j < 7; j++) x0[j] = clampd(pc[j], lo[j], hi[j]);
    x0[3] = clampd(peak_t - 5.0, lo[3], hi[3]);
    for (int b = 1; b < nb; b++)
        for (int j = 0; j < 7; j++) x0[7 * b + j] = clampd(REL_MEAN[j], lo[7 * b + j], hi[7 * b + j]);
    Fit best;
    for (int r = 0; r < 3; r++) {
        orc_pso_job j = make_job(ORC_CPSO_SWARM_JOINT, 1, 0, D, lo, hi, 200, 1500, 60, 1, SEEDS[r]);
        j.n_bands = nb;
        if (r == 1) {
            j.seeded = 1;
            for (int d = 0; d < D; d++) j.x0[d] = x0[d];
        }
        JobOut o = run_job(B, j);
        if (o.cost < best.cost) {
            best.cost = o.cost;
            best.p = o.x;
        }
    }
    if (best.p.empty()) return false;
    for (int b = 0; b < nb; b++) {
        std::vector<double> p(7);
        for (int j = 0; j < 7; j++) p[(size_t)j] = b == 0 ? best.p[(size_t)j] : best.p[(size_t)j] + best.p[(size_t)(7 * b + j)];
        Posterior q = posterior(B, b, 1, p, method);
        double mc;
        finalize(B, b, 1, p, q, ORC_FINALIZE_T0 | ORC_FINALIZE_ALL | ORC_FINALIZE_MAG_CHI2 | ORC_FINALIZE_SVI_CHI2,
                 mc);
        double* rec = recs + (size_t)bands[(size_t)b].input * ORC_FP_RECORD;
        rec[R_RANK] = b;
        put(rec, 1, p, best.cost, q, bands[(size_t)b].n(), mc, 1);
        rec[R_HAS + 1] = 1.0;
        rec[R_CHI2 + 1] = std::isfinite(best.cost) ? best.cost : NAN;
        if (b == 0) put_mb(rec, fit_multi_bazin(B, bands[0], 0));
    }
    return true;
}

// fit_parametric_model (src/parametric.rs:3563-3739).
void per_band(const std::vector<FitBand>& bands, bool fit_all, int method, int forced, double* recs) {
    Batch B;
    for (const FitBand& b : bands) B.add(b);
    int rank = 0;
    for (int bi = 0; bi < (int)bands.size(); bi++) {
        const FitBand& b = bands[(size_t)bi];
        Selection s;
        if (forced >= 0) {
            for (int k = 0; k < 8; k++) {
                s.chi2[k] = NAN;
                s.has[k] = 0;
            }
            s.model = forced;
            s.fit = single_model(B, b, bi, forced);
            s.chi2[forced] = std::isfinite(s.fit.cost) ? s.fit.cost : NAN;
            s.has[forced] = 1;
            if (fit_all) s.params[forced] = s.fit;
        } else {
            s = model_select(B, b, bi, fit_all);
        }
        if (s.fit.p.empty()) continue;
        Fit pso = s.fit;
        if (bi == 0 && method == 0) pso = profile_t0_search(B, b, bi, s.model, pso);
        Posterior q = posterior(B, bi, s.model, pso.p, method);
        double mc;
        finalize(B, bi, s.model, pso.p, q, ORC_FINALIZE_T0 | ORC_FINALIZE_ALL | ORC_FINALIZE_MAG_CHI2, mc);
        double* rec = recs + (size_t)b.input * ORC_FP_RECORD;
        rec[R_RANK] = rank++;
        put(rec, s.model, pso.p, pso.cost, q, b.n(), mc, 0);
        for (int k = 0; k < 8; k++) {
            rec[R_HAS + k] = s.has[k];
            rec[R_CHI2 + k] = s.chi2[k];
            if (fit_all)
                for (int d = 0; d < 7; d++)
                    rec[R_PARAMS + 7 * k + d] = d < (int)s.params[k].p.size() ? s.params[k].p[(size_t)d] : 0.0;
        }
        if (bi == 0) put_mb(rec, fit_multi_bazin(B, b, bi));
    }
}

}  // namespace

}  // namespace orc

extern "C" int orc_fit_parametric(int64_t n_src, const int64_t* src_off, const int64_t* band_off, const double* t,
                                  const double* flux, const double* err, int fit_all_models, int method,
                                  int forced_model, double* out, int n_threads) {
    if (n_src < 0 || (method != 0 && method != 1) || forced_model < -1 || forced_model > 7) return -1;
    const int64_t nb = src_off[n_src];
    for (int64_t k = 0; k < nb * ORC_FP_RECORD; k++) out[k] = NAN;
    for (int64_t b = 0; b < nb; b++) {
        double* r = out + b * ORC_FP_RECORD;
        r[orc::R_RANK] = -1.0;
        for (int k = 0; k < 8; k++) r[orc::R_HAS + k] = 0.0;
        r[orc::R_MB_HAS] = 0.0;
    }
    orc::parallel_for(n_src, n_threads, [&](int64_t s) {
        const int64_t b0 = src_off[s], b1 = src_off[s + 1];
        double* recs = out + b0 * ORC_FP_RECORD;
        int with_data = 0;
        for (int64_t b = b0; b < b1; b++) with_data += band_off[b + 1] > band_off[b];
        std::vector<orc::FitBand> bands = orc::prepare(band_off, b0, b1, t, flux, err);
        if (with_data >= 2 && !fit_all_models && forced_model < 0 && orc::multiband(bands, method, recs)) return;
        orc::per_band(bands, fit_all_models != 0, method, forced_model, recs);
    });
    return 0;
}
