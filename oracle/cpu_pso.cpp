// cpu_pso.cpp -- oracle of the CPU path's swarm fits.
//
// TEST INFRASTRUCTURE ONLY (parity checker for the HIP product).
//
//   pso_minimize / pso_minimize_seeded (src/parametric.rs:1160-1266,
//   1646-1757): w 0.9 -> 0.4 linearly, c1 = c2 = 1.5, v_max = half the box,
//   v0 = 0.2 v_max (2u - 1), particle 0 at the seed when seeded, clamp of
//   the velocity, wall absorption, the global best updated after every
//   particle (asynchronous), stall when the relative gain is <= 1 % for
//   stall_iters iterations (baseline = the initial best);
//   PsoCost::cost_from_slice (:1303-1370), optionally with t0 held fixed
//   (profile_t0_search, :1767-1849); the joint Villar cost of
//   fit_parametric_multiband (:3905-4000); lbfgs_refine (:1453-1497) with
//   PsoCost::gradient (:1383-1446) projected at the bounds.
// The reference draws from SmallRng; this restatement and the device use
// Philox4x32-10 keyed by the job's seed (the reference stream lives in the
// rand crate and is not reproducible here).  argmin's More-Thuente line
// search is replaced by backtracking Armijo (c1 = 1e-4, halving, <= 20
// trials).  Sums over observations are formed as the device forms them: 64
// strided partial sums (observation i into partial i mod 64, in order),
// reduced by the xor butterfly over offsets 32, 16, ..., 1; dot products of
// search vectors likewise over their components.
#include "lcf_oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace orc {

namespace {

const int T0_IDX[8] = {2, 3, 2, 1, 1, 1, 1, 3};
const double LN10 = 2.30258509299404568402;
const double REL_MEAN[7] = {-0.015 * LN10, -0.001, -0.095 * LN10, -1.181, -0.195 * LN10, -0.323 * LN10,
                            -0.025 * LN10};
const double REL_SIG[7] = {0.08 * LN10, 0.008, 0.25 * LN10, 1.5, 0.35 * LN10, 0.25 * LN10, 0.30 * LN10};

double bfly(const double* v) {
    double a[64], b[64];
    for (int l = 0; l < 64; l++) a[l] = v[l];
    for (int off = 32; off >= 1; off >>= 1) {
        for (int l = 0; l < 64; l++) b[l] = a[l] + a[l ^ off];
        for (int l = 0; l < 64; l++) a[l] = b[l];
    }
    return a[0];
}

struct Obs {
    const double* t;
    const double* y;  // flux, or the upper-limit flux for upper limits
    const double* v;
    const int32_t* up;
    int n;
};

double villar_constraint(double lg, double beta, double ltr, double ltf) {
    const double gamma = std::exp(lg), tr = std::exp(ltr), tf = std::exp(ltf);
    // f64::max ignores a NaN operand, as fmax does
    const double c1 = std::fmax(gamma * beta - 1.0, 0.0);
    const double c2 = std::fmax(std::exp(-gamma / tr) * (tf / tr - 1.0) - 1.0, 0.0);
    const double c3 = std::fmax(beta * tf - 1.0 + beta * gamma, 0.0);
    return c1 + c2 + c3;
}

// NLL sum of one band (NaN when a prediction is non-finite).
double band_nll(int m, const double* p, const Obs& o) {
    const int np = n_params(m);
    const double se = std::exp(p[np - 1]);
    const double se2 = se * se;
    std::vector<double> pred(std::max(o.n, 1));
    eval_times(m, p, o.t, o.n, pred.data());
    if (m == 7) {
        double mx = -INFINITY;
        for (int i = 0; i < o.n; i++)
            if (!o.up[i]) mx = std::fmax(mx, pred[i]);
        if (mx > 1e-10 && std::isfinite(mx)) {
            const double sc = std::min(std::max(1.0 / mx, 0.1), 10.0);
            for (int i = 0; i < o.n; i++) pred[i] *= sc;
        }
    }
    double part[64] = {0};
    for (int i = 0; i < o.n; i++) {
        if (!std::isfinite(pred[i])) return NAN;
        const double tv = o.v[i] + se2;
        double& s = part[i & 63];
        if (o.up[i]) s -= log_normal_cdf((o.y[i] - pred[i]) / std::sqrt(tv));
        else {
            const double d = pred[i] - o.y[i];
            s += d * d / tv + std::log(tv);
        }
    }
    return bfly(part);
}

double pso_cost(int m, const double* p, const Obs& o, bool prior) {
    if (m == 1) {
        const double c = villar_constraint(p[2], p[1], p[4], p[5]);
        if (c > 0.0) return 1e12 + 10000.0 * c;
    }
    double nll = band_nll(m, p, o);
    if (std::isnan(nll)) return 1e99;
    const double nn = (double)std::max(o.n, 1);
    if (prior) {
        double pc[7], pw[7], lp = 0.0;
        population_priors(m, pc, pw);
        for (int j = 0; j < n_params(m); j++)
            if (pw[j] > 0.0) {
                const double z = (p[j] - pc[j]) / pw[j];
                lp += 0.5 * z * z;
            }
        nll += lp / nn;
    }
    return nll / nn;
}

struct JobCtx {
    const orc_pso_job* J;
    const int64_t* off;
    const double *t, *y, *v;
    const int32_t* up;
};

Obs band_obs(const JobCtx& c, int b) {
    const int64_t a = c.off[b];
    return Obs{c.t + a, c.y + a, c.v + a, c.up + a, (int)(c.off[b + 1] - a)};
}

// multi_bazin_cost (src/parametric.rs:983-1006): K components as written,
// the device's strided partial sums and butterfly.
double multi_bazin_job_cost(int K, const double* xs, const Obs& o) {
    const double se = std::exp(xs[4 * K + 1]);
    const double se2 = se * se;
    double part[64] = {0};
    for (int i = 0; i < o.n; i++) {
        double f = 0.0;
        for (int c = 0; c < K; c++) {
            const double a = std::exp(xs[4 * c]), tr = std::exp(xs[4 * c + 2]), tf = std::exp(xs[4 * c + 3]);
            const double dt = o.t[i] - xs[4 * c + 1];
            const double e_fall = std::exp(-dt / tf);
            const double sig = 1.0 / (1.0 + std::exp(-dt / tr));
            f += a * e_fall * sig;
        }
        f += xs[4 * K];
        if (!std::isfinite(f)) return 1e99;
        const double tv = o.v[i] + se2;
        double& s = part[i & 63];
        if (o.up[i]) s -= log_normal_cdf((o.y[i] - f) / std::sqrt(tv));
        else {
            const double d = f - o.y[i];
            s += d * d / tv + std::log(tv);
        }
    }
    return bfly(part) / (double)std::max(o.n, 1);
}

double job_cost(const JobCtx& c, const double* xs) {
    const orc_pso_job& J = *c.J;
    if (J.kind == ORC_CPSO_SWARM_MULTIBAZIN) return multi_bazin_job_cost(J.model, xs, band_obs(c, J.band));
    if (J.kind == ORC_CPSO_SWARM_JOINT) {
        const int B = J.n_bands;
        const int n_tot = (int)(c.off[J.band + B] - c.off[J.band]);
        const double* ref = xs;
        const double rc = villar_constraint(ref[2], ref[1], ref[4], ref[5]);
        if (rc > 0.0) return 1e12 + 10000.0 * rc;
        double bal = 0.0;
        for (int b = 0; b < B; b++) {
            double p[7];
            for (int j = 0; j < 7; j++) p[j] = b == 0 ? ref[j] : ref[j] + xs[7 * b + j];
            if (b > 0) {
                const double cc = villar_constraint(p[2], p[1], p[4], p[5]);
                if (cc > 0.0) return 1e12 + 10000.0 * cc;
            }
            Obs o = band_obs(c, J.band + b);
            const double nll = band_nll(1, p, o);
            if (std::isnan(nll)) return 1e99;
            if (o.n > 0) bal += (nll / (double)o.n) * ((double)n_tot / (double)B);
        }
        const double nt = (double)std::max(n_tot, 1);
        double pc[7], pw[7], lp = 0.0;
        population_priors(1, pc, pw);
        for (int j = 0; j < 7; j++)
            if (pw[j] > 0.0) {
                const double z = (ref[j] - pc[j]) / pw[j];
                lp += 0.5 * z * z;
            }
        for (int b = 1; b < B; b++)
            for (int j = 0; j < 7; j++)
                if (REL_SIG[j] > 0.0) {
                    const double z = (xs[7 * b + j] - REL_MEAN[j]) / REL_SIG[j];
                    lp += 0.5 * z * z;
                }
        return (bal + lp) / nt;
    }
    const int m = J.model, np = n_params(m), t0i = T0_IDX[m];
    double p[7];
    for (int j = 0; j < np; j++)
        p[j] = !J.fixed_t0 ? xs[j] : (j < t0i ? xs[j] : (j == t0i ? J.t0_value : xs[j - 1]));
    return pso_cost(m, p, band_obs(c, J.band), J.use_prior != 0);
}

void draw(const orc_pso_job& J, uint32_t slot, int p, int d, double& u1, double& u2) {
    const uint32_t tag = (uint32_t)J.model | ((uint32_t)J.kind << 8) | (J.fixed_t0 ? 0x10000u : 0u);
    const uint32_t ctr[4] = {J.stream, tag, slot, (uint32_t)p * 32u + (uint32_t)(d >> 1)};
    const uint32_t key[2] = {(uint32_t)J.seed, (uint32_t)(J.seed >> 32)};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    const double s = 1.0 / 16777216.0;
    u1 = (double)(((d & 1) ? r[2] : r[0]) >> 8) * s;
    u2 = (double)(((d & 1) ? r[3] : r[1]) >> 8) * s;
}

void swarm(const JobCtx& c, double* out_x, double* out_cost, int32_t* out_it) {
    const orc_pso_job& J = *c.J;
    const int D = J.dim, P = J.n_particles;
    std::vector<double> pos(P * D), vel(P * D), pb(P * D), pbc(P), vmax(D);
    for (int d = 0; d < D; d++) vmax[d] = 0.5 * (J.hi[d] - J.lo[d]);
    double gb[64] = {0}, gcost = INFINITY, xs[64] = {0};
    for (int p = 0; p < P; p++) {
        for (int d = 0; d < D; d++) {
            double u1, u2;
            draw(J, 0u, p, d, u1, u2);
            const double x = (J.seeded && p == 0) ? J.x0[d] : J.lo[d] + u1 * (J.hi[d] - J.lo[d]);
            pos[p * D + d] = x;
            vel[p * D + d] = vmax[d] * 0.2 * (2.0 * u2 - 1.0);
            pb[p * D + d] = x;
            xs[d] = x;
        }
        const double cst = job_cost(c, xs);
        pbc[p] = cst;
        if (cst < gcost) {
            gcost = cst;
            for (int d = 0; d < D; d++) gb[d] = xs[d];
        }
    }
    double prev = gcost;
    int stall = 0, it = 0;
    const double inv = 1.0 / (double)J.max_iters;
    while (it < J.max_iters) {
        const double w = 0.9 - (0.9 - 0.4) * (double)it * inv;
        for (int p = 0; p < P; p++) {
            for (int d = 0; d < D; d++) {
                double u1, u2;
                draw(J, (uint32_t)it + 1u, p, d, u1, u2);
                const double x = pos[p * D + d], v = vel[p * D + d], b = pb[p * D + d];
                double vn = w * v + 1.5 * u1 * (b - x) + 1.5 * u2 * (gb[d] - x);
                vn = vn < -vmax[d] ? -vmax[d] : (vn > vmax[d] ? vmax[d] : vn);
                const double np_ = x + vn;
                if (np_ <= J.lo[d]) { pos[p * D + d] = J.lo[d]; vel[p * D + d] = 0.0; }
                else if (np_ >= J.hi[d]) { pos[p * D + d] = J.hi[d]; vel[p * D + d] = 0.0; }
                else { pos[p * D + d] = np_; vel[p * D + d] = vn; }
                xs[d] = pos[p * D + d];
            }
            const double cst = job_cost(c, xs);
            if (cst < pbc[p]) {
                pbc[p] = cst;
                for (int d = 0; d < D; d++) pb[p * D + d] = xs[d];
                if (cst < gcost) {
                    gcost = cst;
                    for (int d = 0; d < D; d++) gb[d] = xs[d];
                }
            }
        }
        it++;
        const bool improved = prev - gcost > 0.01 * std::fmax(std::fabs(prev), 1e-10);
        if (improved) {
            stall = 0;
            prev = gcost;
        } else if (++stall >= J.stall_iters) {
            break;
        }
    }
    for (int d = 0; d < 64; d++) out_x[d] = d < D ? gb[d] : 0.0;
    *out_cost = gcost;
    *out_it = it;
}

// PsoCost::gradient at p (clamped already) with the device's sum order.
void pso_grad(int m, const double* p, const Obs& o, double* g) {
    const double PI = 3.14159265358979323846;
    const int np = n_params(m), se_idx = np - 1;
    const double se = std::exp(p[se_idx]), se2 = se * se;
    double acc[7][64] = {{0}};
    bool bad = false;
    double mg[7];
    for (int i = 0; i < o.n; i++) {
        const double pred = eval_scalar(m, p, o.t[i]);
        if (!std::isfinite(pred)) { bad = true; continue; }
        eval_grad(m, p, o.t[i], mg);
        const double tv = o.v[i] + se2;
        const int l = i & 63;
        if (o.up[i]) {
            const double sv = std::sqrt(tv);
            const double z = (o.y[i] - pred) / sv;
            const double phi = std::exp(-0.5 * z * z) / std::sqrt(2.0 * PI);
            const double cdf = 0.5 * (1.0 + erf_approx(z / 1.41421356237309504880));
            const double ratio = phi / std::max(cdf, 1e-300);
            for (int j = 0; j < np; j++) acc[j][l] += ratio * mg[j] / sv;
            acc[se_idx][l] += ratio * z * se2 / tv;
        } else {
            const double d = pred - o.y[i];
            const double sc = 2.0 * d / tv;
            for (int j = 0; j < np; j++) acc[j][l] += sc * mg[j];
            acc[se_idx][l] += (-d * d / (tv * tv) + 1.0 / tv) * (2.0 * se2);
        }
    }
    const double nn = (double)std::max(o.n, 1);
    for (int j = 0; j < np; j++) g[j] = bad ? 0.0 : bfly(acc[j]) / nn;
}

double dot(const double* u, const double* v, int D) {
    double part[64] = {0};
    for (int d = 0; d < D; d++) part[d] = u[d] * v[d];
    return bfly(part);
}

void polish(const JobCtx& c, double* out_x, double* out_cost, int32_t* out_it) {
    const orc_pso_job& J = *c.J;
    const int D = J.dim, m = J.model;
    const Obs o = band_obs(c, J.band);
    auto clampv = [&](const double* x, double* xc) {
        for (int d = 0; d < D; d++) xc[d] = std::fmin(std::fmax(x[d], J.lo[d]), J.hi[d]);
    };
    auto bcost = [&](const double* x) {
        double xc[64] = {0};
        clampv(x, xc);
        return pso_cost(m, xc, o, J.use_prior != 0);
    };
    auto bgrad = [&](const double* x, double* g) {
        double xc[64] = {0};
        clampv(x, xc);
        pso_grad(m, xc, o, g);
        for (int d = 0; d < D; d++)
            if ((x[d] <= J.lo[d] && g[d] > 0.0) || (x[d] >= J.hi[d] && g[d] < 0.0)) g[d] = 0.0;
    };
    const int M = 7;
    double x[64], g[64], bx[64], S[7][64], Y[7][64], rho[7];
    for (int d = 0; d < D; d++) x[d] = J.x0[d];
    double f = bcost(x);
    bgrad(x, g);
    for (int d = 0; d < D; d++) bx[d] = x[d];
    double bf = f;
    int nm = 0, head = 0, it = 0;
    for (; it < J.max_iters; it++) {
        if (std::sqrt(dot(g, g, D)) < 1e-6) break;
        double q[64], al[7];
        for (int d = 0; d < D; d++) q[d] = g[d];
        for (int k = 0; k < nm; k++) {
            const int pk = (head - 1 - k + M) % M;
            al[k] = rho[pk] * dot(S[pk], q, D);
            for (int d = 0; d < D; d++) q[d] -= al[k] * Y[pk][d];
        }
        if (nm > 0) {
            const int pk = (head - 1 + M) % M;
            const double sc = dot(S[pk], Y[pk], D) / dot(Y[pk], Y[pk], D);
            for (int d = 0; d < D; d++) q[d] *= sc;
        }
        for (int k = nm - 1; k >= 0; k--) {
            const int pk = (head - 1 - k + M) % M;
            const double b = rho[pk] * dot(Y[pk], q, D);
            for (int d = 0; d < D; d++) q[d] += S[pk][d] * (al[k] - b);
        }
        double dd[64];
        for (int d = 0; d < D; d++) dd[d] = -q[d];
        double gd = dot(g, dd, D);
        if (!(gd < 0.0)) {
            nm = 0;
            for (int d = 0; d < D; d++) dd[d] = -g[d];
            gd = dot(g, dd, D);
        }
        double step = 1.0, xt[64], ft = f;
        bool ok = false;
        for (int ls = 0; ls < 20; ls++) {
            for (int d = 0; d < D; d++) xt[d] = x[d] + step * dd[d];
            ft = bcost(xt);
            if (std::isfinite(ft) && ft <= f + 1e-4 * step * gd) { ok = true; break; }
            step *= 0.5;
        }
        if (!ok) break;
        double gt[64], s[64], yv[64];
        bgrad(xt, gt);
        for (int d = 0; d < D; d++) { s[d] = xt[d] - x[d]; yv[d] = gt[d] - g[d]; }
        const double sy = dot(s, yv, D);
        if (sy > 1e-10) {
            for (int d = 0; d < D; d++) { S[head][d] = s[d]; Y[head][d] = yv[d]; }
            rho[head] = 1.0 / sy;
            head = (head + 1) % M;
            nm = std::min(nm + 1, M);
        }
        for (int d = 0; d < D; d++) { x[d] = xt[d]; g[d] = gt[d]; }
        f = ft;
        if (f < bf) {
            bf = f;
            for (int d = 0; d < D; d++) bx[d] = x[d];
        }
    }
    double xc[64];
    clampv(bx, xc);
    const double fc = bcost(xc);
    const bool keep = fc < J.start_cost && std::isfinite(fc);
    for (int d = 0; d < 64; d++) out_x[d] = d < D ? (keep ? xc[d] : J.x0[d]) : 0.0;
    *out_cost = keep ? fc : J.start_cost;
    *out_it = it;
}

}  // namespace

}  // namespace orc

extern "C" int orc_cpu_pso(const int64_t* off, const double* t, const double* flux, const double* obs_var,
                           const int32_t* is_upper, const double* upper_flux, int64_t n_obs,
                           const orc_pso_job* jobs, int64_t n_jobs, double* out_x, double* out_cost,
                           int32_t* out_iters, int n_threads) {
    std::vector<double> y((size_t)std::max<int64_t>(n_obs, 1));
    for (int64_t i = 0; i < n_obs; i++) y[(size_t)i] = is_upper[i] ? upper_flux[i] : flux[i];
    for (int64_t k = 0; k < n_jobs; k++) {
        const orc_pso_job& J = jobs[k];
        if (J.dim <= 0 || J.dim > 64 || J.kind < 0 || J.kind > 3) return -1;
        if (J.kind == ORC_CPSO_SWARM_MULTIBAZIN) {
            if (J.model < 1 || J.model > 4 || J.dim != 4 * J.model + 2) return -1;
        } else if (J.kind != ORC_CPSO_SWARM_JOINT && orc::n_params(J.model) < 0) {
            return -1;
        }
    }
    orc::parallel_for(n_jobs, n_threads, [&](int64_t k) {
        orc::JobCtx c{&jobs[k], off, t, y.data(), obs_var, is_upper};
        if (jobs[k].kind == ORC_CPSO_POLISH) orc::polish(c, out_x + k * 64, out_cost + k, out_iters + k);
        else orc::swarm(c, out_x + k * 64, out_cost + k, out_iters + k);
    });
    return 0;
}
