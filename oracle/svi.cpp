// svi.cpp -- oracle restatement of the batched SVI fit of the reference's GPU
// path (cuda/svi.cu:48-330, host src/gpu_cuda.rs:1315-1422).
//
// TEST INFRASTRUCTURE ONLY (parity checker for the HIP product).
//
// Per job (one band x one model): mean-field Gaussian guide over the
// internal parameters, mu initialised at the PSO point, log sigma at -1;
// each step draws n_samples reparameterised samples theta = mu + sigma*eps,
// accumulates the Gaussian / censored log-likelihood and its gradient with the
// analytic model gradient (forward differences for MetzgerKN), adds the
// Gaussian prior, forms the ELBO (+ entropy), and takes an Adam step on
// (mu, log sigma) with log sigma clamped to [-6, 2].  Early stop on an EMA of
// the ELBO (min 50 steps, stall 20).
//
// Noise streams:
//   ORC_SVI_RNG_REFERENCE: the reference's xorshift64 seeded 42 + job*123456789,
//     Box-Muller (cos branch) per parameter, samples in order (svi.cu:17-34,124).
//   ORC_SVI_RNG_PHILOX: the stream the HIP kernels use: Philox4x32-10 keyed by
//     the seed, counter (job, step, sample, pair); u1 = ((x>>9)+1/2) 2^-23,
//     u2 = (y>>8) 2^-24, eps[2q] = r cos(2 pi u2), eps[2q+1] = r sin(2 pi u2).
#include "lcf_oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace orc {

static const double PI = 3.14159265358979323846;

// Forward-difference model gradient (cuda/svi.cu:180-192): h = max(1e-5, 1e-5|theta_j|),
// sigma_x slot zero.
static void fd_grad(int m, const double* th, int np, int se, double t, double pred, double* g) {
    double tp[7];
    for (int k = 0; k < np; k++) tp[k] = th[k];
    for (int j = 0; j < np; j++) {
        if (j == se) { g[j] = 0.0; continue; }
        double h = std::max(1e-5, 1e-5 * std::fabs(th[j]));
        tp[j] = th[j] + h;
        double pp = (m == 7) ? metzger_single(tp, t) : eval_scalar(m, tp, t);
        tp[j] = th[j];
        g[j] = std::isfinite(pp) ? (pp - pred) / h : 0.0;
    }
    g[se] = 0.0;
}

// Log-likelihood of one sample and its gradient w.r.t. theta
// (cuda/svi.cu:171-226).
static double loglik_grad(int m, const double* th, int np, int se, const double* t, const double* f,
                          const double* v, const int32_t* up, const double* uf, int64_t n,
                          int grad_mode, double* dll) {
    for (int j = 0; j < np; j++) dll[j] = 0.0;
    double se2 = std::exp(th[se]);
    se2 *= se2;
    double ll = 0.0, g[7];
    for (int64_t i = 0; i < n; i++) {
        double pred = (m == 7) ? metzger_single(th, t[i]) : eval_scalar(m, th, t[i]);
        if (!std::isfinite(pred)) continue;
        if (m == 7 || (grad_mode & 1)) fd_grad(m, th, np, se, t[i], pred, g);
        else eval_grad(m, th, t[i], g);
        double tv = v[i] + se2, st = std::sqrt(tv), inv = 1.0 / tv;
        if (up[i]) {
            double z = (uf[i] - pred) / st;
            ll += log_normal_cdf(z);
            // d ln Phi / dz: the reference's phi / max(Phi, 1e-300) with Phi
            // from erf_approx, or (grad_mode bit 2) the exact derivative of
            // the log_normal_cdf summed above
            double ratio, dpred;
            if (grad_mode & 2) {
                ratio = dlog_normal_cdf(z);
                dpred = -ratio / st;
            } else {
                double phi = std::exp(-0.5 * z * z) / std::sqrt(2.0 * PI);
                double cdf = std::max(0.5 * (1.0 + erf_approx(z * 0.70710678118654752440)), 1e-300);
                ratio = phi / cdf;
                dpred = -phi / (cdf * st);
            }
            for (int j = 0; j < np; j++)
                if (j != se && std::isfinite(g[j])) dll[j] += dpred * g[j];
            double dz = -(uf[i] - pred) * se2 / (st * tv);
            dll[se] += ratio * dz;
        } else {
            double r = f[i] - pred, r2 = r * r;
            ll += -0.5 * (r2 * inv + std::log(2.0 * PI * tv));
            for (int j = 0; j < np; j++)
                if (j != se && std::isfinite(g[j])) dll[j] += r * inv * g[j];
            dll[se] += (r2 * inv * inv - inv) * se2;
        }
    }
    return ll;
}

struct XorShift {
    uint64_t s;
    uint64_t next() {
        uint64_t x = s;
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        s = x;
        return x;
    }
    double uniform() { return (double)(next() & 0x1FFFFFFFFFFFFFull) / (double)0x1FFFFFFFFFFFFFull; }
    double normal() {
        double u1 = std::max(uniform(), 1e-10), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * PI * u2);
    }
};

void philox_normals(uint64_t seed, uint32_t job, uint32_t step, uint32_t sample, int np, double* eps) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int q = 0; q < (np + 1) / 2; q++) {
        uint32_t c[4] = {job, step, sample, (uint32_t)q}, r[4];
        orc_philox4x32_10(c, key, r);
        double u1 = ((double)(r[0] >> 9) + 0.5) * (1.0 / 8388608.0);
        double u2 = (double)(r[1] >> 8) * (1.0 / 16777216.0);
        double rad = std::sqrt(-2.0 * std::log(u1));
        eps[2 * q] = rad * std::cos(2.0 * PI * u2);
        if (2 * q + 1 < np) eps[2 * q + 1] = rad * std::sin(2.0 * PI * u2);
    }
}

// One Monte-Carlo ELBO estimate and its gradient for given noise
// eps[S][np] (the body of one step of cuda/svi.cu:140-244).
double elbo_grad_step(int m, int np, int se, const double* mu, const double* ls, const double* pc,
                      const double* pw, const double* eps, int S, const double* t, const double* f,
                      const double* v, const int32_t* up, const double* uf, int64_t n, int grad_mode,
                      double* gmu, double* gls) {
    double sig[7], th[7], dll[7];
    for (int j = 0; j < np; j++) { sig[j] = std::exp(ls[j]); gmu[j] = 0.0; gls[j] = 0.0; }
    double esum = 0.0;
    for (int s = 0; s < S; s++) {
        for (int j = 0; j < np; j++) th[j] = mu[j] + sig[j] * eps[s * np + j];
        double ll = loglik_grad(m, th, np, se, t, f, v, up, uf, n, grad_mode, dll);
        double lp = 0.0;
        for (int j = 0; j < np; j++) {
            double d = th[j] - pc[j], w2 = pw[j] * pw[j];
            lp += -0.5 * d * d / w2;
            double df = dll[j] - d / w2;
            gmu[j] += df;
            gls[j] += df * sig[j] * eps[s * np + j];
        }
        esum += ll + lp;
    }
    for (int j = 0; j < np; j++) { gmu[j] /= S; gls[j] /= S; }
    double ent = 0.0;
    for (int j = 0; j < np; j++) ent += ls[j];
    ent += 0.5 * np * std::log(2.0 * PI * 2.71828182845904523536);
    return esum / S + ent;
}

struct SviOut {
    double mu[7], ls[7], elbo;
    int steps;
    double gmax;  // largest |gradient| handed to Adam over the fit (inf if any was non-finite)
};

static SviOut svi_job(int m, int64_t job, const double* pso, const double* pc, const double* pw,
                      const double* t, const double* f, const double* v, const int32_t* up,
                      const double* uf, int64_t n, int n_steps, int S, double lr, int rng_mode,
                      uint64_t seed, int grad_mode, uint32_t flags) {
    SviOut o;
    const int np = n_params(m), se = np - 1;
    for (int j = 0; j < 7; j++) { o.mu[j] = 0.0; o.ls[j] = 0.0; }
    o.steps = 0;
    o.gmax = 0.0;
    if (n <= 0) { o.elbo = -1e99; return o; }
    double mu[7], ls[7], am[14] = {0}, av[14] = {0}, gmu[7], gls[7];
    for (int j = 0; j < np; j++) { mu[j] = pso[j]; ls[j] = -1.0; }
    const int min_it = (flags & ORC_SVI_CPU_STOPPING) ? 200 : 50;
    const int stall_it = (flags & ORC_SVI_CPU_STOPPING) ? 50 : 20;
    XorShift xs{42ull + (uint64_t)job * 123456789ull};
    double final_elbo = -1e300, ema = -1e300, best = -1e300;
    int stall = 0, t_adam = 0;
    std::vector<double> epsv((size_t)S * np);
    for (int step = 0; step < n_steps; step++) {
        for (int s = 0; s < S; s++) {
            if (rng_mode == ORC_SVI_RNG_REFERENCE)
                for (int j = 0; j < np; j++) epsv[s * np + j] = xs.normal();
            else
                philox_normals(seed, (flags & ORC_SVI_SHARED_STREAM) ? 0u : (uint32_t)job, (uint32_t)step,
                               (uint32_t)s, np, &epsv[s * np]);
        }
        final_elbo = elbo_grad_step(m, np, se, mu, ls, pc, pw, epsv.data(), S, t, f, v, up, uf, n,
                                    grad_mode, gmu, gls);
        o.steps = step + 1;
        // Adam (cuda/svi.cu:266-291)
        t_adam++;
        double bc1 = 1.0 - std::pow(0.9, (double)t_adam), bc2 = 1.0 - std::pow(0.999, (double)t_adam);
        for (int j = 0; j < np; j++) {
            double ga = std::fabs(gmu[j]), gb = std::fabs(gls[j] + 1.0);
            o.gmax = std::max(o.gmax, std::isfinite(ga) ? ga : INFINITY);
            o.gmax = std::max(o.gmax, std::isfinite(gb) ? gb : INFINITY);
            double g = -gmu[j];
            if (std::isfinite(g)) {
                am[j] = 0.9 * am[j] + 0.1 * g;
                av[j] = 0.999 * av[j] + 0.001 * g * g;
                mu[j] -= lr * (am[j] / bc1) / (std::sqrt(av[j] / bc2) + 1e-8);
            }
            double gl = -(gls[j] + 1.0);
            if (std::isfinite(gl)) {
                am[np + j] = 0.9 * am[np + j] + 0.1 * gl;
                av[np + j] = 0.999 * av[np + j] + 0.001 * gl * gl;
                ls[j] -= lr * (am[np + j] / bc1) / (std::sqrt(av[np + j] / bc2) + 1e-8);
            }
        }
        for (int j = 0; j < np; j++) ls[j] = std::max(-6.0, std::min(2.0, ls[j]));
        // early stop (cuda/svi.cu:297-312)
        if (!(flags & ORC_SVI_FIXED_STEPS) && step >= min_it) {
            ema = (ema < -1e290) ? final_elbo : 0.1 * final_elbo + 0.9 * ema;
            if (ema - best > 0.01 * std::max(std::fabs(best), 1e-10)) { best = ema; stall = 0; }
            else if (++stall >= stall_it) break;
        }
    }
    const double infl = (flags & ORC_SVI_INFLATE_SIGMA) ? std::log(4.0) : 0.0;
    for (int j = 0; j < np; j++) { o.mu[j] = mu[j]; o.ls[j] = std::max(-6.0, std::min(2.0, ls[j])) + infl; }
    o.elbo = final_elbo;
    return o;
}

}  // namespace orc

extern "C" {

int orc_svi_normals(uint64_t seed, uint32_t job, uint32_t step, uint32_t sample, int np, double* eps) {
    if (np <= 0 || np > 7) return -1;
    orc::philox_normals(seed, job, step, sample, np, eps);
    return 0;
}

double orc_svi_elbo_grad(int m, const double* mu, const double* ls, const double* pc, const double* pw,
                         const double* eps, int S, const double* t, const double* f, const double* v,
                         const int32_t* up, const double* uf, int64_t n, int grad_mode, double* gmu,
                         double* gls) {
    int np = orc::n_params(m);
    if (np < 0 || S <= 0) return NAN;
    double e = orc::elbo_grad_step(m, np, np - 1, mu, ls, pc, pw, eps, S, t, f, v, up, uf, n, grad_mode, gmu, gls);
    for (int j = 0; j < np; j++) gls[j] += 1.0;  // d(entropy)/d(log sigma_j); the step folds it into Adam
    return e;
}

int orc_batch_svi(int64_t n_jobs, const int64_t* off, const double* t, const double* f, const double* v,
                  const int32_t* up, const double* uf, const int32_t* model_ids, const double* pso,
                  const double* pc, const double* pw, int n_steps, int n_samples, double lr, int rng_mode,
                  uint64_t seed, int grad_mode, uint32_t flags, double* out_mu, double* out_ls,
                  double* out_elbo, int32_t* out_steps, double* out_gmax, int n_threads) {
    if (n_jobs < 0 || n_steps < 0 || n_samples <= 0 || n_samples > 64) return -1;
    for (int64_t j = 0; j < n_jobs; j++)
        if (orc::n_params(model_ids[j]) < 0) return -1;
    orc::parallel_for(n_jobs, n_threads, [&](int64_t j) {
        int64_t a = off[j], b = off[j + 1];
        orc::SviOut o = orc::svi_job(model_ids[j], j, pso + j * 7, pc + j * 7, pw + j * 7, t + a, f + a,
                                     v + a, up + a, uf + a, b - a, n_steps, n_samples, lr, rng_mode, seed,
                                     grad_mode | ((flags & ORC_SVI_EXACT_TAIL) ? 2 : 0), flags);
        for (int d = 0; d < 7; d++) { out_mu[j * 7 + d] = o.mu[d]; out_ls[j * 7 + d] = o.ls[d]; }
        out_elbo[j] = o.elbo;
        if (out_steps) out_steps[j] = o.steps;
        if (out_gmax) out_gmax[j] = o.gmax;
    });
    return 0;
}

// svi_prior_for_model (src/parametric.rs:3259-3273): prior_params widths,
// centers moved to the PSO point except sigma_x.
int orc_svi_prior_for_model(int m, const double* pso, double* centers, double* widths) {
    int np = orc::n_params(m);
    if (np < 0) return -1;
    double c[7], w[7];
    orc::prior_params(m, c, w);
    for (int j = 0; j < np; j++) {
        centers[j] = (j == np - 1) ? c[j] : pso[j];
        widths[j] = w[j];
    }
    return 0;
}

}  // extern "C"
