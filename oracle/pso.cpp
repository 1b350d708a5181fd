// pso.cpp -- oracle restatement of the PSO cost, bounds, priors and the
// GPU-path batched PSO / model selection.
//
// TEST INFRASTRUCTURE ONLY (parity checker for the HIP product).
#include "lcf_oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// likelihood helpers: Abramowitz & Stegun 7.1.26 erfc, as in
// erf_approx / log_normal_cdf (src/parametric.rs:867-901).
// ---------------------------------------------------------------------------
static inline double as_poly(double t) {
    return t * (0.254829592 + t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429))));
}

double erf_approx(double x) {
    double a = std::fabs(x);
    double t = 1.0 / (1.0 + 0.3275911 * a);
    double val = 1.0 - as_poly(t) * std::exp(-a * a);
    return x >= 0.0 ? val : -val;
}

double log_normal_cdf(double x) {
    const double PI = 3.14159265358979323846;
    if (x > 8.0) return 0.0;
    if (x < -30.0) return -0.5 * x * x - 0.5 * std::log(2.0 * PI) - std::log(-x);
    double z = -x * 0.70710678118654752440;
    double t = 1.0 / (1.0 + 0.3275911 * std::fabs(z));
    double erfc_z = as_poly(t) * std::exp(-z * z);
    double phi = z >= 0.0 ? 0.5 * erfc_z : 1.0 - 0.5 * erfc_z;
    return std::log(std::max(phi, 1e-300));
}

// d/dx log_normal_cdf(x), differentiated from the form above (not from the
// exact normal CDF): the A&S polynomial Q(t) = t P(t), t = 1 / (1 + p |z|),
// z = -x / sqrt 2, with dQ/dt in Horner form; 0 above 8; -x - 1/x on the
// asymptote below -30.
double dlog_normal_cdf(double x) {
    const double S2 = 0.70710678118654752440, p = 0.3275911;
    if (x > 8.0) return 0.0;
    if (x < -30.0) return -x - 1.0 / x;
    const double z = -x * S2, az = std::fabs(z);
    const double t = 1.0 / (1.0 + p * az);
    const double q = as_poly(t);
    const double dq = 0.254829592 + t * (2.0 * -0.284496736 + t * (3.0 * 1.421413741 +
                      t * (4.0 * -1.453152027 + t * (5.0 * 1.061405429))));
    const double dt = -p * t * t;  // dt / d|z|
    if (z >= 0.0) {  // ln(0.5 Q) - z^2, z = |z|, dz/dx = -1/sqrt 2
        return -S2 * (dq * dt / q - 2.0 * az);
    }
    // ln(1 - 0.5 Q e^{-z^2}), |z| = x / sqrt 2
    const double e = std::exp(-az * az);
    const double E = 0.5 * q * e;
    const double dE = 0.5 * e * (dq * dt - 2.0 * az * q);
    return -S2 * dE / (1.0 - E);
}

// ---------------------------------------------------------------------------
// bounds and priors
// ---------------------------------------------------------------------------
// GpuModelName::pso_bounds (src/gpu_types.rs:64-103)
static const double GPU_LO[8][7] = {
    {-3.0, -1.0, -100.0, -2.0, -2.0, -5.0},
    {-1.5, -0.02, 1.0, -60.0, -1.0, 1.0, -6.0},
    {-3.0, -1.0, -100.0, -2.0, -1.0, 0.5, -5.0},
    {-3.0, -100.0, 0.5, -3.0, -5.0},
    {-3.0, -100.0, 0.0, 0.5, -5.0},
    {-3.0, -100.0, 0.1, -1.0, -5.0},
    {-3.0, -100.0, -2.0, -2.0, 0.5, -5.0},
    {-3.0, -2.0, -1.0, -2.0, -5.0}};
static const double GPU_HI[8][7] = {
    {3.0, 1.0, 100.0, 5.0, 6.0, 0.0},
    {1.5, 0.05, 5.0, 40.0, 3.0, 5.5, 0.0},
    {3.0, 1.0, 100.0, 5.0, 6.0, 4.0, 0.0},
    {3.0, 100.0, 4.5, 3.0, 0.0},
    {3.0, 100.0, 6.0, 4.5, 0.0},
    {3.0, 100.0, 3.0, 4.0, 0.0},
    {3.0, 100.0, 6.0, 3.0, 5.0, 0.0},
    {-0.5, -0.5, 2.0, 1.0, 0.0}};

void pso_bounds_gpu(int m, double* lo, double* hi) {
    int np = n_params(m);
    for (int d = 0; d < np; d++) { lo[d] = GPU_LO[m][d]; hi[d] = GPU_HI[m][d]; }
}

// pso_bounds (src/parametric.rs:1541-1579): identical to the GPU table except
// the Bazin baseline B, which is +-0.3 on the CPU path.
void pso_bounds_cpu(int m, double* lo, double* hi) {
    pso_bounds_gpu(m, lo, hi);
    if (m == 0) { lo[1] = -0.3; hi[1] = 0.3; }
}

// population_priors (src/parametric.rs:1588-1621)
void population_priors(int m, double* c, double* w) {
    const double L = 2.30258509299404568402;  // ln 10
    int np = n_params(m);
    if (m == 1) {
        const double cc[7] = {-0.009 * L, 0.009, 1.327 * L, -12.0, 0.360 * L, 1.454 * L, -1.033 * L};
        const double ww[7] = {0.10 * L, 0.010, 0.40 * L, 15.0, 0.35 * L, 0.40 * L, 0.50 * L};
        for (int d = 0; d < 7; d++) { c[d] = cc[d]; w[d] = ww[d]; }
    } else if (m == 0) {
        const double cc[6] = {0.0, 0.0, 0.0, 1.5, 3.0, -3.0};
        const double ww[6] = {2.0, 0.2, 30.0, 2.5, 2.5, 2.0};
        for (int d = 0; d < 6; d++) { c[d] = cc[d]; w[d] = ww[d]; }
    } else {
        for (int d = 0; d < np; d++) { c[d] = 0.0; w[d] = 5.0; }
    }
}

// prior_params (src/parametric.rs:2072-2115)
void prior_params(int m, double* c, double* w) {
    int np = n_params(m);
    for (int d = 0; d < np; d++) { c[d] = 0.0; w[d] = 2.0; }
    switch (m) {
        case 7: {
            const double cc[5] = {-1.75, -1.25, 0.5, -0.5, 0.0};
            const double ww[5] = {1.25, 0.75, 1.5, 1.5, 2.0};
            for (int d = 0; d < 5; d++) { c[d] = cc[d]; w[d] = ww[d]; }
            break;
        }
        case 2: c[5] = 1.67; w[5] = 1.0; break;
        case 3: c[2] = 2.3; w[2] = 1.0; break;
        case 4: c[2] = 3.0; w[2] = 1.5; c[3] = 2.3; w[3] = 1.0; break;
        case 5: c[2] = 0.5; w[2] = 1.0; c[3] = 1.0; w[3] = 2.0; break;
        case 6: c[2] = 1.0; w[2] = 2.0; c[3] = 0.8; w[3] = 1.0; c[4] = 2.0; w[4] = 1.0; break;
        default: break;
    }
}

void parallel_for(int64_t n, int n_threads, const std::function<void(int64_t)>& fn) {
    if (n_threads <= 1 || n <= 1) {
        for (int64_t i = 0; i < n; i++) fn(i);
        return;
    }
    int nt = (int)std::min<int64_t>(n_threads, n);
    std::vector<std::thread> th;
    std::atomic<int64_t> next(0);
    for (int k = 0; k < nt; k++)
        th.emplace_back([&]() {
            for (;;) {
                int64_t i = next.fetch_add(1);
                if (i >= n) break;
                fn(i);
            }
        });
    for (auto& x : th) x.join();
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
// SC'11): 10 rounds of two 32x32->64 multiplies with the published multipliers
// and Weyl key increments.
// ---------------------------------------------------------------------------
static void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
    uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
    uint32_t k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// 24-bit uniform in [0,1), exactly representable in FP32 and FP64.
static inline double u24(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

// ---------------------------------------------------------------------------
// costs
// ---------------------------------------------------------------------------
static double nll_gpu_kernel(int m, const double* p, const double* t, const double* f,
                             const double* v, const int32_t* up, const double* uf, int64_t n) {
    // batch_pso_cost (cuda/models.cu:516-584)
    if (n <= 0) return 1e99;
    int np = n_params(m);
    double se = std::exp(p[np - 1]);
    double se2 = se * se;
    double nll = 0.0;
    for (int64_t i = 0; i < n; i++) {
        double pred = (m == 7) ? metzger_single(p, t[i]) : eval_scalar(m, p, t[i]);
        if (!std::isfinite(pred)) return 1e99;
        double tv = v[i] + se2;
        if (up[i]) {
            double z = (uf[i] - pred) / std::sqrt(tv);
            nll -= log_normal_cdf(z);
        } else {
            double d = pred - f[i];
            nll += d * d / tv + std::log(tv);
        }
    }
    return nll / (double)n;
}

static double prior_penalty(int m, const double* p) {
    double c[7], w[7];
    population_priors(m, c, w);
    int np = n_params(m);
    double s = 0.0;
    for (int j = 0; j < np; j++)
        if (w[j] > 0.0) {
            double z = (p[j] - c[j]) / w[j];
            s += 0.5 * z * z;
        }
    return s;
}

double pso_cost_gpu(int m, const double* p, const double* t, const double* f, const double* v,
                    const int32_t* up, const double* uf, int64_t n) {
    // kernel cost + host prior / n_obs^2 (src/gpu_cuda.rs:461-480)
    double c = nll_gpu_kernel(m, p, t, f, v, up, uf, n);
    double nn = (double)std::max<int64_t>(n, 1);
    return c + prior_penalty(m, p) / (nn * nn);
}

// ---------------------------------------------------------------------------
// batched PSO, GPU semantics (src/gpu_cuda.rs:368-567)
// ---------------------------------------------------------------------------
struct SwarmOut {
    double params[7];
    double cost;
    int iters;
};

static SwarmOut pso_one_source(int m, int64_t s, const double* t, const double* f, const double* v,
                               const int32_t* up, const double* uf, int64_t n, int P, int max_iters,
                               int stall_iters, uint64_t seed, uint32_t flags) {
    const int np = n_params(m);
    double lo[7], hi[7], vmax[7];
    pso_bounds_gpu(m, lo, hi);
    for (int d = 0; d < np; d++) vmax[d] = 0.5 * (hi[d] - lo[d]);
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const int npair = (np + 1) / 2;

    std::vector<double> pos(P * np), vel(P * np), pbest(P * np, 0.0), pcost(P, INFINITY);
    double gbest[7] = {0, 0, 0, 0, 0, 0, 0};
    double gcost = INFINITY;

    // initial positions/velocities: one Philox block per (particle, dim pair)
    // at iteration slot 0xFFFFFFFF; x,y feed dim 2k and z,w feed dim 2k+1.
    for (int p = 0; p < P; p++)
        for (int k = 0; k < npair; k++) {
            uint32_t ctr[4] = {(uint32_t)s, (uint32_t)p, 0xFFFFFFFFu, (uint32_t)k}, r[4];
            philox(ctr, key, r);
            for (int h = 0; h < 2; h++) {
                int d = 2 * k + h;
                if (d >= np) break;
                pos[p * np + d] = lo[d] + u24(r[2 * h]) * (hi[d] - lo[d]);
                vel[p * np + d] = vmax[d] * 0.2 * (2.0 * u24(r[2 * h + 1]) - 1.0);
            }
        }

    // velocity-update stream of each particle: xoshiro128** seeded from the
    // Philox block at iteration slot 0xFFFFFFFE
    std::vector<Xoshiro128ss> rng(P);
    for (int p = 0; p < P; p++) {
        uint32_t ctr[4] = {(uint32_t)s, (uint32_t)p, 0xFFFFFFFEu, 0u}, r[4];
        philox(ctr, key, r);
        rng[p].seed(r);
    }

    const double w_max = 0.9, w_min = 0.4, c1 = 1.5, c2 = 1.5;
    const double inv_max = 1.0 / (double)max_iters;
    double prev = INFINITY;
    int stall = 0, it_done = 0;
    bool first_eval = true;
    std::vector<double> cost(P);
    for (int it = 0; it < max_iters; it++) {
        double w = w_max - (w_max - w_min) * (double)it * inv_max;
        for (int p = 0; p < P; p++) cost[p] = pso_cost_gpu(m, &pos[p * np], t, f, v, up, uf, n);
        it_done = it + 1;
        // personal / global bests, particle order, strict '<'
        for (int p = 0; p < P; p++) {
            if (cost[p] < pcost[p]) {
                pcost[p] = cost[p];
                std::memcpy(&pbest[p * np], &pos[p * np], sizeof(double) * np);
                if (cost[p] < gcost) {
                    gcost = cost[p];
                    std::memcpy(gbest, &pos[p * np], sizeof(double) * np);
                }
            }
        }
        if ((flags & ORC_PSO_STALL_FROM_FIRST_EVAL) && first_eval) prev = gcost;
        first_eval = false;
        // velocity / position update (the swarm is not yet done here)
        for (int p = 0; p < P; p++)
            for (int d = 0; d < np; d++) {
                {
                    double r1 = u24(rng[p].next());
                    double r2 = u24(rng[p].next());
                    double x = pos[p * np + d];
                    double vv = w * vel[p * np + d] + c1 * r1 * (pbest[p * np + d] - x) + c2 * r2 * (gbest[d] - x);
                    vv = std::min(std::max(vv, -vmax[d]), vmax[d]);
                    double nx = x + vv;
                    if (nx <= lo[d]) { pos[p * np + d] = lo[d]; vel[p * np + d] = 0.0; }
                    else if (nx >= hi[d]) { pos[p * np + d] = hi[d]; vel[p * np + d] = 0.0; }
                    else { pos[p * np + d] = nx; vel[p * np + d] = vv; }
                }
            }
        // stall rule (src/gpu_cuda.rs:537-552).  With prev starting at +inf
        // the reference never registers an improvement, so a swarm stops
        // after exactly stall_iters iterations; reproduced as is.
        bool improved = prev - gcost > 0.01 * std::max(std::fabs(prev), 1e-10);
        if (improved) { stall = 0; prev = gcost; }
        else if (++stall >= stall_iters) break;
    }
    SwarmOut o;
    for (int d = 0; d < 7; d++) o.params[d] = d < np ? gbest[d] : 0.0;
    o.cost = gcost;
    o.iters = it_done;
    return o;
}

}  // namespace orc

extern "C" {

double orc_log_normal_cdf(double x) { return orc::log_normal_cdf(x); }
double orc_dlog_normal_cdf(double x) { return orc::dlog_normal_cdf(x); }
double orc_erf_approx(double x) { return orc::erf_approx(x); }

int orc_pso_bounds(int m, double* lo, double* hi) {
    if (orc::n_params(m) < 0) return -1;
    orc::pso_bounds_gpu(m, lo, hi);
    return 0;
}
int orc_pso_bounds_cpu(int m, double* lo, double* hi) {
    if (orc::n_params(m) < 0) return -1;
    orc::pso_bounds_cpu(m, lo, hi);
    return 0;
}
int orc_population_priors(int m, double* c, double* w) {
    if (orc::n_params(m) < 0) return -1;
    orc::population_priors(m, c, w);
    return 0;
}
int orc_prior_params(int m, double* c, double* w) {
    if (orc::n_params(m) < 0) return -1;
    orc::prior_params(m, c, w);
    return 0;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    orc::philox(ctr, key, out);
}

double orc_pso_cost_gpu(int m, const double* p, const double* t, const double* f, const double* v,
                        const int32_t* up, const double* uf, int64_t n) {
    return orc::pso_cost_gpu(m, p, t, f, v, up, uf, n);
}

// PsoCost::cost_from_slice (src/parametric.rs:1303-1370)
double orc_pso_cost_cpu(int m, const double* p, const double* t, const double* f, const double* v,
                        const int32_t* up, const double* uf, int64_t n) {
    if (m == 1) {
        double beta = p[1], gamma = std::exp(p[2]), tr = std::exp(p[4]), tf = std::exp(p[5]);
        double c1 = std::max(gamma * beta - 1.0, 0.0);
        double c2 = std::max(std::exp(-gamma / tr) * (tf / tr - 1.0) - 1.0, 0.0);
        double c3 = std::max(beta * tf - 1.0 + beta * gamma, 0.0);
        double c = c1 + c2 + c3;
        if (c > 0.0) return 1e12 + 10000.0 * c;
    }
    int np = orc::n_params(m);
    double se = std::exp(p[np - 1]), se2 = se * se;
    std::vector<double> pred(std::max<int64_t>(n, 1));
    orc::eval_times(m, p, t, n, pred.data());
    if (m == 7) {
        double mx = -INFINITY;
        for (int64_t i = 0; i < n; i++)
            if (!up[i]) mx = std::max(mx, pred[i]);
        if (mx > 1e-10 && std::isfinite(mx)) {
            double sc = std::min(std::max(1.0 / mx, 0.1), 10.0);
            for (int64_t i = 0; i < n; i++) pred[i] *= sc;
        }
    }
    double nn = (double)std::max<int64_t>(n, 1);
    double nll = 0.0;
    for (int64_t i = 0; i < n; i++) {
        if (!std::isfinite(pred[i])) return 1e99;
        double tv = v[i] + se2;
        if (up[i]) nll -= orc::log_normal_cdf((uf[i] - pred[i]) / std::sqrt(tv));
        else {
            double d = pred[i] - f[i];
            nll += d * d / tv + std::log(tv);
        }
    }
    double c[7], w[7];
    orc::population_priors(m, c, w);
    double lp = 0.0;
    for (int j = 0; j < np; j++)
        if (w[j] > 0.0) {
            double z = (p[j] - c[j]) / w[j];
            lp += 0.5 * z * z;
        }
    nll += lp / nn;
    return nll / nn;
}

// PsoCost::gradient (src/parametric.rs:1383-1446); omits the prior term.
int orc_pso_cost_grad_cpu(int m, const double* p, const double* t, const double* f, const double* v,
                          const int32_t* up, const double* uf, int64_t n, double* grad) {
    const double PI = 3.14159265358979323846;
    int np = orc::n_params(m);
    if (np < 0 || m == 7) return -1;
    int se_idx = np - 1;
    double se = std::exp(p[se_idx]), se2 = se * se;
    double nn = (double)std::max<int64_t>(n, 1);
    double g[7] = {0}, mg[7];
    for (int64_t i = 0; i < n; i++) {
        double pred = orc::eval_scalar(m, p, t[i]);
        if (!std::isfinite(pred)) {
            for (int j = 0; j < np; j++) grad[j] = 0.0;
            return 0;
        }
        double tv = v[i] + se2;
        orc::eval_grad(m, p, t[i], mg);
        if (up[i]) {
            double sv = std::sqrt(tv);
            double z = (uf[i] - pred) / sv;
            double phi = std::exp(-0.5 * z * z) / std::sqrt(2.0 * PI);
            double cdf = 0.5 * (1.0 + orc::erf_approx(z / std::sqrt(2.0)));
            double ratio = phi / std::max(cdf, 1e-300);
            for (int j = 0; j < np; j++) g[j] += ratio * mg[j] / sv;
            g[se_idx] += ratio * z * se2 / tv;
        } else {
            double d = pred - f[i];
            double sc = 2.0 * d / tv;
            for (int j = 0; j < np; j++) g[j] += sc * mg[j];
            g[se_idx] += (-d * d / (tv * tv) + 1.0 / tv) * (2.0 * se2);
        }
    }
    for (int j = 0; j < np; j++) grad[j] = g[j] / nn;
    return 0;
}

int orc_batch_pso(int m, int64_t n_src, const int64_t* off, const double* t, const double* f,
                  const double* v, const int32_t* up, const double* uf, int P, int max_iters,
                  int stall_iters, uint64_t seed, uint32_t flags, double* out_params,
                  double* out_cost, int32_t* out_iters, int n_threads) {
    int np = orc::n_params(m);
    if (np < 0 || n_src < 0 || P <= 0 || max_iters <= 0 || stall_iters < 0) return -1;
    orc::parallel_for(n_src, n_threads, [&](int64_t s) {
        int64_t a = off[s], b = off[s + 1];
        orc::SwarmOut o = orc::pso_one_source(m, s, t + a, f + a, v + a, up + a, uf + a, b - a, P,
                                              max_iters, stall_iters, seed, flags);
        for (int d = 0; d < np; d++) out_params[s * np + d] = o.params[d];
        out_cost[s] = o.cost;
        if (out_iters) out_iters[s] = o.iters;
    });
    return 0;
}

// batch_model_select (src/gpu_cuda.rs:573-620): gate model first; sources
// whose gate cost is >= gate continue through the remaining models in list
// order, replacing the best only on a strictly lower cost.  Seed 42 for every
// model as in the reference.
int orc_batch_model_select(int64_t n_src, const int64_t* off, const double* t, const double* f,
                           const double* v, const int32_t* up, const double* uf,
                           const int32_t* models, int n_models, int P, int max_iters,
                           int stall_iters, double gate, uint32_t flags, int32_t* out_model,
                           double* out_params, double* out_cost, int n_threads) {
    if (n_models <= 0) return -1;
    for (int k = 0; k < n_models; k++)
        if (orc::n_params(models[k]) < 0) return -1;
    orc::parallel_for(n_src, n_threads, [&](int64_t s) {
        int64_t a = off[s], b = off[s + 1];
        int best_m = models[0];
        orc::SwarmOut best = orc::pso_one_source(models[0], s, t + a, f + a, v + a, up + a, uf + a,
                                                 b - a, P, max_iters, stall_iters, 42, flags);
        if (best.cost >= gate) {
            for (int k = 1; k < n_models; k++) {
                orc::SwarmOut o = orc::pso_one_source(models[k], s, t + a, f + a, v + a, up + a,
                                                      uf + a, b - a, P, max_iters, stall_iters, 42, flags);
                if (o.cost < best.cost) { best = o; best_m = models[k]; }
            }
        }
        out_model[s] = best_m;
        for (int d = 0; d < 7; d++) out_params[s * 7 + d] = best.params[d];
        out_cost[s] = best.cost;
    });
    return 0;
}

}  // extern "C"
