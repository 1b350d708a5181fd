// multibazin.cpp -- oracle restatement of the GPU greedy MultiBazin fit.
//
// TEST INFRASTRUCTURE ONLY (parity checker for the HIP product).
//
// GpuContext::batch_pso_multi_bazin (src/gpu_cuda.rs:650-963) with its cost
// kernel batch_pso_cost_multi_bazin (cuda/models.cu:286-367), in FP64 with the
// reference's expressions:
//   * K = 1..4 components, parameters [ln A, t0, ln tau_rise, ln tau_fall] x K,
//     B, ln sigma_x; bounds [-3, T_lo, -2, -2] .. [3, T_hi, 5, 6], B in
//     [-0.3, 0.3], ln sigma_x in [-5, 0], T_lo/T_hi = batch time range -/+ 30;
//   * cost = NLL / n (no population prior), 1e99 for a non-finite prediction
//     or an empty band;
//   * synchronous swarm update, inertia 0.9 -> 0.4, c1 = c2 = 1.5, velocity
//     clamp half the range, wall absorption, the reference's stall rule;
//   * from K = 2, particle 0 starts at the K - 1 solution plus a component at
//     the largest detection residual [ln max(r, 1e-10), t_peak, 1, 1] (then
//     clamped), with velocity (hi - lo) 0.02 (2u - 1);
//   * BIC = 2 cost n + (4K + 2) ln n, best K by strict '<', stop once a K's BIC
//     exceeds the previous K's by more than 2.
// Random numbers: initial positions and velocities from Philox4x32-10 keyed by
// the seed, counter (band, particle, 0xFFFFFFFF, 0x4D420000 + (K << 8) +
// dimension pair); the velocity updates from a per-particle xoshiro128**
// stream seeded by the block (band, particle, 0xFFFFFFFE, 0x4D420000 +
// (K << 8)), r1 and r2 drawn per dimension in order.  They stand in for the
// reference's SmallRng streams (not reproducible here; parity unpinned).
#include "lcf_oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace orc {

static inline double u24mb(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

// multi_bazin_at (cuda/models.cu:297-313)
static double mb_eval(const double* p, int k, double t) {
    double sum = 0.0;
    for (int c = 0; c < k; c++) {
        const double* q = p + 4 * c;
        double a = std::exp(q[0]), t0 = q[1], tr = std::exp(q[2]), tf = std::exp(q[3]);
        double dt = t - t0;
        sum += a * std::exp(-dt / tf) * (1.0 / (1.0 + std::exp(-(dt / tr))));
    }
    return sum + p[4 * k];
}

// batch_pso_cost_multi_bazin (cuda/models.cu:319-367)
static double mb_cost(const double* p, int k, const double* t, const double* f, const double* v,
                      const int32_t* up, const double* uf, int64_t n) {
    if (n <= 0) return 1e99;
    double se = std::exp(p[4 * k + 1]);
    double se2 = se * se;
    double nll = 0.0;
    for (int64_t i = 0; i < n; i++) {
        double pred = mb_eval(p, k, t[i]);
        if (!std::isfinite(pred)) return 1e99;
        double tv = v[i] + se2;
        if (up[i]) {
            nll -= log_normal_cdf((uf[i] - pred) / std::sqrt(tv));
        } else {
            double d = pred - f[i];
            nll += d * d / tv + std::log(tv);
        }
    }
    return nll / (double)n;
}

struct MbOut {
    int best_k;
    double params[18], cost, bic, pk_cost[4], pk_bic[4];
};

static MbOut mb_source(int64_t s, const double* t, const double* f, const double* v, const int32_t* up,
                       const double* uf, int64_t n, double t_lo, double t_hi, int P, int max_iters,
                       int stall_iters, uint64_t seed) {
    MbOut o;
    o.best_k = 1;
    o.cost = INFINITY;
    o.bic = INFINITY;
    for (int d = 0; d < 18; d++) o.params[d] = 0.0;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    std::vector<double> prev;
    bool stopped = false;
    for (int k = 1; k <= 4; k++) {
        if (stopped) {
            o.pk_cost[k - 1] = NAN;
            o.pk_bic[k - 1] = NAN;
            continue;
        }
        const int np = 4 * k + 2;
        std::vector<double> lo(np), hi(np), vmax(np);
        for (int c = 0; c < k; c++) {
            lo[4 * c] = -3.0; lo[4 * c + 1] = t_lo; lo[4 * c + 2] = -2.0; lo[4 * c + 3] = -2.0;
            hi[4 * c] = 3.0; hi[4 * c + 1] = t_hi; hi[4 * c + 2] = 5.0; hi[4 * c + 3] = 6.0;
        }
        lo[4 * k] = -0.3; hi[4 * k] = 0.3;
        lo[4 * k + 1] = -5.0; hi[4 * k + 1] = 0.0;
        for (int d = 0; d < np; d++) vmax[d] = 0.5 * (hi[d] - lo[d]);
        // seeded particle from the K - 1 solution (src/gpu_cuda.rs:712-757)
        std::vector<double> seedp;
        if (k > 1) {
            int peak_idx = 0;
            double peak_val = -INFINITY;
            for (int64_t i = 0; i < n; i++) {
                double pred = 0.0;
                for (int c = 0; c < k - 1; c++) {
                    const double* q = &prev[4 * c];
                    double a = std::exp(q[0]), t0 = q[1], tr = std::exp(q[2]), tf = std::exp(q[3]);
                    double dt = t[i] - t0;
                    double sig = 1.0 / (1.0 + std::exp(-dt / tr));
                    pred += a * std::exp(-dt / tf) * sig;
                }
                pred += prev[4 * (k - 1)];
                double resid = f[i] - pred;
                if (!up[i] && resid > peak_val) {
                    peak_val = resid;
                    peak_idx = (int)i;
                }
            }
            seedp.assign(prev.begin(), prev.begin() + 4 * (k - 1));
            seedp.push_back(std::log(std::max(peak_val, 1e-10)));
            seedp.push_back(n > 0 ? t[peak_idx] : 0.0);
            seedp.push_back(1.0);
            seedp.push_back(1.0);
            seedp.push_back(prev[4 * (k - 1)]);
            seedp.push_back(prev[4 * (k - 1) + 1]);
            for (int d = 0; d < np; d++) seedp[d] = std::min(std::max(seedp[d], lo[d]), hi[d]);
        }
        const uint32_t tag = 0x4D420000u + ((uint32_t)k << 8);
        const int npair = (np + 1) / 2;
        std::vector<double> pos(P * np), vel(P * np), pb(P * np, 0.0), pc(P, INFINITY), cost(P), gb(np, 0.0);
        double gc = INFINITY;
        for (int p = 0; p < P; p++)
            for (int q = 0; q < npair; q++) {
                uint32_t ctr[4] = {(uint32_t)s, (uint32_t)p, 0xFFFFFFFFu, tag + (uint32_t)q}, r[4];
                orc_philox4x32_10(ctr, key, r);
                for (int h = 0; h < 2; h++) {
                    int d = 2 * q + h;
                    if (d >= np) break;
                    double ux = u24mb(r[2 * h]), uv = u24mb(r[2 * h + 1]);
                    if (k > 1 && p == 0) {
                        pos[d] = seedp[d];
                        vel[d] = (hi[d] - lo[d]) * 0.02 * (2.0 * uv - 1.0);
                    } else {
                        pos[p * np + d] = lo[d] + ux * (hi[d] - lo[d]);
                        vel[p * np + d] = vmax[d] * 0.2 * (2.0 * uv - 1.0);
                    }
                }
            }
        std::vector<orc::Xoshiro128ss> rng(P);
        for (int p = 0; p < P; p++) {
            uint32_t ctr[4] = {(uint32_t)s, (uint32_t)p, 0xFFFFFFFEu, tag}, r[4];
            orc_philox4x32_10(ctr, key, r);
            rng[p].seed(r);
        }
        const double w_max = 0.9, w_min = 0.4, c1 = 1.5, c2 = 1.5;
        const double inv_max = 1.0 / (double)max_iters;
        double prev_g = INFINITY;
        int stall = 0;
        for (int it = 0; it < max_iters; it++) {
            double w = w_max - (w_max - w_min) * (double)it * inv_max;
            for (int p = 0; p < P; p++) cost[p] = mb_cost(&pos[p * np], k, t, f, v, up, uf, n);
            for (int p = 0; p < P; p++)
                if (cost[p] < pc[p]) {
                    pc[p] = cost[p];
                    std::copy(&pos[p * np], &pos[p * np] + np, &pb[p * np]);
                    if (cost[p] < gc) {
                        gc = cost[p];
                        std::copy(&pos[p * np], &pos[p * np] + np, gb.begin());
                    }
                }
            for (int p = 0; p < P; p++)
                for (int d = 0; d < np; d++) {
                    {
                        double r1 = u24mb(rng[p].next());
                        double r2 = u24mb(rng[p].next());
                        double x = pos[p * np + d];
                        double vv = w * vel[p * np + d] + c1 * r1 * (pb[p * np + d] - x) + c2 * r2 * (gb[d] - x);
                        vv = std::min(std::max(vv, -vmax[d]), vmax[d]);
                        double nx = x + vv;
                        if (nx <= lo[d]) { pos[p * np + d] = lo[d]; vel[p * np + d] = 0.0; }
                        else if (nx >= hi[d]) { pos[p * np + d] = hi[d]; vel[p * np + d] = 0.0; }
                        else { pos[p * np + d] = nx; vel[p * np + d] = vv; }
                    }
                }
            bool improved = prev_g - gc > 0.01 * std::max(std::fabs(prev_g), 1e-10);
            if (improved) { stall = 0; prev_g = gc; }
            else if (++stall >= stall_iters) break;
        }
        const double nn = (double)n;
        const double bic = 2.0 * gc * nn + (double)np * std::log(nn);
        o.pk_cost[k - 1] = gc;
        o.pk_bic[k - 1] = bic;
        if (bic < o.bic) {
            o.bic = bic;
            o.cost = gc;
            o.best_k = k;
            for (int d = 0; d < 18; d++) o.params[d] = d < np ? gb[d] : 0.0;
        }
        prev = gb;
        if (k > 1 && bic > o.pk_bic[k - 2] + 2.0) stopped = true;
    }
    return o;
}

}  // namespace orc

extern "C" double orc_multi_bazin_cost(const double* p, int k, const double* t, const double* f,
                                      const double* v, const int32_t* up, const double* uf, int64_t n) {
    if (k < 1 || k > 4) return NAN;
    return orc::mb_cost(p, k, t, f, v, up, uf, n);
}

extern "C" int orc_batch_multi_bazin(int64_t n_src, const int64_t* off, const double* t, const double* f,
                                     const double* v, const int32_t* up, const double* uf, int P,
                                     int max_iters, int stall_iters, uint64_t seed, int32_t* best_k,
                                     double* params, double* cost, double* bic, double* per_k_cost,
                                     double* per_k_bic, int n_threads) {
    if (n_src < 0 || P <= 0 || max_iters <= 0 || stall_iters < 0) return -1;
    double tmin = INFINITY, tmax = -INFINITY;
    for (int64_t i = 0; i < (n_src > 0 ? off[n_src] : 0); i++) {
        tmin = std::min(tmin, t[i]);
        tmax = std::max(tmax, t[i]);
    }
    orc::parallel_for(n_src, n_threads, [&](int64_t s) {
        int64_t a = off[s], b = off[s + 1];
        orc::MbOut o = orc::mb_source(s, t + a, f + a, v + a, up + a, uf + a, b - a, tmin - 30.0, tmax + 30.0,
                                      P, max_iters, stall_iters, seed);
        best_k[s] = o.best_k;
        for (int d = 0; d < 18; d++) params[s * 18 + d] = o.params[d];
        cost[s] = o.cost;
        bic[s] = o.bic;
        for (int k = 0; k < 4; k++) {
            per_k_cost[s * 4 + k] = o.pk_cost[k];
            per_k_bic[s * 4 + k] = o.pk_bic[k];
        }
    });
    return 0;
}
