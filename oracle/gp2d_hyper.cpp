// gp2d_hyper.cpp -- oracle of the 2D (time x log10 wavelength) GP
// hyper-parameter fit: negative log marginal likelihood, its adjoint gradient
// and the L-BFGS driver.
//
// TEST INFRASTRUCTURE ONLY (parity checker for the HIP product).
//
// The reference selects the 2D GP's hyper-parameters from a grid
// (src/gp2d.rs:336-439, 516-724) and has no gradient; the optimised form is
// new, so this file restates the textbook quantities it is built from:
//   K_ij = a exp(-(t_i - t_j)^2 / (2 lt^2) - (w_i - w_j)^2 / (2 lw^2))
//          + delta_ij (nv_i + sn^2),           theta = ln (a, lt, lw, sn)
//   r = y - mean(y),  K = L L^T,  z = L^-1 r,  alpha = K^-1 r
//   NLML = 1/2 z^T z + sum ln L_ii + n/2 ln 2 pi   (Rasmussen & Williams 2.30)
//   dNLML/dtheta_k = 1/2 tr((K^-1 - alpha alpha^T) dK/dtheta_k)   (R & W 5.9)
// with dK/d ln a = a k, dK/d ln lt = a k dt^2 / lt^2, dK/d ln lw =
// a k dw^2 / lw^2, dK/d ln sn = 2 sn^2 I; K^-1 = X^T X with X = L^-1.
// The driver is limited-memory BFGS (Nocedal & Wright, algorithm 7.4; m = 6
// pairs, H0 = s^T y / y^T y), a projected backtracking Armijo line search
// (c1 = 1e-4, halving, at most 30 trials, first trial capped to a step of 2
// in every log-parameter, given up once the step moves theta by <= 1e-10),
// box bounds on theta, and a final re-evaluation at
// the returned point; convergence is tested on the projected gradient
// (components pressing against their bound count as zero).  The device kernel
// runs the same algorithm.
#include "lcf_oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace orc {

namespace {

const double LN2PI = 1.83787706640934548356;

struct Gp2dData {
    const double* t;
    const double* w;
    const double* y;
    const double* nv;
    int n;
};

// NLML at theta; on success L (row-major n x n lower) and z are filled.
bool nlml_eval(const Gp2dData& d, const double* th, std::vector<double>& L, std::vector<double>& z,
               double& f) {
    const int n = d.n;
    const double a = std::exp(th[0]), lt = std::exp(th[1]), lw = std::exp(th[2]), sn = std::exp(th[3]);
    const double it = 1.0 / (2.0 * lt * lt), iw = 1.0 / (2.0 * lw * lw), sn2 = sn * sn;
    double ybar = 0.0;
    for (int i = 0; i < n; i++) ybar += d.y[i];
    ybar /= n;
    L.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j <= i; j++) {
            const double dt = d.t[i] - d.t[j], dw = d.w[i] - d.w[j];
            double k = a * std::exp(-(dt * dt * it + dw * dw * iw));
            if (i == j) k += d.nv[i] + sn2;
            L[(size_t)i * n + j] = k;
        }
    // Cholesky-Crout, terms subtracted in increasing k
    for (int j = 0; j < n; j++) {
        double s = L[(size_t)j * n + j];
        for (int k = 0; k < j; k++) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        if (!(s > 0.0)) return false;
        const double ljj = std::sqrt(s);
        L[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; i++) {
            double v = L[(size_t)i * n + j];
            for (int k = 0; k < j; k++) v -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = v / ljj;
        }
    }
    z.assign(n, 0.0);
    double q = 0.0, ld = 0.0;
    for (int i = 0; i < n; i++) {
        double v = d.y[i] - ybar;
        for (int k = 0; k < i; k++) v -= L[(size_t)i * n + k] * z[k];
        z[i] = v / L[(size_t)i * n + i];
        q += z[i] * z[i];
        ld += std::log(L[(size_t)i * n + i]);
    }
    f = 0.5 * q + ld + 0.5 * n * LN2PI;
    return std::isfinite(f);
}

// Gradient from the factor of the last successful nlml_eval at theta.
void grad_eval(const Gp2dData& d, const double* th, const std::vector<double>& L, const std::vector<double>& z,
               double* g) {
    const int n = d.n;
    const double a = std::exp(th[0]), lt = std::exp(th[1]), lw = std::exp(th[2]), sn = std::exp(th[3]);
    const double it = 1.0 / (2.0 * lt * lt), iw = 1.0 / (2.0 * lw * lw), sn2 = sn * sn;
    // X = L^-1, column by column
    std::vector<double> X((size_t)n * n, 0.0);
    for (int c = 0; c < n; c++)
        for (int r = c; r < n; r++) {
            double s = r == c ? 1.0 : 0.0;
            for (int q = c; q < r; q++) s -= L[(size_t)r * n + q] * X[(size_t)q * n + c];
            X[(size_t)r * n + c] = s / L[(size_t)r * n + r];
        }
    std::vector<double> al(n, 0.0);
    for (int j = 0; j < n; j++) {
        double s = 0.0;
        for (int i = j; i < n; i++) s += X[(size_t)i * n + j] * z[i];
        al[j] = s;
    }
    double ga = 0.0, gt = 0.0, gw = 0.0, gs = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j <= i; j++) {
            double kinv = 0.0;
            for (int k = i; k < n; k++) kinv += X[(size_t)k * n + i] * X[(size_t)k * n + j];
            const double W = kinv - al[i] * al[j];
            const double dt = d.t[i] - d.t[j], dw = d.w[i] - d.w[j];
            const double ak = a * std::exp(-(dt * dt * it + dw * dw * iw));
            const double m = i == j ? 1.0 : 2.0;
            ga += m * W * ak;
            gt += m * W * ak * (dt * dt) * (2.0 * it);
            gw += m * W * ak * (dw * dw) * (2.0 * iw);
            if (i == j) gs += W * 2.0 * sn2;
        }
    g[0] = 0.5 * ga;
    g[1] = 0.5 * gt;
    g[2] = 0.5 * gw;
    g[3] = 0.5 * gs;
}

}  // namespace

// Box on theta = ln (a, lt, lw, sn).
const double GP2D_TH_LO[4] = {-13.815510557964274, -2.302585092994046, -6.907755278982137, -13.815510557964274};
const double GP2D_TH_HI[4] = {13.815510557964274, 9.210340371976184, 2.302585092994046, 2.302585092994046};

int gp2d_train_one(const Gp2dData& d, const double* th0, int max_iters, double gtol, double* th_out,
                   double* f_out, double* g_out, int* iters_out, int* evals) {
    const int M = 6;
    double th[4], g[4], f;
    for (int k = 0; k < 4; k++) th[k] = std::min(std::max(th0[k], GP2D_TH_LO[k]), GP2D_TH_HI[k]);
    std::vector<double> L, z;
    *iters_out = 0;
    evals[0] = 1;
    evals[1] = 0;
    if (!nlml_eval(d, th, L, z, f)) {
        for (int k = 0; k < 4; k++) { th_out[k] = th[k]; g_out[k] = NAN; }
        *f_out = NAN;
        return 3;
    }
    grad_eval(d, th, L, z, g);
    evals[1]++;
    double S[6][4], Y[6][4], rho[6];
    int nm = 0, head = 0, status = 1, it = 0;
    for (; it < max_iters; it++) {
        // projected gradient: a component pressing against its bound is spent
        // and held fixed for this iteration (zeroed in the recursion's input
        // and in the direction)
        bool pin[4];
        double gi = 0.0;
        for (int k = 0; k < 4; k++) {
            pin[k] = (th[k] <= GP2D_TH_LO[k] && g[k] > 0.0) || (th[k] >= GP2D_TH_HI[k] && g[k] < 0.0);
            gi = std::max(gi, pin[k] ? 0.0 : std::fabs(g[k]));
        }
        if (gi <= gtol) { status = 0; break; }
        double gfn = 0.0;
        for (int k = 0; k < 4; k++)
            if (!pin[k]) gfn += g[k] * g[k];
        const double gsc = 1.0 / std::max(1.0, std::sqrt(gfn));
        // two-loop recursion, newest pair first
        double q[4], al[6];
        for (int k = 0; k < 4; k++) q[k] = pin[k] ? 0.0 : g[k];
        for (int m = 0; m < nm; m++) {
            const int p = (head - 1 - m + M) % M;
            double s = 0.0;
            for (int k = 0; k < 4; k++) s += S[p][k] * q[k];
            al[m] = rho[p] * s;
            for (int k = 0; k < 4; k++) q[k] -= al[m] * Y[p][k];
        }
        double gam = 1.0;
        if (nm > 0) {
            const int p = (head - 1 + M) % M;
            double sy = 0.0, yy = 0.0;
            for (int k = 0; k < 4; k++) { sy += S[p][k] * Y[p][k]; yy += Y[p][k] * Y[p][k]; }
            gam = sy / yy;
        } else {
            gam = gsc;
        }
        for (int k = 0; k < 4; k++) q[k] *= gam;
        for (int m = nm - 1; m >= 0; m--) {
            const int p = (head - 1 - m + M) % M;
            double s = 0.0;
            for (int k = 0; k < 4; k++) s += Y[p][k] * q[k];
            const double b = rho[p] * s;
            for (int k = 0; k < 4; k++) q[k] += S[p][k] * (al[m] - b);
        }
        double dd[4], gd = 0.0;
        for (int k = 0; k < 4; k++) { dd[k] = pin[k] ? 0.0 : -q[k]; gd += g[k] * dd[k]; }
        bool steepest = false;
        auto to_steepest = [&]() {  // scaled free gradient; the curvature memory is dropped
            nm = 0;
            for (int k = 0; k < 4; k++) dd[k] = pin[k] ? 0.0 : -g[k] * gsc;
            steepest = true;
        };
        auto first_step = [&]() {
            double dmax = 0.0;
            for (int k = 0; k < 4; k++) dmax = std::max(dmax, std::fabs(dd[k]));
            return dmax > 2.0 ? 2.0 / dmax : 1.0;
        };
        if (!(gd < 0.0)) to_steepest();  // not a descent direction
        double step = first_step();
        double tt[4], ft = NAN;
        bool ok = false;
        for (int ls = 0; ls < 30; ls++) {
            double dec = 0.0, mv = 0.0;
            for (int pass = 0; pass < 2; pass++) {
                dec = 0.0;
                mv = 0.0;
                for (int k = 0; k < 4; k++) {
                    tt[k] = std::min(std::max(th[k] + step * dd[k], GP2D_TH_LO[k]), GP2D_TH_HI[k]);
                    dec += g[k] * (tt[k] - th[k]);
                    mv = std::max(mv, std::fabs(tt[k] - th[k]));
                }
                // an L-BFGS step that the projection leaves without decrease
                // or movement: restart along the free steepest descent
                if (steepest || (dec < 0.0 && !(mv <= 1e-10))) break;
                to_steepest();
                step = first_step();
            }
            if (mv <= 1e-10) break;  // the step no longer moves theta: no further decrease
            evals[0]++;
            if (nlml_eval(d, tt, L, z, ft) && ft <= f + 1e-4 * dec) { ok = true; break; }
            step *= 0.5;
        }
        if (!ok) { status = 2; break; }
        double gn[4];
        grad_eval(d, tt, L, z, gn);
        evals[1]++;
        double sy = 0.0, sk[4], yk[4];
        for (int k = 0; k < 4; k++) {
            sk[k] = tt[k] - th[k];
            yk[k] = gn[k] - g[k];
            sy += sk[k] * yk[k];
        }
        if (sy > 1e-12) {  // curvature pair kept only when s^T y > 0
            for (int k = 0; k < 4; k++) { S[head][k] = sk[k]; Y[head][k] = yk[k]; }
            rho[head] = 1.0 / sy;
            head = (head + 1) % M;
            nm = std::min(nm + 1, M);
        }
        for (int k = 0; k < 4; k++) { th[k] = tt[k]; g[k] = gn[k]; }
        f = ft;
    }
    *iters_out = it;
    // final state at the returned point
    nlml_eval(d, th, L, z, f);
    grad_eval(d, th, L, z, g);
    evals[0]++;
    evals[1]++;
    for (int k = 0; k < 4; k++) { th_out[k] = th[k]; g_out[k] = g[k]; }
    *f_out = f;
    return status;
}

}  // namespace orc

extern "C" int orc_gp2d_nlml_grad(const double* t, const double* w, const double* y, const double* nv, int n,
                                  const double* theta, double* nlml, double* grad) {
    orc::Gp2dData d{t, w, y, nv, n};
    std::vector<double> L, z;
    double f;
    if (n <= 0 || !orc::nlml_eval(d, theta, L, z, f)) return 1;
    *nlml = f;
    orc::grad_eval(d, theta, L, z, grad);
    return 0;
}

extern "C" int orc_batch_gp2d_train(int64_t n_src, const int64_t* off, const double* t, const double* w,
                                    const double* y, const double* nv, const double* theta0, int max_iters,
                                    double gtol, double* theta, double* nlml, double* grad, int32_t* iters,
                                    int32_t* status, int32_t* evals, int n_threads) {
    orc::parallel_for(n_src, n_threads, [&](int64_t s) {
        const int64_t a = off[s];
        orc::Gp2dData d{t + a, w + a, y + a, nv + a, (int)(off[s + 1] - a)};
        int it = 0;
        if (d.n <= 0) {
            for (int k = 0; k < 4; k++) { theta[s * 4 + k] = theta0[s * 4 + k]; grad[s * 4 + k] = NAN; }
            nlml[s] = NAN;
            iters[s] = 0;
            status[s] = 3;
            evals[2 * s] = evals[2 * s + 1] = 0;
            return;
        }
        int ev[2];
        status[s] = orc::gp2d_train_one(d, theta0 + s * 4, max_iters, gtol, theta + s * 4, nlml + s, grad + s * 4,
                                        &it, ev);
        iters[s] = it;
        evals[2 * s] = ev[0];
        evals[2 * s + 1] = ev[1];
    });
    return 0;
}
