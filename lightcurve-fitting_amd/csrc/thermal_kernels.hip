// thermal_kernels.hip -- batched blackbody colour-temperature fit for gfx950,
// FP64 (fit_thermal, src/thermal.rs:222-418).
//
// One 64-thread workgroup (a single wavefront) per source:
//   1. the reference-band GP: the prefit dense GP of the state row, or
//      fit_ref_band_gp (src/thermal.rs:167-212) with the <= 32-point
//      training set one row per lane in registers (Cholesky by lane reads,
//      each element reduced in the reference's k order), the triangular solves
//      and the train RMS in the reference's summation order;
//   2. colour observations: every lane predicts the reference band at its own
//      observations of the other bands; finite ones are compacted in order;
//   3. both particle-swarm runs at once, lane = (run, particle): each lane
//      moves its particle and sums its own cost over the colours in the
//      reference order; the run's global best by lane reads in particle order.
// The swarm follows argmin 0.11's ParticleSwarm (inertia 1/(2 ln 2),
// cognitive = social = 1/2 + ln 2, box [3, 6] x [-0.05, 0.01], 10 particles,
// 30 iterations, two runs); random numbers are Philox4x32-10 on the counter
// (0, 64 run + particle, iteration or 0xFFFFFFFF, 0x54480000): every source
// draws the same stream, as every fit_thermal call of the reference starts
// its own generator, so a source's result does not depend on its batch.
// Built without multiply-add contraction, like the reference's arithmetic.
#include "thermal_kernels.hpp"
#include "lcf_device.hpp"

namespace lcf {

namespace {

constexpr double TH_HC_OVER_K = 1.4388e8;
constexpr double TH_LOG10_E = 0.434294481903251827651128918916605082;
constexpr int TH_P = 10;
constexpr int TH_ITERS = 30;
constexpr uint32_t TH_TAG = 0x54480000u;

__device__ __forceinline__ double th_u24(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

__device__ __forceinline__ double th_rbf(double a, double b, double amp, double it) {
    double d = a - b;
    return amp * ::exp(-d * d * it);
}

// bb_color_mag (src/thermal.rs:75-92)
__device__ __forceinline__ double bb_color_mag(double lref, double l, double temp) {
    const double x_ref = TH_HC_OVER_K / (lref * temp);
    const double x = TH_HC_OVER_K / (l * temp);
    if (x > 500.0 || x_ref > 500.0) return -2.5 * (3.0 * ::log10(l / lref) + (x - x_ref) * TH_LOG10_E);
    const double q = l / lref;
    const double ratio = q * q * q * (::exp(x) - 1.0) / (::exp(x_ref) - 1.0);
    if (ratio <= 0.0 || !__builtin_isfinite(ratio)) return NAN;
    return -2.5 * ::log10(ratio);
}

struct ThLds {
    double xs[TH_MAX_GP];         // GP training times
    double al[TH_MAX_GP];         // GP weights
    double zz[TH_MAX_GP];         // forward-solve result
    double L[TH_MAX_GP * TH_MAX_GP];  // factor, row-major
    double pv[TH_MAX_GP];         // predictions at the training points
    int cnt[2];
};

// fit_ref_band_gp into the state row; returns true when a candidate was kept.
__device__ __forceinline__ bool fit_ref_gp(const ThermalArgs& a, ThLds& s, const double* t, const double* m,
                                           const double* e, int n, double dur, double* st) {
    const int lane = lane_id();
    const int cap = a.cap;
    const int k = n <= 30 ? n : 25;  // subsample size (src/thermal.rs:168-172)
    if (k > TH_MAX_GP || k > cap || k <= 0) return false;
    // this lane's training point (src/gp.rs:5-27)
    int src = lane;
    if (n > k) {
        const double step = (double)n / (double)k;
        src = (int)::floor(((double)lane + 0.5) * step);
        src = src < n - 1 ? src : n - 1;
    }
    const bool own = lane < k;
    const double xr = own ? t[src] : 0.0;
    const double yr = own ? m[src] : 0.0;
    double nvr = 0.0;
    if (own) {
        const double er = e[src];
        nvr = fmax(er * er, 1e-6);
    }
    if (own) s.xs[lane] = xr;
    __syncthreads();
    double ym = 0.0, omin = INFINITY;  // sequential over the subsample, as the reference
    for (int i = 0; i < k; i++) {
        const double yi = __shfl(yr, i, WAVE);
        ym += yi;
        omin = fmin(omin, yi);
    }
    ym = ym / (double)k;
    const double amps[2] = {0.1, 0.3}, fac[3] = {6.0, 12.0, 24.0};
    double best = INFINITY;
    bool have = false;
    for (int ca = 0; ca < 2; ca++)
        for (int cf = 0; cf < 3; cf++) {
            const double amp = amps[ca];
            const double ls = fmax(dur / fac[cf], 0.1);
            const double it = 0.5 / (ls * ls);
            // row `lane` of K + diag(max(nv, 1e-10)) in registers
            double v[TH_MAX_GP];
#pragma unroll
            for (int c = 0; c < TH_MAX_GP; c++) v[c] = (own && c <= lane && c < k) ? th_rbf(xr, s.xs[c], amp, it) : 0.0;
#pragma unroll
            for (int c = 0; c < TH_MAX_GP; c++)
                if (c == lane && own) v[c] += fmax(nvr, 1e-10);
            // Cholesky: column j scaled, then columns c > j of rows >= c lose l_ij l_cj
            bool bad = false;
#pragma unroll
            for (int j = 0; j < TH_MAX_GP; j++) {
                if (j < k) {
                    const double piv = __shfl(v[j], j, WAVE);
                    bad = bad || piv <= 0.0;
                    const double ljj = ::sqrt(piv);
                    v[j] = lane == j ? ljj : (lane > j ? v[j] / ljj : v[j]);
#pragma unroll
                    for (int c = j + 1; c < TH_MAX_GP; c++) {
                        if (c < k) {
                            const double lcj = __shfl(v[j], c, WAVE);
                            if (lane >= c) v[c] -= v[j] * lcj;
                        }
                    }
                }
            }
            if (bad) continue;  // uniform: every lane read the same pivots
#pragma unroll
            for (int c = 0; c < TH_MAX_GP; c++)
                if (own && c <= lane) s.L[lane * TH_MAX_GP + c] = v[c];
            // forward solve L z = y - ym, right-looking: row i loses l_ij z_j in increasing j
            double b = own ? yr - ym : 0.0;
#pragma unroll
            for (int j = 0; j < TH_MAX_GP; j++) {
                if (j < k) {
                    const double zj = __shfl(b, j, WAVE) / __shfl(v[j], j, WAVE);
                    if (lane == j) b = zj;
                    else if (lane > j) b -= v[j] * zj;
                }
            }
            if (own) s.zz[lane] = b;
            __syncthreads();
            // backward solve L^T alpha = z, each row's sum in increasing j (solve_lt)
            for (int i = k - 1; i >= 0; i--) {
                double sacc = s.zz[i];
                for (int j = i + 1; j < k; j++) sacc -= s.L[j * TH_MAX_GP + i] * s.al[j];
                const double ai = sacc / s.L[i * TH_MAX_GP + i];
                __syncthreads();
                if (lane == 0) s.al[i] = ai;
                __syncthreads();
            }
            // predictions at the subsample (per lane), RMS in point order
            if (own) {
                double dot = 0.0;
                for (int i = 0; i < k; i++) dot += th_rbf(xr, s.xs[i], amp, it) * s.al[i];
                s.pv[lane] = dot + ym;
            }
            __syncthreads();
            double rss = 0.0;
            for (int i = 0; i < k; i++) {
                const double yi = __shfl(yr, i, WAVE);
                const double r = yi - s.pv[i];
                rss += r * r;
            }
            const double rms = ::sqrt(rss / (double)k);
            // reject fits whose minimum over the whole band is > 6 mag from the data
            double pmin = INFINITY;
            for (int q = lane; q < n; q += WAVE) {
                double dot = 0.0;
                for (int i = 0; i < k; i++) dot += th_rbf(t[q], s.xs[i], amp, it) * s.al[i];
                pmin = fmin(pmin, dot + ym);
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) pmin = fmin(pmin, __shfl_xor(pmin, off, WAVE));
            if (__builtin_isfinite(pmin) && fabs(pmin - omin) > 6.0) {
                __syncthreads();
                continue;
            }
            if (__builtin_isfinite(rms) && rms < best) {
                best = rms;
                have = true;
                if (own) {
                    st[lane] = s.al[lane];
                    st[cap + lane] = xr;
                }
                if (lane == 0) {
                    st[2 * cap] = amp;
                    st[2 * cap + 1] = it;
                    st[2 * cap + 2] = ym;
                    st[2 * cap + 3] = (double)k;
                }
            }
            __syncthreads();
        }
    return have;
}

// ThermalCost::cost (src/thermal.rs:116-140) over the compacted colours.
__device__ __forceinline__ double th_cost(const double* obs, int nc, double lref, double lt0, double rate) {
    double chi2 = 0.0;
    for (int i = 0; i < nc; i++) {
        const double* o = obs + 4 * i;
        const double temp = ::pow(10.0, lt0 + rate * o[0]);
        const double mc = bb_color_mag(lref, o[3], temp);
        if (!__builtin_isfinite(mc)) return 1e10;
        const double r = o[1] - mc;
        chi2 += r * r / o[2];
    }
    return chi2 / (double)(nc > 1 ? nc : 1);
}

__device__ __forceinline__ void thermal_source(const ThermalArgs& a, ThLds& s, int src) {
    const int lane = lane_id();
    double* out = a.out + (int64_t)src * 8;
    double* st = a.gp_state + (int64_t)src * (2 * a.cap + 4);
    const int64_t c0 = a.col_off[src], c1 = a.col_off[src + 1];
    const int ncol = (int)(c1 - c0);
    if (lane < 8) out[lane] = NAN;
    for (int i = lane; i < ncol; i += WAVE) a.col_used[c0 + i] = 0;
    if (lane == 0) a.status[src] = 0;
    const double dur = a.duration[src];
    if (!(dur > 0.0)) return;
    if (!(st[2 * a.cap + 3] > 0.0)) {
        const int64_t r0 = a.ref_off[src];
        for (int q = lane; q < 2 * a.cap + 4; q += WAVE) st[q] = 0.0;
        __syncthreads();
        if (!fit_ref_gp(a, s, a.ref_t + r0, a.ref_m + r0, a.ref_e + r0, (int)(a.ref_off[src + 1] - r0), dur, st))
            return;
        __syncthreads();
    }
    const int gm = (int)st[2 * a.cap + 3];
    const double gamp = st[2 * a.cap], git = st[2 * a.cap + 1], gym = st[2 * a.cap + 2];
    // 2. colours, compacted in observation order
    double* obs = a.scratch + 4 * c0;
    int nc = 0;
    for (int q0 = 0; q0 < ncol; q0 += WAVE) {
        const int q = q0 + lane;
        double rm = NAN;
        if (q < ncol) {
            const double tq = a.col_t[c0 + q];
            double dot = 0.0;
            for (int i = 0; i < gm; i++) dot += th_rbf(tq, st[a.cap + i], gamp, git) * st[i];
            rm = dot + gym;
        }
        const bool ok = q < ncol && __builtin_isfinite(rm);
        const unsigned long long bal = __ballot(ok);
        const int below = __popcll(bal & ((1ull << lane) - 1ull));
        if (ok) {
            const double ee = a.col_e[c0 + q];
            const double err = ::sqrt(ee * ee + 0.02 * 0.02);
            double* o = obs + 4 * (nc + below);
            o[0] = a.col_t[c0 + q];
            o[1] = rm - a.col_m[c0 + q];
            o[2] = err * err + 1e-10;
            o[3] = a.col_lambda[c0 + q];
            a.col_used[c0 + q] = 1;
        }
        nc += __popcll(bal);
    }
    if (lane == 7) out[7] = (double)nc;
    if (lane == 0) a.status[src] = 1;
    if (nc < 3) return;
    __syncthreads();  // colours written
    // 3. both swarm runs, lane = 10 run + particle
    const double lref = a.lambda_ref[src];
    const double lo[2] = {3.0, -0.05}, hi[2] = {6.0, 0.01};
    const double w = 1.0 / (2.0 * ::log(2.0)), cc = 0.5 + ::log(2.0), cs = 0.5 + ::log(2.0);
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    const bool act = lane < 2 * TH_P;
    const int run = lane / TH_P, p = lane - run * TH_P;
    double x[2] = {0.0, 0.0}, v[2] = {0.0, 0.0}, pb[2] = {0.0, 0.0}, pc = INFINITY, c = INFINITY;
    if (act) {
        U4 r = philox4x32_10(0u, (uint32_t)(64 * run + p), 0xFFFFFFFFu, TH_TAG, k0, k1);
        const uint32_t ru[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int d = 0; d < 2; d++) {
            const double rg = hi[d] - lo[d];
            x[d] = lo[d] + th_u24(ru[d]) * rg;
            v[d] = -rg + th_u24(ru[2 + d]) * (2.0 * rg);
            pb[d] = x[d];
        }
        c = th_cost(obs, nc, lref, x[0], x[1]);
        pc = c;
    }
    // global best of each run: first minimum over its particles
    double gp[2], gc;
    {
        const int base = run < 2 ? run * TH_P : 0;
        int g = base;
        double bc = __shfl(c, base, WAVE);
        for (int q = 1; q < TH_P; q++) {
            const double cq = __shfl(c, base + q, WAVE);
            if (cq < bc) {
                bc = cq;
                g = base + q;
            }
        }
        gc = bc;
        gp[0] = __shfl(x[0], g, WAVE);
        gp[1] = __shfl(x[1], g, WAVE);
    }
    for (int itn = 0; itn < TH_ITERS; itn++) {
        if (act) {
            U4 r = philox4x32_10(0u, (uint32_t)(64 * run + p), (uint32_t)itn, TH_TAG, k0, k1);
            const uint32_t ru[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const double mom = v[d] * w;
                const double po = (pb[d] - x[d]) * th_u24(ru[d]) * cc;
                const double pg = (gp[d] - x[d]) * th_u24(ru[2 + d]) * cs;
                v[d] = mom + po + pg;
                const double nx = x[d] + v[d];
                x[d] = fmin(fmax(nx, lo[d]), hi[d]);
            }
            c = th_cost(obs, nc, lref, x[0], x[1]);
            if (c < pc) {
                pc = c;
                pb[0] = x[0];
                pb[1] = x[1];
            }
        }
        // a particle below the run's best takes it, first in particle order
        const int base = run < 2 ? run * TH_P : 0;
        int g = -1;
        double bc = gc;
        for (int q = 0; q < TH_P; q++) {
            const double cq = __shfl(c, base + q, WAVE);
            if (cq < bc) {
                bc = cq;
                g = base + q;
            }
        }
        if (g >= 0) {
            gc = bc;
            gp[0] = __shfl(x[0], g, WAVE);
            gp[1] = __shfl(x[1], g, WAVE);
        }
        wave_reconverge();
    }
    // best of the two runs, spread of their best positions
    const double p00 = __shfl(gp[0], 0, WAVE), p01 = __shfl(gp[1], 0, WAVE), c0r = __shfl(gc, 0, WAVE);
    const double p10 = __shfl(gp[0], TH_P, WAVE), p11 = __shfl(gp[1], TH_P, WAVE), c1r = __shfl(gc, TH_P, WAVE);
    if (lane == 0) {
        const bool b1 = c1r < c0r;
        const double bp0 = b1 ? p10 : p00, bp1 = b1 ? p11 : p01, bcst = b1 ? c1r : c0r;
        const double m0 = (p00 + p10) / 2.0, m1 = (p01 + p11) / 2.0;
        const double v0 = ((p00 - m0) * (p00 - m0) + (p10 - m0) * (p10 - m0)) / 2.0;
        const double v1 = ((p01 - m1) * (p01 - m1) + (p11 - m1) * (p11 - m1)) / 2.0;
        double tl = -INFINITY;
        for (int i = 0; i < nc; i++) tl = fmax(tl, obs[4 * i]);
        const double lt = bp0 + bp1 * tl;
        out[0] = bp0;
        out[1] = __builtin_isfinite(lt) ? fmin(fmax(lt, 2.0), 7.0) : NAN;
        out[2] = bp1;
        out[3] = ::sqrt(v0);
        out[4] = ::sqrt(v1);
        out[5] = __builtin_isfinite(bcst) ? bcst : NAN;
        out[6] = bcst;
        a.status[src] = 2;
    }
}

}  // namespace

__global__ __launch_bounds__(WAVE) void thermal_kernel(ThermalArgs a) {
    __shared__ ThLds s;
    thermal_source(a, s, blockIdx.x);
}

hipError_t launch_thermal(const ThermalArgs& a, hipStream_t st) {
    if (a.n_src <= 0) return hipSuccess;
    if (a.cap < TH_MAX_GP) return hipErrorInvalidValue;
    hipLaunchKernelGGL(thermal_kernel, dim3(a.n_src), dim3(WAVE), 0, st, a);
    return hipGetLastError();
}

}  // namespace lcf
