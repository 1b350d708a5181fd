// eval_kernels.hip -- GpuContext::eval_batch: forward model evaluation of many
// parameter draws at many times, FP64, with the reference's expressions
// evaluated as written (src/parametric.rs:199-510 / cuda/models.cu:46-131), so
// that non-finite outputs (overflowing exponentials) match the reference too.
#include "eval_kernels.hpp"
#include "models_fit.hpp"

namespace lcf {

__device__ __forceinline__ double sgm(double x) { return 1.0 / (1.0 + ::exp(-x)); }
__device__ __forceinline__ double sp6(double x) { return ::log(1.0 + ::exp(x)) + 1e-6; }

__device__ double eval_direct(int m, const double* p, double t) {
    switch (m) {
    case 0: {
        double dt = t - p[2];
        return ::exp(p[0]) * ::exp(-dt / ::exp(p[4])) * sgm(dt / ::exp(p[3])) + p[1];
    }
    case 1: {
        double a = ::exp(p[0]), beta = p[1], g = ::exp(p[2]), ph = t - p[3];
        double sr = sgm(ph / ::exp(p[4]));
        double w = sgm(10.0 * (ph - g));
        return a * sr * ((1.0 - w) * (1.0 - beta * ph) + w * ((1.0 - beta * g) * ::exp((g - ph) / ::exp(p[5]))));
    }
    case 2: {
        double ph = t - p[2];
        return ::exp(p[0]) * sgm(ph / ::exp(p[3])) * ::pow(1.0 + sp6(ph) / ::exp(p[4]), -p[5]) + p[1];
    }
    case 3: {
        double ps = sp6(t - p[1]);
        double f = sgm(p[3]);
        double heat = f * ::exp(-ps / 8.8) + (1.0 - f) * ::exp(-ps / 111.3);
        double x = ps / ::exp(p[2]);
        return ::exp(p[0]) * heat * (1.0 - ::exp(-x * x));
    }
    case 4: {
        double ps = sp6(t - p[1]);
        double w = 1.0 + ps / ::exp(p[2]);
        double x = ps / ::exp(p[3]);
        return ::exp(p[0]) * (1.0 / (w * w)) * (1.0 - ::exp(-x * x));
    }
    case 5: {
        double ph = t - p[1];
        double ps = sp6(ph);
        double r = ps / ::exp(p[3]);
        return ::exp(p[0]) * sgm(ph * 5.0) * ::pow(ps, -p[2]) * ::exp(-r * r);
    }
    case 6: {
        double ps = sp6(t - p[1]);
        double lr = ::log(ps / ::exp(p[2]));
        return ::exp(p[0]) * ::pow(::exp(2.0 * p[3] * lr) + ::exp(2.0 * p[4] * lr), -0.5);
    }
    default:
        return 0.0;
    }
}

__global__ void eval_closed_form(int m, const double* __restrict__ params, const double* __restrict__ times,
                                 double* __restrict__ out, int64_t n_draws, int64_t n_times, int np) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_draws * n_times) return;
    int64_t d = idx / n_times, i = idx - d * n_times;
    double p[7];
    for (int j = 0; j < np; j++) p[j] = params[d * np + j];
    out[idx] = eval_direct(m, p, times[i]);
}

// MetzgerKN: one thread per draw integrates the 200-point log grid up to
// 1.05 x the largest phase, normalises by the peak and interpolates
// (src/parametric.rs:544-648, cuda/models.cu:146-282).  The grid lives in the
// thread's private memory, as in the reference.
__global__ void eval_metzger(const double* __restrict__ params, const double* __restrict__ times,
                             double* __restrict__ out, int64_t n_draws, int64_t n_times, int np) {
    int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n_draws) return;
    const double* p = params + d * np;
    const double MSUN = 1.989e33, CCGS = 2.998e10, DAY = 86400.0, PI = 3.14159265358979323846;
    double t0 = p[3];
    double phase_max = 0.01;
    for (int64_t i = 0; i < n_times; i++) phase_max = fmax(phase_max, times[i] - t0);
    double* o = out + d * n_times;
    if (phase_max <= 0.01) {
        for (int64_t i = 0; i < n_times; i++) o[i] = 0.0;
        return;
    }
    const int NG = 200;
    double gt[NG], gl[NG];
    double m_ej = ::pow(10.0, p[0]) * MSUN, v_ej = ::pow(10.0, p[1]) * CCGS, kr = ::pow(10.0, p[2]);
    double lmin = ::log(0.01), lmax = ::log(phase_max * 1.05);
    for (int i = 0; i < NG; i++) gt[i] = ::exp(lmin + (lmax - lmin) * (double)i / (double)(NG - 1));
    const double xn0 = 0.8, scale = 1e40;
    double e0 = 0.5 * m_ej * v_ej * v_ej;
    double e_th = e0 / scale, e_kin = e0 / scale, v = v_ej, r = gt[0] * DAY * v;
    for (int i = 0; i < NG; i++) {
        double td = gt[i], ts = td * DAY;
        double ef = 0.34 * ::pow(td, 0.74);
        double eth = 0.36 * (::exp(-0.56 * td) + (ef > 1e-10 ? ::log(1.0 + ef) / ef : 1.0));
        double xn = xn0 * ::exp(-ts / 900.0);
        double tt = 0.5 - ::atan((ts - 1.3) / 0.11) / PI;
        if (tt < 1e-30) tt = 1e-30;
        double l_heat = m_ej * (3.2e14 * xn + 2e18 * eth * ::pow(tt, 1.3)) / scale;
        double kappa = 0.4 * (xn0 - xn) + kr * (1.0 - xn0);
        double t_diff = 3.0 * kappa * m_ej / (4.0 * PI * CCGS * v * ts) + r / CCGS;
        double l_rad = (e_th > 0.0 && t_diff > 0.0) ? e_th / t_diff : 0.0;
        gl[i] = l_rad;
        double l_pdv = r > 0.0 ? e_th * v / r : 0.0;
        if (i < NG - 1) {
            double dts = (gt[i + 1] - td) * DAY;
            e_th += (l_heat - l_pdv - l_rad) * dts;
            if (e_th < 0.0) e_th = 0.0;
            e_kin += l_pdv * dts;
            v = ::sqrt(2.0 * e_kin * scale / m_ej);
            if (v > CCGS) v = CCGS;
            r += v * dts;
        }
    }
    double lp = -INFINITY;
    for (int i = 0; i < NG; i++) lp = fmax(lp, gl[i]);
    if (lp <= 0.0 || !__builtin_isfinite(lp)) {
        for (int64_t i = 0; i < n_times; i++) o[i] = 0.0;
        return;
    }
    for (int i = 0; i < NG; i++) gl[i] /= lp;
    for (int64_t k = 0; k < n_times; k++) {
        double ph = times[k] - t0, val;
        if (ph <= 0.0) val = 0.0;
        else if (ph <= gt[0]) val = gl[0];
        else if (ph >= gt[NG - 1]) val = gl[NG - 1];
        else {
            int lo = 0, hi = NG - 1;
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (gt[mid] < ph) lo = mid; else hi = mid;
            }
            val = gl[lo] + (ph - gt[lo]) / (gt[hi] - gt[lo]) * (gl[hi] - gl[lo]);
        }
        o[k] = val;
    }
}

hipError_t launch_eval(int model, const double* params, const double* times, double* out,
                       int64_t n_draws, int64_t n_times, int np, hipStream_t st) {
    if (model == 7) {
        int64_t blocks = (n_draws + 63) / 64;
        hipLaunchKernelGGL(eval_metzger, dim3((unsigned)blocks), dim3(64), 0, st, params, times, out,
                           n_draws, n_times, np);
    } else {
        int64_t total = n_draws * n_times;
        int64_t blocks = (total + 255) / 256;
        hipLaunchKernelGGL(eval_closed_form, dim3((unsigned)blocks), dim3(256), 0, st, model, params,
                           times, out, n_draws, n_times, np);
    }
    return hipGetLastError();
}

}  // namespace lcf
