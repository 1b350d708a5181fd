// bb_kernels.hip -- blackbody temperature of many SEDs at once, FP64:
// fit_bb_temperature (src/gp2d.rs:275-324) with planck_relative_mag
// (:264-271), one thread per SED: 200 log10 T values on [3, 5.5], the
// weighted least-squares magnitude offset and chi2 / max(n - 2, 1) at each,
// first strict minimum.  Each thread runs the reference's loops in order
// (built without multiply-add contraction).
#include "bb_kernels.hpp"

namespace lcf {

namespace {

constexpr double BB_HC_OVER_K = 1.4387769e8;  // h c / k_B, Angstrom K (src/gp2d.rs:267)
constexpr double BB_LN_10 = 2.302585092994045684017991454684364208;

__device__ __forceinline__ double planck_relative_mag(double lam, double temp) {
    const double x = BB_HC_OVER_K / (lam * temp);
    const double log_b = -5.0 * ::log(lam) - ::log(::exp(x) - 1.0);
    return -2.5 * log_b / BB_LN_10;
}

}  // namespace

__global__ __launch_bounds__(256) void bb_fit_kernel(BbArgs a) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.n_sets) return;
    const int64_t o0 = a.off[s];
    const int n = (int)(a.off[s + 1] - o0);
    const double* lam = a.lam + o0;
    const double* mag = a.mag + o0;
    const double* err = a.err + o0;
    if (n < 2) {
        a.ok[s] = 0;
        a.log_t[s] = a.chi2[s] = a.offset[s] = NAN;
        return;
    }
    double best_chi2 = INFINITY, best_log_t = 4.0, best_off = 0.0;
    const double nf = (double)n;
    for (int i = 0; i < 200; i++) {
        const double log_t = 3.0 + (double)i * 2.5 / (200.0 - 1.0);
        const double temp = ::pow(10.0, log_t);
        double sw = 0.0, swd = 0.0;
        for (int j = 0; j < n; j++) {
            const double w = 1.0 / fmax(err[j] * err[j], 1e-10);
            sw += w;
            swd += w * (mag[j] - planck_relative_mag(lam[j], temp));
        }
        const double off = swd / sw;
        double chi = 0.0;
        for (int j = 0; j < n; j++) {
            const double r = mag[j] - planck_relative_mag(lam[j], temp) - off;
            chi += r * r / fmax(err[j] * err[j], 1e-10);
        }
        chi = chi / fmax(nf - 2.0, 1.0);
        if (chi < best_chi2) {
            best_chi2 = chi;
            best_log_t = log_t;
            best_off = off;
        }
    }
    a.ok[s] = 1;
    a.log_t[s] = best_log_t;
    a.chi2[s] = best_chi2;
    a.offset[s] = best_off;
}

hipError_t launch_bb(const BbArgs& a, hipStream_t st) {
    if (a.n_sets <= 0) return hipSuccess;
    const int64_t nb = (a.n_sets + 255) / 256;
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(bb_fit_kernel, dim3((unsigned)nb), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace lcf
