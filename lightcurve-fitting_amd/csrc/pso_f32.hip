// pso_f32.hip -- FP32 instantiations of the PSO kernels (compiled with the
// default floating-point contraction) and the one-time set-up of the
// censored-likelihood table they read.
#include <mutex>

#define LCF_PSO_F32_UNIT 1
#include "pso_swarm.hpp"

namespace lcf {
const void* pso_kernel_f32(int model, int ppl, bool persistent) { return pick_kernel_r<float>(model, ppl, persistent); }
const void* mb_kernel_f32(int ppl) { return pick_mb_r<float>(ppl); }
const void* cost64_kernel(int model) { return pick_cost64(model); }
const void* pso_mixed_kernel_f32() { return (const void*)pso_mixed_kernel<float, 1>; }

namespace {

// One thread per interval: the reference's FP64 log_normal_cdf at the four
// interpolation nodes, Newton forward differences -> cubic coefficients in u.
__global__ void build_lnphi_table() {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= f32::LNPHI_N) return;
    const double x0 = (double)f32::LNPHI_X0 + (double)i / (double)f32::LNPHI_SCALE;
    const double h = 1.0 / (3.0 * (double)f32::LNPHI_SCALE);
    double f[4];
    for (int k = 0; k < 4; k++) f[k] = -log_normal_cdf<double>(x0 + (double)k * h);
    const double d1 = f[1] - f[0];
    const double d2 = f[2] - 2.0 * f[1] + f[0];
    const double d3 = f[3] - 3.0 * f[2] + 3.0 * f[1] - f[0];
    f32::g_lnphi[i] = make_float4((float)f[0], (float)(3.0 * (d1 - 0.5 * d2 + d3 / 3.0)),
                                  (float)(4.5 * (d2 - d3)), (float)(4.5 * d3));
}

constexpr int MAX_DEVICES = 256;
std::mutex g_mu;
bool g_built[MAX_DEVICES];

}  // namespace

// Builds the table on the current device the first time it is needed and
// waits for it, so that every later launch on any stream sees it.
hipError_t pso_f32_prepare(hipStream_t st) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_built[dev]) return hipSuccess;
    hipLaunchKernelGGL(build_lnphi_table, dim3((f32::LNPHI_N + 63) / 64), dim3(64), 0, st);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) g_built[dev] = true;
    return e;
}
}  // namespace lcf
