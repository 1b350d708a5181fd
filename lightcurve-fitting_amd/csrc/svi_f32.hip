// svi_f32.hip -- FP32 instantiations of the SVI kernels (compiled with the
// default floating-point contraction) and the one-time set-up of the
// upper-limit tail table they read.
#include <mutex>

#define LCF_SVI_F32_UNIT 1
#include "svi_job.hpp"

namespace lcf {
const void* svi_kernel_f32(int model, bool fd, bool persistent) { return svi_ptr_m<float>(model, fd, persistent); }

namespace {

// One thread per interval: the FP64 log_normal_cdf and dlog_normal_cdf at the
// four interpolation nodes, Newton forward differences -> cubic coefficients
// in u (the same construction for both functions).
__device__ float4 svi_cubic(const double (&f)[4]) {
    const double d1 = f[1] - f[0];
    const double d2 = f[2] - 2.0 * f[1] + f[0];
    const double d3 = f[3] - 3.0 * f[2] + 3.0 * f[1] - f[0];
    return make_float4((float)f[0], (float)(3.0 * (d1 - 0.5 * d2 + d3 / 3.0)), (float)(4.5 * (d2 - d3)),
                       (float)(4.5 * d3));
}

__global__ void build_svi_tail_table() {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= SVI_TAIL_N) return;
    const double z0 = (double)SVI_TAIL_Z0 + (double)i / (double)SVI_TAIL_SCALE;
    const double h = 1.0 / (3.0 * (double)SVI_TAIL_SCALE);
    double v[4], d[4];
    for (int k = 0; k < 4; k++) {
        const double z = z0 + (double)k * h;
        v[k] = log_normal_cdf<double>(z);
        d[k] = dlog_normal_cdf<double>(z);
    }
    g_svi_tail[2 * i] = svi_cubic(v);
    g_svi_tail[2 * i + 1] = svi_cubic(d);
}

constexpr int MAX_DEVICES = 256;
std::mutex g_mu;
bool g_built[MAX_DEVICES];

}  // namespace

// Builds the table on the current device the first time it is needed and
// waits for it, so that every later launch on any stream sees it.
hipError_t svi_f32_prepare(hipStream_t st) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_built[dev]) return hipSuccess;
    hipLaunchKernelGGL(build_svi_tail_table, dim3((SVI_TAIL_N + 63) / 64), dim3(64), 0, st);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) g_built[dev] = true;
    return e;
}
}  // namespace lcf
