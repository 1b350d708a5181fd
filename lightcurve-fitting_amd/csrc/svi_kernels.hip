// svi_kernels.hip -- launch side of the batched SVI kernels (the device code
// lives in svi_job.hpp, instantiated by svi_f32.hip / svi_f64.hip).
#include "svi_kernels.hpp"

namespace lcf {

const void* svi_kernel_f32(int model, bool fd, bool persistent);
const void* svi_kernel_f64(int model, bool fd, bool persistent);

size_t svi_lds_bytes(bool fp64, int tile) { return (size_t)tile * (fp64 ? sizeof(double4) : sizeof(float4)); }

hipError_t launch_svi(int model, const SviArgs& a, bool fp64, bool fd, bool persistent, int grid,
                      hipStream_t st) {
    const void* k = fp64 ? svi_kernel_f64(model, fd, persistent) : svi_kernel_f32(model, fd, persistent);
    if (!k || grid <= 0) return hipErrorInvalidValue;
    if (!fp64) {
        hipError_t e = svi_f32_prepare(st);
        if (e != hipSuccess) return e;
    }
    void* args[] = {const_cast<SviArgs*>(&a)};
    return hipLaunchKernel(k, dim3(grid), dim3(64), args, svi_lds_bytes(fp64, a.tile), st);
}

hipError_t svi_occupancy(int model, bool fp64, bool fd, int tile, int* blocks_per_cu) {
    const void* k = fp64 ? svi_kernel_f64(model, fd, true) : svi_kernel_f32(model, fd, true);
    if (!k) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, 64, svi_lds_bytes(fp64, tile));
}

}  // namespace lcf
