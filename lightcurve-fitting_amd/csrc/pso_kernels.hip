// pso_kernels.hip -- launch side of the batched PSO kernels (the device code
// lives in pso_swarm.hpp, instantiated by pso_f32.hip / pso_f64.hip) and the
// per-band model selection kernel.
#include "pso_kernels.hpp"

namespace lcf {

const void* pso_kernel_f32(int model, int ppl, bool persistent);
const void* pso_kernel_f64(int model, int ppl, bool persistent);
const void* mb_kernel_f32(int ppl);
const void* mb_kernel_f64(int ppl);
const void* cost64_kernel(int model);
const void* pso_mixed_kernel_f32();

static const void* pick_kernel(int model, bool fp64, int ppl, bool persistent) {
    return fp64 ? pso_kernel_f64(model, ppl, persistent) : pso_kernel_f32(model, ppl, persistent);
}

// Best model per band over the fitted slots: strict '<' in slot order
// (src/gpu_cuda.rs:595-616), gated slots included only when they ran.
__global__ void select_best(const double* __restrict__ mp, const double* __restrict__ mc,
                            const int32_t* __restrict__ models, int n_models, int n_bands,
                            int32_t* best_model, double* best_params, double* best_cost) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_bands) return;
    int bs = 0;
    double bc = mc[b];
    for (int s = 1; s < n_models; s++) {
        double c = mc[(int64_t)s * n_bands + b];
        if (c < bc) { bc = c; bs = s; }
    }
    if (best_model) best_model[b] = models[bs];
    if (best_cost) best_cost[b] = bc;
    if (best_params)
        for (int d = 0; d < 7; d++) best_params[(int64_t)b * 7 + d] = mp[((int64_t)bs * n_bands + b) * 7 + d];
}

size_t pso_lds_bytes(bool fp64, int tile) {
    return (size_t)tile * (fp64 ? sizeof(double4) : sizeof(float4));
}

hipError_t launch_pso(int model, const PsoArgs& a, bool fp64, int ppl, bool persistent, int grid,
                      hipStream_t st) {
    const void* k = pick_kernel(model, fp64, ppl, persistent);
    if (!k || grid <= 0) return hipErrorInvalidValue;
    if (!fp64) {
        hipError_t e = pso_f32_prepare(st);
        if (e != hipSuccess) return e;
    }
    void* args[] = {const_cast<PsoArgs*>(&a)};
    return hipLaunchKernel(k, dim3(grid), dim3(64), args, pso_lds_bytes(fp64, a.tile), st);
}

hipError_t launch_pso_mixed(const PsoArgs& a, const PsoMix& mx, int grid, hipStream_t st) {
    if (grid <= 0 || mx.n < 1 || mx.n > 8) return hipErrorInvalidValue;
    for (int i = 0; i < mx.n; i++)
        if (mx.model[i] < 0 || mx.model[i] > 6) return hipErrorInvalidValue;
    hipError_t e = pso_f32_prepare(st);
    if (e != hipSuccess) return e;
    void* args[] = {const_cast<PsoArgs*>(&a), const_cast<PsoMix*>(&mx)};
    return hipLaunchKernel(pso_mixed_kernel_f32(), dim3(grid), dim3(64), args, pso_lds_bytes(false, a.tile), st);
}

hipError_t pso_mixed_occupancy(int tile, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, pso_mixed_kernel_f32(), 64,
                                                        pso_lds_bytes(false, tile));
}

hipError_t launch_cost64(int model, const PsoArgs& a, hipStream_t st) {
    const void* k = cost64_kernel(model);
    if (!k) return hipSuccess;  // no FP32 cost-loop form: nothing to re-cost
    if (a.n_bands <= 0 || !a.obs64) return hipErrorInvalidValue;
    void* args[] = {const_cast<PsoArgs*>(&a)};
    return hipLaunchKernel(k, dim3(a.n_bands), dim3(64), args, 0, st);
}

hipError_t pso_occupancy(int model, bool fp64, int ppl, int tile, int* blocks_per_cu) {
    const void* k = pick_kernel(model, fp64, ppl, true);
    if (!k) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, 64, pso_lds_bytes(fp64, tile));
}

hipError_t launch_select_best(const double* mp, const double* mc, const int32_t* models, int n_models,
                              int n_bands, int32_t* best_model, double* best_params,
                              double* best_cost, hipStream_t st) {
    hipLaunchKernelGGL(select_best, dim3((n_bands + 255) / 256), dim3(256), 0, st, mp, mc, models,
                       n_models, n_bands, best_model, best_params, best_cost);
    return hipGetLastError();
}

hipError_t launch_multi_bazin(const MbArgs& a, bool fp64, int ppl, int grid, hipStream_t st) {
    const void* k = fp64 ? mb_kernel_f64(ppl) : mb_kernel_f32(ppl);
    if (!k || grid <= 0) return hipErrorInvalidValue;
    if (!fp64) {
        hipError_t e = pso_f32_prepare(st);
        if (e != hipSuccess) return e;
    }
    void* args[] = {const_cast<MbArgs*>(&a)};
    return hipLaunchKernel(k, dim3(grid), dim3(64), args, pso_lds_bytes(fp64, a.tile), st);
}

hipError_t multi_bazin_occupancy(bool fp64, int ppl, int tile, int* blocks_per_cu) {
    const void* k = fp64 ? mb_kernel_f64(ppl) : mb_kernel_f32(ppl);
    if (!k) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, 64, pso_lds_bytes(fp64, tile));
}

}  // namespace lcf
