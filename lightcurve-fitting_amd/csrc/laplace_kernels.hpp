// laplace_kernels.hpp -- launch interface of the batched Laplace kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lcf {

// Bands longer than this many observations are rejected (whole band in LDS).
constexpr int LAPLACE_MAX_OBS = 4096;

struct LaplaceArgs {
    const void* obs;          // packed double4 {t, y, var, flag}
    const int64_t* offsets;   // [n_jobs + 1]
    const int32_t* order;     // job indices of this launch's model
    int n_tasks;
    int* queue;
    int max_n;                // largest band (LDS records)
    int mode;                 // 0: 4-point FD of f (reference), 1: FD of the analytic gradient
    const double* map;        // [n_jobs][7]
    double* out_log_sigma;    // [n_jobs][7]
    double* out_elbo;         // [n_jobs]
    double* out_hessian;      // [n_jobs][36] or null
};

hipError_t launch_laplace(int model, const LaplaceArgs& a, int grid, hipStream_t st);
hipError_t laplace_occupancy(int model, int max_n, int* blocks_per_cu);

}  // namespace lcf
