// svi_kernels.hpp -- launch interface of the batched SVI kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace lcf {

constexpr uint32_t LCF_SVI_CPU_STOPPING_FLAG = 1u;  // min 200 / stall 50 (svi_fit constants)
constexpr uint32_t LCF_SVI_INFLATE_FLAG = 2u;       // add ln 4 to the returned log sigma
constexpr uint32_t LCF_SVI_FIXED_STEPS_FLAG = 8u;   // no early stop: all n_steps steps
constexpr uint32_t LCF_SVI_EXACT_TAIL_FLAG = 16u;   // upper-limit adjoint = d log_normal_cdf / dz
constexpr uint32_t LCF_SVI_SHARED_STREAM_FLAG = 32u; // every job draws the noise of job 0

struct SviArgs {
    const void* obs;           // packed {t, y, var, flag} records (float4 / double4)
    const int64_t* offsets;    // [n_jobs + 1]
    const int32_t* order;      // job indices of this launch's model, decreasing size
    int n_tasks;
    int task_base;
    int* queue;
    const double* mu_in;       // [n_jobs][7] initial means (PSO point)
    const double* ls_in;       // [n_jobs][7] log sigma (gradient mode) or null (= -1)
    const double* prior_c;     // [n_jobs][7]
    const double* prior_w;     // [n_jobs][7]
    int n_steps, n_samples;
    double lr;
    uint64_t seed;
    uint32_t flags;
    int tile;
    int grad_only;             // evaluate one step's ELBO estimate and gradient only
    int step0;                 // noise step index used in gradient mode
    double* out_mu;            // [n_jobs][7] (gradient mode: dELBO/dmu)
    double* out_ls;            // [n_jobs][7] (gradient mode: dELBO/dlog sigma)
    double* out_elbo;          // [n_jobs]
    int32_t* out_steps;        // [n_jobs] or null
};

size_t svi_lds_bytes(bool fp64, int tile);
// Builds the FP32 kernels' upper-limit tail table on the current device the
// first time it is called there (one launch, then a wait on `st`); later
// calls only check a flag.  launch_svi calls it before every FP32 launch.
hipError_t svi_f32_prepare(hipStream_t st);
hipError_t launch_svi(int model, const SviArgs& a, bool fp64, bool fd, bool persistent, int grid,
                      hipStream_t st);
hipError_t svi_occupancy(int model, bool fp64, bool fd, int tile, int* blocks_per_cu);

}  // namespace lcf
