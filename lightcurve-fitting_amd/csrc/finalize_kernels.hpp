// finalize_kernels.hpp -- launch interface of the batched finalize kernel:
// profile refinement of the posterior (t0 grid with amplitude refit, per
// parameter scans) and the magnitude-space chi2 of each band's fit (FP64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lcf {

// Bands longer than this many observations are rejected (whole band in LDS).
constexpr int FIN_MAX_OBS = 4096;

constexpr int FIN_T0 = 1;        // profile_t0_refine
constexpr int FIN_ALL = 2;       // profile_refine_all
constexpr int FIN_MAG_CHI2 = 4;  // magnitude chi2
constexpr int FIN_SVI_CHI2 = 8;  // chi2 from the posterior mean only (no PSO fallback)

struct FinalizeArgs {
    const void* obs;          // packed double4 {t, flux or upper flux, obs var, upper flag}
    const int64_t* offsets;   // [n_bands + 1]
    const int32_t* order;     // band indices of this launch's model
    int n_tasks;
    int* queue;
    int max_n;
    int flags;
    const double* flux;       // [n_obs] normalised flux (upper limits included)
    const double* flux_err;   // [n_obs] normalised flux error
    const double* mag_obs;    // [n_obs] observed magnitude
    const double* peak;       // [n_bands] band peak flux
    const double* pso;        // [n_bands][7]
    const double* mu_in;      // [n_bands][7]
    const double* ls_in;      // [n_bands][7]
    double* mu_out;           // [n_bands][7]
    double* ls_out;           // [n_bands][7]
    double* mag_chi2;         // [n_bands]
};

hipError_t launch_finalize(int model, const FinalizeArgs& a, int grid, hipStream_t st);
hipError_t finalize_occupancy(int model, int max_n, int* blocks_per_cu);

}  // namespace lcf
