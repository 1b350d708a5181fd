// thermal_kernels.hpp -- launch interface of the batched blackbody
// colour-temperature fit (fit_thermal, FP64).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace lcf {

constexpr int TH_MAX_GP = 32;  // reference-band GP points held one row per lane

struct ThermalArgs {
    int n_src;
    const int64_t* ref_off;    // [n_src + 1] reference-band observations
    const double* ref_t;
    const double* ref_m;
    const double* ref_e;
    const int64_t* col_off;    // [n_src + 1] other bands' observations
    const double* col_t;
    const double* col_m;
    const double* col_e;
    const double* col_lambda;  // effective wavelength of each observation's band
    const double* lambda_ref;  // [n_src]
    const double* duration;    // [n_src]
    double* gp_state;          // [n_src][2 cap + 4], m = 0: fit the reference GP here
    int cap;
    uint64_t seed;
    double* scratch;           // [4 * n_col]: colour observations (t, colour, error^2 + 1e-10, lambda), compacted
    int32_t* status;           // [n_src] 0 None, 1 no fit (< 3 colours), 2 fitted
    double* out;               // [n_src][8]
    int32_t* col_used;         // [n_col]
};

hipError_t launch_thermal(const ThermalArgs& a, hipStream_t st);

}  // namespace lcf
