// bb_kernels.hpp -- launch interface of the batched blackbody temperature
// grid fit (fit_bb_temperature, FP64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lcf {

struct BbArgs {
    int64_t n_sets;
    const int64_t* off;    // [n_sets + 1] magnitudes per set (one SED)
    const double* lam;     // wavelengths, Angstrom
    const double* mag;
    const double* err;
    double* log_t;         // [n_sets]
    double* chi2;
    double* offset;
    int32_t* ok;           // 1 when the set has >= 2 points
};

hipError_t launch_bb(const BbArgs& a, hipStream_t st);

}  // namespace lcf
