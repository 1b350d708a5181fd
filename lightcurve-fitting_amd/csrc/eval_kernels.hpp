// eval_kernels.hpp -- launch interface of the forward-evaluation kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lcf {

hipError_t launch_eval(int model, const double* params, const double* times, double* out,
                       int64_t n_draws, int64_t n_times, int np, hipStream_t st);

}  // namespace lcf
