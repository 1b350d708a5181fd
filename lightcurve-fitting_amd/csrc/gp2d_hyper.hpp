// gp2d_hyper.hpp -- launch interface of the 2D GP hyper-parameter fit:
// negative log marginal likelihood, its Cholesky adjoint gradient and an
// L-BFGS driver per source, FP64, one workgroup per source.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lcf {

constexpr int GPH_TILE = 64;      // GEMM tile and Cholesky block width
constexpr int GPH_MAX_N = 2047;   // points per source
constexpr int GPH_THREADS = 256;

struct Gp2dHyperArgs {
    const int64_t* off;       // [n_src + 1]
    const double* t;          // time
    const double* w;          // log10 wavelength
    const double* y;          // magnitude
    const double* nv;         // noise variance
    const int32_t* order;     // sources by decreasing size
    int n_tasks;
    int* queue;
    const double* theta0;     // [n_src][4] ln (a, l_t, l_w, sigma_n)
    int max_iters;            // < 0: NLML and gradient at theta0 only
    int fd_grad;              // 1: L-BFGS on central finite differences of the NLML (the plain form)
    double gtol;
    int ld;                   // leading dimension of the slot matrices (multiple of 64)
    double* work;             // [slots][2 ld ld + 2 ld]
    const double* qt;         // queries (may be null)
    const double* qw;
    int n_query;
    int q_per_src;            // 1: queries [n_src][n_query]; 0: shared
    double* theta;            // [n_src][4]
    double* nlml;             // [n_src]
    double* grad;             // [n_src][4]
    int32_t* iters;           // [n_src]
    int32_t* status;          // [n_src]
    int32_t* evals;           // [n_src][2] NLML and gradient evaluations (may be null)
    double* pred;             // [n_src][n_query] (may be null)
    double* stdv;             // [n_src][n_query]
};

// Doubles per workgroup slot: A (K, then L), X (L^-1), z, alpha.
__host__ __device__ inline size_t gp2d_hyper_slot_doubles(int ld) {
    return 2 * (size_t)ld * ld + 2 * (size_t)ld;
}
hipError_t launch_gp2d_hyper(const Gp2dHyperArgs& a, int grid, hipStream_t st);
hipError_t gp2d_hyper_occupancy(int* blocks_per_cu);

}  // namespace lcf
