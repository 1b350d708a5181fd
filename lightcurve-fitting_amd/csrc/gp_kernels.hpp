// gp_kernels.hpp -- launch interface of the batched dense GP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace lcf {

// Largest training set factorised in LDS: the packed FP64 triangle of a
// 200-point kernel matrix plus alpha and the training indices fill the 160 KB
// a gfx950 workgroup can address.
constexpr int GP_LDS_CAP = 200;

struct GpArgs {
    const double* t;           // [n_obs] observation times
    const double* w;           // [n_obs] log10 wavelengths (2D fit) or null (1D fit)
    const double* query_w;     // query wavelengths, same shape as query (2D) or null
    const double* lsws;        // wavelength length-scale candidates (2D)
    const double* lsws_band;   // [n_bands][n_lsw] per-source candidates (null: use lsws)
    int n_lsw;                 // 0 in 1D
    const double* y;           // [n_obs] values (magnitudes)
    const double* nv;          // [n_obs] noise variances
    const int64_t* offsets;    // [n_bands + 1]
    const int32_t* order;      // band indices, decreasing size
    int n_tasks;
    int* queue;
    const double* query;       // [n_query] or [n_bands][n_query]
    int n_query;
    int query_per_band;
    const double* amps;
    int n_amp;
    const double* lss;         // [n_ls] shared length-scale candidates, or
    const double* lss_band;    // [n_bands][n_ls] per-band candidates (null: use lss)
    const double* min_ls;      // [n_bands] smallest admissible length scale, or null
    int n_ls;
    int cap;                   // training-set cap (subsample size)
    double ls_floor;           // candidates below these length scales are skipped
    double lw_floor;
    int subsample_rule;        // 0: GPU rule, 1: CPU 1D rule, 2: CPU 2D rule
    int criterion;             // 0: NLML (GPU path), 1: train RMS (CPU dense path)
    double* state;             // [n_bands][state_stride]
    int state_stride;          // >= 2*cap + 4
    double* pred;              // [n_bands][n_query]
    double* stdv;              // [n_bands][n_query]
    double* dpred;             // [n_bands][n_query] or null
    double* pred_obs;          // [n_obs]
};

size_t gp_lds_bytes(int cap);
hipError_t launch_gp_fit(const GpArgs& a, int grid, hipStream_t st);
hipError_t gp_occupancy(int cap, int* blocks_per_cu);

}  // namespace lcf
