// np_kernels.hpp -- launch interface of the batched nonparametric band
// pipeline (GP fit + light-curve features, FP64).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace lcf {

constexpr int NP_NFEAT = 50;       // NonparametricBandResult fields, band name excluded
constexpr int NP_FINE_MAX = 2000;  // 1-day fine grid cap (src/nonparametric.rs:515-518)
constexpr int NP_MAX_IND = 32;     // FITC inducing points held per thread
constexpr int NP_CHUNK = 64;       // observations per FITC accumulation chunk
constexpr int NP_CAP_MAX = 200;    // dense training set factorised in LDS

struct NpArgs {
    const int64_t* offsets;  // [n_bands + 1] raw observations per band
    const double* t;         // [n_obs] times
    const double* mag;       // [n_obs] magnitudes
    const double* err;       // [n_obs] magnitude errors
    const double* t_lo;      // [n_bands] grid start (source's first time)
    const double* t_hi;      // [n_bands] grid end (source's last time)
    const int32_t* order;    // band indices of this launch, decreasing size
    int n_tasks;
    int* queue;
    int n_grid;
    const double* amps;
    int n_amp;
    const double* lsf;       // length-scale divisors: l = max(duration / f, 0.1)
    int n_ls;
    int dense_max;           // dense GP up to this many detections, FITC above
    int max_sub;             // subsample of the dense GP kept with a FITC fit
    int n_ind;               // FITC inducing points (<= NP_MAX_IND)
    int cap;                 // dense capacity of this launch's LDS
    int fitc;                // this launch's LDS holds the FITC work area
    int state_cap;           // state rows: alpha[state_cap], x[state_cap], amp, 1/(2 l^2), mean, m
    double* dx;              // [n_obs] scratch: detections (times, mags, errors), CSR aligned
    double* dy;
    double* de;
    double* buf;             // [n_obs] scratch
    int32_t* status;         // [n_bands] 0 no result, 1 dense, 2 FITC
    double* feat;            // [n_bands][NP_NFEAT]
    double* pred;            // [n_bands][n_grid]
    double* stdv;            // [n_bands][n_grid]
    double* dpred;           // [n_bands][n_grid]
    double* state;           // [n_bands][2 state_cap + 4]
};

size_t np_lds_bytes(int cap, int fitc);
hipError_t launch_np(const NpArgs& a, int grid, hipStream_t st);
hipError_t np_occupancy(int cap, int fitc, int* blocks_per_cu);

}  // namespace lcf
