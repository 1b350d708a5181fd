// pso_kernels.hpp -- host-visible launch interface of the batched PSO kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace lcf {

struct PsoArgs {
    const void* obs;            // packed records: float4 (FP32) or double4 (FP64) {t, y, var, flag}
    const double4* obs64;       // the FP64 records (the FP32 swarms' final FP64 cost)
    const int64_t* offsets;     // [n_bands + 1] CSR offsets into obs
    const int32_t* order;       // [n_tasks] band indices, decreasing size
    int n_tasks;
    int task_base;              // first task of a one-shot launch
    int n_bands;
    int* queue;                 // persistent-launch work counter (zeroed before launch)
    int slot;                   // output slot of this model
    const double* gate_cost;    // slot-0 costs when gated (skip band if < gate), else null
    double gate;
    int n_particles, max_iters, stall_iters;
    uint64_t seed;
    uint32_t flags;             // bit 0: stall baseline from the first evaluation
    int tile;                   // LDS observation tile (records per wave)
    double* model_params;       // [n_slots][n_bands][7]
    double* model_cost;         // [n_slots][n_bands]
    int32_t* model_iters;       // [n_slots][n_bands]
};

// Model families of one mixed launch (models 0-6, in output-slot order).
struct PsoMix {
    int n;
    int model[8];
};

// Greedy MultiBazin fit (batch_pso_multi_bazin, src/gpu_cuda.rs:650-963).
struct MbArgs {
    const void* obs;            // packed records of the swarm's precision
    const double4* obs64;       // FP64 records (residual seeding)
    const int64_t* offsets;
    const int32_t* order;       // band indices, decreasing size
    int n_tasks;
    int* queue;
    double t_lo, t_hi;          // t0 bounds: global time range -/+ 30 days
    int n_particles, max_iters, stall_iters;
    uint64_t seed;
    int tile;
    int32_t* best_k;            // [n_bands]
    double* params;             // [n_bands][18] best-K parameters, zero padded
    double* cost;               // [n_bands]
    double* bic;                // [n_bands]
    double* per_k_cost;         // [n_bands][4] (NaN after the early stop)
    double* per_k_bic;          // [n_bands][4]
};

constexpr int MB_MAX_K = 4;
constexpr int MB_MAX_PARAMS = 4 * MB_MAX_K + 2;

hipError_t launch_multi_bazin(const MbArgs& a, bool fp64, int ppl, int grid, hipStream_t st);
hipError_t multi_bazin_occupancy(bool fp64, int ppl, int tile, int* blocks_per_cu);

size_t pso_lds_bytes(bool fp64, int tile);
hipError_t launch_pso(int model, const PsoArgs& a, bool fp64, int ppl, bool persistent, int grid,
                      hipStream_t st);
hipError_t pso_occupancy(int model, bool fp64, int ppl, int tile, int* blocks_per_cu);
// One persistent FP32 launch over every (band, model) task of mx (models 0-6,
// one particle per lane; a.n_tasks = bands x mx.n, a.order the band order).
hipError_t launch_pso_mixed(const PsoArgs& a, const PsoMix& mx, int grid, hipStream_t st);
hipError_t pso_mixed_occupancy(int tile, int* blocks_per_cu);
// FP64 re-costing of the best positions of an FP32 swarm launch (slot a.slot,
// every band of the batch; models 0-6, others a no-op).
hipError_t launch_cost64(int model, const PsoArgs& a, hipStream_t st);
// Builds the FP32 cost loop's censored-likelihood table on the current device
// the first time it is called there (one launch, then a wait on `st`); later
// calls only check a flag.
hipError_t pso_f32_prepare(hipStream_t st);
hipError_t launch_select_best(const double* mp, const double* mc, const int32_t* models, int n_models,
                              int n_bands, int32_t* best_model, double* best_params,
                              double* best_cost, hipStream_t st);

}  // namespace lcf
