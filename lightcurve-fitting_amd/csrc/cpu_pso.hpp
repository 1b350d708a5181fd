// cpu_pso.hpp -- launch interface of the CPU-path fitting engine: the
// reference's pso_minimize / pso_minimize_seeded (asynchronous global best),
// its per-band and joint multi-band costs, and the L-BFGS polish, one
// wavefront per job (FP64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lcf {

constexpr int CPSO_MAX_DIM = 64;        // search dimensions (lanes)
constexpr int CPSO_MAX_PARTICLES = 256;
constexpr int CPSO_MAX_OBS = 4096;      // observations of one job (LDS)
constexpr int CPSO_MAX_BANDS = 9;       // joint fit: 7 + 7 (bands - 1) <= 64

enum CpsoKind : int32_t {
    CPSO_SWARM_BAND = 0,   // PsoCost of one band (src/parametric.rs:1303-1370)
    CPSO_SWARM_JOINT = 1,  // joint Villar cost of a band group (:3905-4000)
    CPSO_POLISH = 2,       // lbfgs_refine from x0 (:1453-1497)
    CPSO_SWARM_MULTIBAZIN = 3,  // multi_bazin_cost of one band, K = model components (:983-1006)
};

struct CpuPsoJob {
    int32_t kind;
    int32_t model;        // model id (CPSO_SWARM_BAND, CPSO_POLISH); components K (CPSO_SWARM_MULTIBAZIN)
    int32_t band;         // first band (batch index)
    int32_t n_bands;      // joint: consecutive bands, reference band first
    int32_t dim;          // search dimensions
    int32_t n_particles;
    int32_t max_iters;    // swarm iterations (polish: L-BFGS iterations)
    int32_t stall_iters;
    int32_t use_prior;    // population priors in PsoCost
    int32_t fixed_t0;     // t0 held at t0_value and left out of the search vector
    int32_t seeded;       // particle 0 starts at x0
    uint32_t stream;      // Philox counter word (band or group index)
    uint64_t seed;        // Philox key (the reference's SmallRng seed)
    double t0_value;
    double start_cost;    // polish: cost at x0
};

struct CpuPsoArgs {
    const void* obs;             // packed double4 {t, flux or upper flux, obs var, upper flag}
    const int64_t* offsets;      // [n_bands + 1]
    const CpuPsoJob* jobs;
    int n_jobs;
    int* queue;
    const double* lo;            // [n_jobs][64]
    const double* hi;
    const double* x0;
    double* scratch;             // [grid][slot_doubles] swarm state
    int64_t slot_doubles;
    int max_obs;
    double* out_x;               // [n_jobs][64]
    double* out_cost;            // [n_jobs]
    int32_t* out_iters;          // [n_jobs]
};

hipError_t launch_cpu_pso(const CpuPsoArgs& a, int grid, hipStream_t st);
hipError_t cpu_pso_occupancy(int max_obs, int* blocks_per_cu);

}  // namespace lcf
