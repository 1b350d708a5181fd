// model_tables.hpp -- per-model constant tables shared by the fitting and
// finalize kernels: PSO box bounds and population priors.
#pragma once
#include <hip/hip_runtime.h>

namespace lcf {

// Box bounds (GpuModelName::pso_bounds, src/gpu_types.rs:64-103) and
// population priors (src/parametric.rs:1588-1621; the Villar entries are the
// log10 table values times ln 10), indexed [model][param].
static __constant__ double c_lo[8][7] = {
    {-3.0, -1.0, -100.0, -2.0, -2.0, -5.0, 0},
    {-1.5, -0.02, 1.0, -60.0, -1.0, 1.0, -6.0},
    {-3.0, -1.0, -100.0, -2.0, -1.0, 0.5, -5.0},
    {-3.0, -100.0, 0.5, -3.0, -5.0, 0, 0},
    {-3.0, -100.0, 0.0, 0.5, -5.0, 0, 0},
    {-3.0, -100.0, 0.1, -1.0, -5.0, 0, 0},
    {-3.0, -100.0, -2.0, -2.0, 0.5, -5.0, 0},
    {-3.0, -2.0, -1.0, -2.0, -5.0, 0, 0}};
static __constant__ double c_hi[8][7] = {
    {3.0, 1.0, 100.0, 5.0, 6.0, 0.0, 0},
    {1.5, 0.05, 5.0, 40.0, 3.0, 5.5, 0.0},
    {3.0, 1.0, 100.0, 5.0, 6.0, 4.0, 0.0},
    {3.0, 100.0, 4.5, 3.0, 0.0, 0, 0},
    {3.0, 100.0, 6.0, 4.5, 0.0, 0, 0},
    {3.0, 100.0, 3.0, 4.0, 0.0, 0, 0},
    {3.0, 100.0, 6.0, 3.0, 5.0, 0.0, 0},
    {-0.5, -0.5, 2.0, 1.0, 0.0, 0, 0}};
#define LN10 2.30258509299404568402
static __constant__ double c_pc[8][7] = {
    {0.0, 0.0, 0.0, 1.5, 3.0, -3.0, 0},
    {-0.009 * LN10, 0.009, 1.327 * LN10, -12.0, 0.360 * LN10, 1.454 * LN10, -1.033 * LN10},
    {0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0, 0}};
static __constant__ double c_pw[8][7] = {
    {2.0, 0.2, 30.0, 2.5, 2.5, 2.0, 0},
    {0.10 * LN10, 0.010, 0.40 * LN10, 15.0, 0.35 * LN10, 0.40 * LN10, 0.50 * LN10},
    {5, 5, 5, 5, 5, 5, 5}, {5, 5, 5, 5, 5, 5, 5}, {5, 5, 5, 5, 5, 5, 5},
    {5, 5, 5, 5, 5, 5, 5}, {5, 5, 5, 5, 5, 5, 5}, {5, 5, 5, 5, 5, 5, 5}};
#undef LN10

// The CPU path's pso_bounds (src/parametric.rs:1541-1579) equals the table
// above except for the Bazin baseline, bounded to +-0.3 there.
constexpr double BAZIN_B_CPU = 0.3;

}  // namespace lcf
