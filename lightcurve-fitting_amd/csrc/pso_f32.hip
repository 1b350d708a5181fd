// pso_f32.hip -- FP32 instantiations of the PSO kernels (compiled with the
// default floating-point contraction).
#include "pso_swarm.hpp"

namespace lcf {
const void* pso_kernel_f32(int model, int ppl, bool persistent) { return pick_kernel_r<float>(model, ppl, persistent); }
const void* mb_kernel_f32(int ppl) { return pick_mb_r<float>(ppl); }
}  // namespace lcf
