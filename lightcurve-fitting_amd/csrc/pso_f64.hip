// pso_f64.hip -- FP64 instantiations of the PSO kernels.  Built with
// -ffp-contract=off so that no multiply-add is fused: the FP64 swarm then
// rounds like the reference's (and the oracle's) unfused arithmetic.
#include "pso_swarm.hpp"

namespace lcf {
const void* pso_kernel_f64(int model, int ppl, bool persistent) { return pick_kernel_r<double>(model, ppl, persistent); }
const void* mb_kernel_f64(int ppl) { return pick_mb_r<double>(ppl); }
}  // namespace lcf
