// svi_f64.hip -- FP64 instantiations of the SVI kernels.  Built with
// -ffp-contract=off so that no multiply-add is fused: the FP64 optimiser then
// rounds like the unfused arithmetic of the reference and the oracle.
#include "svi_job.hpp"

namespace lcf {
const void* svi_kernel_f64(int model, bool fd, bool persistent) { return svi_ptr_m<double>(model, fd, persistent); }
}  // namespace lcf
