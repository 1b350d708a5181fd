// svi_f32.hip -- FP32 instantiations of the SVI kernels (compiled with the
// default floating-point contraction).
#include "svi_job.hpp"

namespace lcf {
const void* svi_kernel_f32(int model, bool fd, bool persistent) { return svi_ptr_m<float>(model, fd, persistent); }
}  // namespace lcf
