// mp_forward_multi_s3.hip -- the looping (more than one sub-step) forward
// kernel for 3 slot(s) of 64 cells per wavefront, FP64 and mixed precision.
#include "mp_forward.h"

namespace shield {
template hipError_t launch_kernels<3, float, true>(const FwdArgs<float>&, bool, unsigned, hipStream_t);
template hipError_t launch_kernels<3, double, true>(const FwdArgs<double>&, bool, unsigned, hipStream_t);
template int waves_kernels<3, float, true>(bool);
template int waves_kernels<3, double, true>(bool);
}  // namespace shield
