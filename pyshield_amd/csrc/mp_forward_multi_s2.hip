// mp_forward_multi_s2.hip -- the looping (more than one sub-step) forward
// kernel for 2 slot(s) of 64 cells per wavefront, FP64 and mixed precision.
#include "mp_forward.h"

namespace shield {
template hipError_t launch_kernels<2, float, true>(const FwdArgs<float>&, bool, unsigned, hipStream_t);
template hipError_t launch_kernels<2, double, true>(const FwdArgs<double>&, bool, unsigned, hipStream_t);
template int waves_kernels<2, float, true>(bool);
template int waves_kernels<2, double, true>(bool);
}  // namespace shield
