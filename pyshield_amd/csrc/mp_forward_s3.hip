// mp_forward_s3.hip -- the forward kernel for 3 slot(s) of 64 cells per
// wavefront, FP64 and mixed precision, single sub-step (one file per slot
// count so the instantiations compile in parallel; the looping instances are
// in mp_forward_multi_s3.hip).
#include "mp_forward.h"

namespace shield {
extern template hipError_t launch_kernels<3, float, true>(const FwdArgs<float>&, bool, unsigned, hipStream_t);
extern template hipError_t launch_kernels<3, double, true>(const FwdArgs<double>&, bool, unsigned, hipStream_t);
extern template int waves_kernels<3, float, true>(bool);
extern template int waves_kernels<3, double, true>(bool);
template hipError_t launch_kernels<3, float, false>(const FwdArgs<float>&, bool, unsigned, hipStream_t);
template hipError_t launch_kernels<3, double, false>(const FwdArgs<double>&, bool, unsigned, hipStream_t);
template int waves_kernels<3, float, false>(bool);
template int waves_kernels<3, double, false>(bool);
template hipError_t launch_slots<3, float>(const shield_mp_shape&, const shield_mp_params&,
                                            const shield_mp_fields&, const shield_update_fields&, int,
                                            hipStream_t);
template hipError_t launch_slots<3, double>(const shield_mp_shape&, const shield_mp_params&,
                                             const shield_mp_fields&, const shield_update_fields&, int,
                                             hipStream_t);
template int waves_per_simd<3, float>(bool, bool);
template int waves_per_simd<3, double>(bool, bool);
}  // namespace shield
