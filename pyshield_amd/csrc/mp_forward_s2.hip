// mp_forward_s2.hip -- the forward kernel for 2 slot(s) of 64 cells per
// wavefront, FP64 and mixed precision, single sub-step (one file per slot
// count so the instantiations compile in parallel; the looping instances are
// in mp_forward_multi_s2.hip).
#include "mp_forward.h"

namespace shield {
extern template hipError_t launch_kernels<2, float, true>(const FwdArgs<float>&, bool, unsigned, hipStream_t);
extern template hipError_t launch_kernels<2, double, true>(const FwdArgs<double>&, bool, unsigned, hipStream_t);
extern template int waves_kernels<2, float, true>(bool);
extern template int waves_kernels<2, double, true>(bool);
template hipError_t launch_kernels<2, float, false>(const FwdArgs<float>&, bool, unsigned, hipStream_t);
template hipError_t launch_kernels<2, double, false>(const FwdArgs<double>&, bool, unsigned, hipStream_t);
template int waves_kernels<2, float, false>(bool);
template int waves_kernels<2, double, false>(bool);
template hipError_t launch_slots<2, float>(const shield_mp_shape&, const shield_mp_params&,
                                            const shield_mp_fields&, const shield_update_fields&, int,
                                            hipStream_t);
template hipError_t launch_slots<2, double>(const shield_mp_shape&, const shield_mp_params&,
                                             const shield_mp_fields&, const shield_update_fields&, int,
                                             hipStream_t);
template int waves_per_simd<2, float>(bool, bool);
template int waves_per_simd<2, double>(bool, bool);
}  // namespace shield
