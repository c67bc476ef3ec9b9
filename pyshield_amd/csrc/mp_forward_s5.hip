// mp_forward_s5.hip -- the forward kernel for 5 slot(s) of 64 cells per
// wavefront, FP64 and mixed precision (one file per slot count so the
// instantiations compile in parallel).
#include "mp_forward.h"

namespace shield {
template hipError_t launch_slots<5, float>(const shield_mp_shape&, const shield_mp_params&,
                                            const shield_mp_fields&, const shield_update_fields&, int,
                                            hipStream_t);
template hipError_t launch_slots<5, double>(const shield_mp_shape&, const shield_mp_params&,
                                             const shield_mp_fields&, const shield_update_fields&, int,
                                            hipStream_t);
template int waves_per_simd<5, float>(bool, bool);
template int waves_per_simd<5, double>(bool, bool);
}  // namespace shield
