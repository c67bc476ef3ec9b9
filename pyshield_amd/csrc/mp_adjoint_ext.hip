// mp_adjoint_ext.hip -- the adjoint kernel's instances with fix_negative, the
// ice slope limiter and the w transport (fix_negative, z_slope_ice,
// do_sedi_w), compiled apart from the default ones in mp_adjoint.hip so the
// two compile in parallel.
#include "mp_adjoint.h"

namespace shield {
template hipError_t launch_adjoint_kernel<1, true>(unsigned, const AdjArgs&, hipStream_t);
template hipError_t launch_adjoint_kernel<2, true>(unsigned, const AdjArgs&, hipStream_t);
template hipError_t launch_adjoint_kernel<3, true>(unsigned, const AdjArgs&, hipStream_t);
}  // namespace shield
