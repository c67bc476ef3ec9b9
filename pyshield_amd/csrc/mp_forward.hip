// mp_forward.hip -- layout choice and dispatch of the fused microphysics
// forward kernel (the kernel itself is in mp_forward.h, instantiated per slot
// count in mp_forward_s<S>.hip).
#include <stdlib.h>

#include "mp_common.h"

namespace shield {

template <int S, class R>
hipError_t launch_slots(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f, int P,
                        hipStream_t stream);

struct LayoutChoice {
  int S, P;
};

// Columns per wavefront P and slots S for nz levels: the compiled slot count
// that leaves the fewest idle lanes (P = floor(64 S / nz) whole columns in
// 64 S lanes).  SHIELD_MP_SLOTS=<S> forces a slot count (tuning).
LayoutChoice choose_layout(int64_t nz) {
  LayoutChoice best{0, 0};
  if (nz < 1) return best;
  static const int forced = [] {
    const char* e = getenv("SHIELD_MP_SLOTS");
    const int v = e ? atoi(e) : 0;
    return (v >= 1 && v <= 5) ? v : 0;
  }();
  double best_util = -1.0;
  for (int S = 1; S <= 5; ++S) {
    if (forced && S != forced) continue;
    const int64_t P = (64 * S) / nz;
    if (P < 1) continue;
    const int64_t Pc = P > pmax_for(S) ? pmax_for(S) : P;
    const double util = double(Pc * nz) / double(64 * S);
    if (util > best_util + 1e-9) {
      best_util = util;
      best = {S, (int)Pc};
    }
  }
  return best;
}

template <class R>
static hipError_t launch_r(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f,
                           LayoutChoice lc, hipStream_t stream) {
  switch (lc.S) {
    case 1: return launch_slots<1, R>(sh, p, f, lc.P, stream);
    case 2: return launch_slots<2, R>(sh, p, f, lc.P, stream);
    case 3: return launch_slots<3, R>(sh, p, f, lc.P, stream);
    case 4: return launch_slots<4, R>(sh, p, f, lc.P, stream);
    case 5: return launch_slots<5, R>(sh, p, f, lc.P, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_forward(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f, int mode,
                          hipStream_t stream) {
  const LayoutChoice lc = choose_layout(sh.nk);
  if (lc.S == 0) return hipErrorInvalidValue;
  if (mode == SHIELD_MP_FP64) return launch_r<double>(sh, p, f, lc, stream);
  return launch_r<float>(sh, p, f, lc, stream);
}

}  // namespace shield
