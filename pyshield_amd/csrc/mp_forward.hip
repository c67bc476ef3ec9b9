// mp_forward.hip -- layout choice and dispatch of the fused microphysics
// forward kernel (the kernel itself is in mp_forward.h, instantiated per slot
// count in mp_forward_s<S>.hip).
#include <stdlib.h>

#include <atomic>

#include "mp_common.h"

namespace shield {

template <int S, class R>
hipError_t launch_slots(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f,
                        const shield_update_fields& u, int P, hipStream_t stream);

template <int S, class R>
int waves_per_simd(bool multi, bool fused);

struct LayoutChoice {
  int S, P;
};

// Throughput weight of a kernel instance by resident waves per SIMD.  The
// kernel is issue-bound with long dependent chains, so a second and third
// wave hide latency; the weights are fitted to C384 x 79 on MI355X, where
// 3 slots at 3 waves (82% of lanes busy) beat 5 slots at 2 waves (99%) in
// mixed precision and lost to them in FP64 (2 waves each).
static double occupancy_weight(int waves) {
  if (waves <= 0) return 1.0;  // unknown (no device): lanes decide
  if (waves == 1) return 0.5;
  if (waves == 2) return 0.8;
  return 1.0;
}

static int waves_of(int S, int mode, bool multi, bool fused) {
  // queried once per instance and cached as value + 1 (0: not yet queried);
  // concurrent first calls compute the same value
  static std::atomic<int> cache[6][2][2][2];
  const int m = mode == SHIELD_MP_FP64 ? 1 : 0;
  int v = cache[S][m][multi][fused].load(std::memory_order_relaxed);
  if (v == 0) {
    int w = 0;
    switch (S) {
      case 1: w = m ? waves_per_simd<1, double>(multi, fused) : waves_per_simd<1, float>(multi, fused); break;
      case 2: w = m ? waves_per_simd<2, double>(multi, fused) : waves_per_simd<2, float>(multi, fused); break;
      case 3: w = m ? waves_per_simd<3, double>(multi, fused) : waves_per_simd<3, float>(multi, fused); break;
      case 4: w = m ? waves_per_simd<4, double>(multi, fused) : waves_per_simd<4, float>(multi, fused); break;
      case 5: w = m ? waves_per_simd<5, double>(multi, fused) : waves_per_simd<5, float>(multi, fused); break;
    }
    v = w + 1;
    cache[S][m][multi][fused].store(v, std::memory_order_relaxed);
  }
  return v - 1;
}

// Columns per wavefront P and slots S for nz levels: P = floor(64 S / nz)
// whole columns in 64 S lanes, scored by the busy-lane fraction times the
// occupancy weight of that instance (plain or fused).  SHIELD_MP_SLOTS=<S>
// forces a slot count (tuning).
LayoutChoice choose_layout(int64_t nz, int mode, int ntimes, bool fused) {
  LayoutChoice best{0, 0};
  if (nz < 1) return best;
  static const int forced = [] {
    const char* e = getenv("SHIELD_MP_SLOTS");
    const int v = e ? atoi(e) : 0;
    return (v >= 1 && v <= 5) ? v : 0;
  }();
  double best_score = -1.0;
  for (int S = 1; S <= 5; ++S) {
    if (forced && S != forced) continue;
    const int64_t P = (64 * S) / nz;
    if (P < 1) continue;
    const int64_t Pc = P > pmax_for(S) ? pmax_for(S) : P;
    const double util = double(Pc * nz) / double(64 * S);
    const double score = util * occupancy_weight(waves_of(S, mode, ntimes > 1, fused));
    if (score > best_score + 1e-9) {
      best_score = score;
      best = {S, (int)Pc};
    }
  }
  return best;
}

template <class R>
static hipError_t launch_r(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f,
                           const shield_update_fields& u, LayoutChoice lc, hipStream_t stream) {
  switch (lc.S) {
    case 1: return launch_slots<1, R>(sh, p, f, u, lc.P, stream);
    case 2: return launch_slots<2, R>(sh, p, f, u, lc.P, stream);
    case 3: return launch_slots<3, R>(sh, p, f, u, lc.P, stream);
    case 4: return launch_slots<4, R>(sh, p, f, u, lc.P, stream);
    case 5: return launch_slots<5, R>(sh, p, f, u, lc.P, stream);
    default: return hipErrorInvalidValue;
  }
}

// u.delp == NULL: the microphysics alone (the state update is skipped).
hipError_t launch_forward(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f,
                          const shield_update_fields& u, int mode, hipStream_t stream) {
  const LayoutChoice lc = choose_layout(sh.nk, mode, p.ntimes, u.delp != nullptr);
  if (lc.S == 0) return hipErrorInvalidValue;
  if (mode == SHIELD_MP_FP64) return launch_r<double>(sh, p, f, u, lc, stream);
  return launch_r<float>(sh, p, f, u, lc, stream);
}

}  // namespace shield
