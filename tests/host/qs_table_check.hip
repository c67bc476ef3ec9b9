// Host build of the mixed kernel's saturation functions (mp_process.h) for
// tests/test_qs_table.py: the table-based qs_fast / iqs_fast / qs_pick beside
// the reference's FP64 closed forms wqs2 / iqs2 at the same (T, den).
#include "../../pyshield_amd/csrc/mp_process.h"

using namespace shield;

// out[5 i + 0..4]: table water q, closed-form water q, table ice q, closed-form
// ice q, qs_pick(water=false); dq[4 i + 0..3]: table and closed-form dq/dT,
// water then ice
extern "C" void qs_check(int n, const double* t, const double* den, double* out, double* dq) {
  for (int i = 0; i < n; ++i) {
    double g0, g1, g2, g3;
    out[5 * i + 0] = qs_fast(t[i], 0, den[i], g0);
    out[5 * i + 1] = wqs2<double>(t[i], den[i], g1);
    out[5 * i + 2] = iqs_fast(t[i], den[i], g2);
    out[5 * i + 3] = iqs2<double>(t[i], den[i], g3);
    out[5 * i + 4] = qs_pick(t[i], den[i], false);
    dq[4 * i + 0] = g0;
    dq[4 * i + 1] = g1;
    dq[4 * i + 2] = g2;
    dq[4 * i + 3] = g3;
  }
}
