// mp_common.h -- constants, coefficient block and precision-generic math for
// the GFDL microphysics kernels (gfx950).
//
// Precision model: the prognostic state (six water species, temperature, the
// cloud-fraction accumulator) is always FP64 in registers; every process rate
// is computed in the "rate type" R (double for SHIELD_MP_FP64, float for
// SHIELD_MP_MIXED) and applied to the FP64 state as a paired FP64 transfer, so
// total water stays conserved to FP64 rounding in both modes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/shield_mp.h"

// The constants, math and point-wise process functions are host-and-device
// so a host build can check the process code and its derivatives on the CPU.
#define SH_FN __host__ __device__ __forceinline__

namespace shield {

// ndsl.constants, GFS set
constexpr double GRAV = 9.80665;
constexpr double RGRAV = 1.0 / GRAV;
constexpr double RDGAS = 287.05;
constexpr double RVGAS = 461.50;
constexpr double HLV = 2.5e6;
constexpr double HLF = 3.3358e5;
constexpr double CP_AIR = 1004.6;
constexpr double TICE = 273.16;
constexpr double C_LIQ = 4185.5;
constexpr double C_ICE = 1972.0;
constexpr double E00 = 611.21;
constexpr double T_MIN = 178.0;
constexpr double CV_AIR = CP_AIR - RDGAS;
constexpr double CV_VAP = 3.0 * RVGAS;
constexpr double CP_VAP = 4.0 * RVGAS;
constexpr double DC_VAP = CP_VAP - C_LIQ;
constexpr double DC_ICE = C_LIQ - C_ICE;
constexpr double D2ICE = DC_VAP + DC_ICE;
constexpr double LV0 = HLV - DC_VAP * TICE;
constexpr double LI00 = HLF - DC_ICE * TICE;
constexpr double LI2 = LV0 + LI00;
constexpr double T_SAT_MIN = TICE - 160.0;
constexpr double LAT2 = (HLV + HLF) * (HLV + HLF);

// microphysics_funcs.py:7-42
constexpr double VCONS = 6.6280504;
constexpr double VCONG = 87.2382675;
constexpr double NORMS = 942477796.076938;
constexpr double NORMG = 5026548245.74367;
constexpr double VCONR = 2503.23638966667;
constexpr double NORMR = 25132741228.7183;
constexpr double THR = 1.0e-8;
constexpr double THI = 1.0e-8;
constexpr double THG = 1.0e-8;
constexpr double THS = 1.0e-8;
constexpr double AA = -4.14122e-5;
constexpr double BB = -0.00538922;
constexpr double CC = -0.0516344;
constexpr double DD_FS = 0.00216078;
constexpr double EE = 1.9714;
constexpr double VR_MIN = 1.0e-3;
constexpr double VF_MIN = 1.0e-5;
constexpr double P_MIN = 100.0;
constexpr double DT_FR = 8.0;
constexpr double SFCRHO = 1.2;
constexpr double RHOR = 1.0e3;
constexpr double QCMIN = 1.0e-12;
constexpr double QRMIN = 1.0e-8;
constexpr double QVMIN = 1.0e-20;

// ---------------------------------------------------------------- math
SH_FN double mexp(double x) { return exp(x); }
SH_FN double mlog(double x) { return log(x); }
SH_FN double msqrt(double x) { return sqrt(x); }
SH_FN double mpow(double x, double y) { return pow(x, y); }
SH_FN double mabs(double x) { return fabs(x); }
SH_FN float mexp(float x) { return __expf(x); }
SH_FN float mlog(float x) { return __logf(x); }
// v_sqrt_f32 (1 ulp); the FP32 rate laws need no correctly rounded root
SH_FN float msqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
SH_FN float mpow(float x, float y) { return __expf(y * __logf(x)); }
SH_FN float mabs(float x) { return fabsf(x); }
// Powers x^p (x > 0, p a literal).  pw_el is the reference's exp(p*log(x)),
// pw its x**p.  In FP64 they are exactly those expressions.  In FP32 the power
// is split as x^n * 2^((p-n) log2 x) with n the integer nearest p: rounding
// the exponent then costs |p-n| |log2 x| ulp instead of |p| |ln x| ulp (for
// x^(7/3) at x = 1e-4, 3 ulp instead of 21).
SH_FN float pw_split(float x, double p) {
  const double nd = __builtin_rint(p);
  const float r = float(p - nd);
  float y = __builtin_amdgcn_exp2f(r * __builtin_amdgcn_logf(x));
  const int n = int(nd);
  if (n >= 1) y *= x;
  if (n >= 2) y *= x;
  if (n >= 3) y *= x;
  return y;
}
template <class R>
SH_FN R pw_el(R x, double p);
template <>
SH_FN double pw_el<double>(double x, double p) { return exp(p * log(x)); }
template <>
SH_FN float pw_el<float>(float x, double p) { return pw_split(x, p); }
template <class R>
SH_FN R pw(R x, double p);
template <>
SH_FN double pw<double>(double x, double p) { return pow(x, p); }
template <>
SH_FN float pw<float>(float x, double p) { return pw_split(x, p); }

// a / b for the FP64 parts of the process code.  FP64 mode keeps the IEEE
// division of the reference's expression; mixed mode multiplies by a
// reciprocal from v_rcp_f64 refined by one Newton step, which the compiler
// shares between divisions by the same value (lhl / cvm and lhi / cvm, ...).
// Measured on gfx950 over b in [1e-3, 1e7]: v_rcp_f64 is within 4.6e-8
// relative, and one Newton step brings it to 2.2e-15 (two give the same).
// Host builds of the process code (the tests' checks) divide.
SH_FN double rcp64(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rcp(b);
  return fma(fma(-b, r, 1.0), r, r);
#else
  return 1.0 / b;
#endif
}
// The wide (state) type of rate type R: double for the forward kernels' float
// and double, the dual number itself for the adjoint's local Jacobians.
struct Dual;
template <class R>
struct Wide {
  using type = double;
};
template <>
struct Wide<Dual> {
  using type = Dual;
};
template <class R>
using wide_t = typename Wide<R>::type;

template <class R>
SH_FN wide_t<R> fdiv(wide_t<R> a, wide_t<R> b) {
  if constexpr (sizeof(R) == 4)
    return a * rcp64(b);
  else
    return a / b;
}
template <class R>
SH_FN wide_t<R> recip(wide_t<R> b) {
  if constexpr (sizeof(R) == 4)
    return rcp64(b);
  else
    return 1.0 / b;
}

// exp and log for the FP64 saturation functions of the mixed-precision kernel,
// whose arguments are moderate (the exponent of e_sat(T) has |x| < 50, the log
// argument T / tice is positive and normal) and which need ~1e-14, not the
// last ulp.
//  * exp: Cody-Waite reduction x = n ln2 + r, |r| <= ln2 / 2, with ln2 split
//    into a 32-bit head and its remainder so n ln2 is exact, then the Taylor
//    series of exp(r) to r^11 (truncation < 7e-15 relative); no overflow or
//    denormal handling (|x| < 700).
//  * log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with
//    s = (m - 1) / (m + 1), |s| <= 0.172, summed to s^21 (truncation < 3e-17);
//    x must be positive and normal.
SH_FN double exp_sat(double x) {
  constexpr double LOG2E = 1.4426950408889634, LN2_HI = 6.93147180369123816490e-01,
                   LN2_LO = 1.90821492927058770002e-10;
  const double n = __builtin_rint(x * LOG2E);
  const double r = fma(-n, LN2_LO, fma(-n, LN2_HI, x));
  double p = 1.0 / 39916800.0;  // 1/11!
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return __builtin_ldexp(p, int(n));
}
SH_FN double log_sat(double x) {
  constexpr double LN2 = 6.93147180559945286227e-01, RSQRT2 = 0.70710678118654752440;
  int e;
  double m = __builtin_frexp(x, &e);  // [0.5, 1)
  const bool lo = m < RSQRT2;
  m = lo ? m + m : m;
  e = lo ? e - 1 : e;
  const double s = (m - 1.0) * rcp64(m + 1.0);
  const double z = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, z, 1.0 / 19.0);
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  const double s2 = s + s;
  return fma(double(e), LN2, fma(s2 * z, p, s2));
}

template <class T>
SH_FN T mmin(T a, T b) { return a < b ? a : b; }
template <class T>
SH_FN T mmax(T a, T b) { return a > b ? a : b; }
// v_min / v_max for the floating-point cases: one instruction instead of a
// compare and a select per 32-bit half.  They differ from the selects only
// for NaN operands and in the sign of a zero result, which no result of the
// kernel depends on.
SH_FN float mmin(float a, float b) { return __builtin_fminf(a, b); }
SH_FN float mmax(float a, float b) { return __builtin_fmaxf(a, b); }
SH_FN double mmin(double a, double b) { return __builtin_fmin(a, b); }
SH_FN double mmax(double a, double b) { return __builtin_fmax(a, b); }
// microphysics_funcs.py:45-49
template <class T>
SH_FN T dim(T x, T y) {
  T diff = x - y;
  return diff > T(0) ? diff : T(0);
}

// most columns one wavefront of the forward kernel owns, per slot count (the
// size of its per-column LDS tables)
constexpr int pmax_for(int S) { return S >= 3 ? 16 : 64; }

// --------------------------------------------------- coefficient block
// Offsets of the coefficient arrays in Coef::tab: acco[3][4] (row-major),
// cssub[5], crevp[5], csmlt[5], cgmlt[5], cgfr[2].
enum { T_ACCO = 0, T_CSSUB = 12, T_CREVP = 17, T_CSMLT = 22, T_CGMLT = 27, T_CGFR = 32, T_N = 34 };

// The scalars the kernels read, in the rate type R.  Filled on the host from
// shield_mp_params (which is built by shield_mp_setup).
template <class R>
struct Coef {
  R c_air, c_vap, d0_vap, lv00, so3, log_10, tice, tice0, t_wfr, t_sub;
  R fac_rc, cracs, csacr, cgacr, cgacs, csacw, csaci, cgacw, cgaci, cracw, ces0;
  R tab[T_N];
  R dts, rdts, dt_rain, fac_i2s, fac_g2v, fac_v2g, fac_imlt, fac_l2v;
  R ql_mlt, qs_mlt, qi_gen, qi_lim, qi0_crt, qs0_crt, qc_crt;
  R vs_fac, vg_fac, vi_fac, vr_fac, vs_max, vg_max, vi_max, vr_max;
  int32_t const_vi, const_vs, const_vg, const_vr, fast_sat_adj, do_qa, rad_snow, rad_rain;
  int32_t use_ccn, pad_;
  // FP64 copies of the thresholds the state is tested against and of the
  // heat-capacity coefficients (latent heats and cvm are FP64 in both modes)
  double t_sub_d, t_wfr_d, qc_crt_d;
  double c_air_d, c_vap_d, d0_vap_d, lv00_d;
  double cgfr1_d;  // the Bigg coefficient A' of the rain freezing, FP64
};

// The kernels keep their copy of Coef::tab in LDS (filled at their start by
// fill_coef_tab): with the whole block in scalar registers the compiler
// spills them to VGPR lanes (or, for the FP64 block, to scratch memory) and
// reloads them inside the process code.  Each translation unit gets its own
// (per-kernel) copy; a kernel allocates only the one its precision reads.
static __shared__ float coef_tab_lds[T_N];
static __shared__ double coef_tab_lds64[T_N];

// Coefficient-array entry i (a constant after inlining): in device code from
// the LDS copy, in host builds of the process code (the point-wise test
// harness) from the block itself.
template <class R>
SH_FN R ck(const Coef<R>& c, int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(R) == 4)
    return coef_tab_lds[i];
  else if constexpr (sizeof(R) == 8)
    return coef_tab_lds64[i];
  else
    return c.tab[i];
#else
  return c.tab[i];
#endif
}
// Copies Coef::tab into the LDS table of R; the caller synchronises the
// workgroup before the first ck().
template <class R>
__device__ __forceinline__ void fill_coef_tab(const Coef<R>& c, int lane) {
  if (lane < T_N) {
    if constexpr (sizeof(R) == 4)
      coef_tab_lds[lane] = c.tab[lane];
    else if constexpr (sizeof(R) == 8)
      coef_tab_lds64[lane] = c.tab[lane];
  }
}

template <class R>
inline Coef<R> make_coef(const shield_mp_params& p) {
  Coef<R> c;
  const shield_mp_namelist& n = p.nml;
  c.c_air = R(p.c_air); c.c_vap = R(p.c_vap); c.d0_vap = R(p.d0_vap); c.lv00 = R(p.lv00);
  c.so3 = R(p.so3); c.log_10 = R(p.log_10); c.tice = R(n.tice); c.tice0 = R(p.tice0);
  c.t_wfr = R(p.t_wfr); c.t_sub = R(n.t_sub);
  c.fac_rc = R(p.fac_rc); c.cracs = R(p.cracs); c.csacr = R(p.csacr); c.cgacr = R(p.cgacr);
  c.cgacs = R(p.cgacs); c.csacw = R(p.csacw); c.csaci = R(p.csaci); c.cgacw = R(p.cgacw);
  c.cgaci = R(p.cgaci); c.cracw = R(p.cracw); c.ces0 = R(p.ces0);
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 4; ++k) c.tab[T_ACCO + 4 * i + k] = R(p.acco[i][k]);
  for (int i = 0; i < 5; ++i) {
    c.tab[T_CSSUB + i] = R(p.cssub[i]); c.tab[T_CREVP + i] = R(p.crevp[i]);
    c.tab[T_CSMLT + i] = R(p.csmlt[i]); c.tab[T_CGMLT + i] = R(p.cgmlt[i]);
  }
  c.tab[T_CGFR] = R(p.cgfr[0]); c.tab[T_CGFR + 1] = R(p.cgfr[1]);
  c.dts = R(p.dts); c.rdts = R(p.rdts); c.dt_rain = R(p.dt_rain);
  c.fac_i2s = R(p.fac_i2s); c.fac_g2v = R(p.fac_g2v); c.fac_v2g = R(p.fac_v2g);
  c.fac_imlt = R(p.fac_imlt); c.fac_l2v = R(p.fac_l2v);
  c.ql_mlt = R(n.ql_mlt); c.qs_mlt = R(n.qs_mlt); c.qi_gen = R(n.qi_gen); c.qi_lim = R(n.qi_lim);
  c.qi0_crt = R(n.qi0_crt); c.qs0_crt = R(n.qs0_crt); c.qc_crt = R(n.qc_crt);
  c.vs_fac = R(n.vs_fac); c.vg_fac = R(n.vg_fac); c.vi_fac = R(n.vi_fac); c.vr_fac = R(n.vr_fac);
  c.vs_max = R(n.vs_max); c.vg_max = R(n.vg_max); c.vi_max = R(n.vi_max); c.vr_max = R(n.vr_max);
  c.const_vi = n.const_vi; c.const_vs = n.const_vs; c.const_vg = n.const_vg; c.const_vr = n.const_vr;
  c.fast_sat_adj = n.fast_sat_adj; c.do_qa = n.do_qa; c.rad_snow = n.rad_snow; c.rad_rain = n.rad_rain;
  c.use_ccn = p.use_ccn; c.pad_ = 0;
  c.t_sub_d = n.t_sub; c.t_wfr_d = p.t_wfr; c.qc_crt_d = n.qc_crt;
  c.c_air_d = p.c_air; c.c_vap_d = p.c_vap; c.d0_vap_d = p.d0_vap; c.lv00_d = p.lv00;
  c.cgfr1_d = p.cgfr[1];
  return c;
}

// fill_gfs_delp (update_atmos_state.py:21-37) on one column of nk layers in
// memory (k = 0 the top): q below q_min is filled from the layer above it,
// then floored, filled downward from a negative layer and clamped at zero.
// The running value is carried in registers; the reference's three downward
// passes run as one (the borrow takes the layer above after its own borrow,
// before the clamp), with every element getting the reference's operations
// in the reference's order.
SH_FN void fill_column(double* qc, const double* dp, int nk, double q_min) {
  // upward: fill from the layer below (its value as just updated)
  double qn = qc[nk - 1], dn = dp[nk - 1];
  for (int k = nk - 2; k >= 0; --k) {
    double qk = qc[k];
    const double dk = dp[k];
    if (qn < q_min) qk = qk + (qn - q_min) * dn / dk;
    qc[k] = qk;
    qn = qk;
    dn = dk;
  }
  // downward: the floor at q_min (k >= 1), the borrow from a negative layer
  // above, the clamp at zero
  double qp = qc[0], dpv = dp[0];
  qc[0] = qp < 0.0 ? 0.0 : qp;
  for (int k = 1; k < nk; ++k) {
    double qk = qc[k];
    if (qk < q_min) qk = q_min;
    const double dk = dp[k];
    if (qp < 0.0) qk = qk + qp * (dpv) / (dk);
    qc[k] = qk < 0.0 ? 0.0 : qk;
    qp = qk;
    dpv = dk;
  }
}

}  // namespace shield
