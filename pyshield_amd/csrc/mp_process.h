// mp_process.h -- point-wise GFDL microphysics processes for one cell.
//
// Each function follows the pySHiELD process library
// (pySHiELD/functions/microphysics_funcs.py) statement by statement.  The
// cell state `Cell` is FP64; process rates are computed in the rate type R and
// applied as FP64 transfers (s.q_a -= x; s.q_b += x).  With R = double every
// expression has the reference's operation order.  With R = float (mixed
// mode) the rate formulas (collection, evaporation, deposition, melting,
// fall-speed and autoconversion laws) run in FP32, and what decides how much
// of a species survives stays FP64:
//  * saturation mixing ratios and their gradients, from the FP64 temperature
//    and density, and every difference between a saturation value and the
//    state (qsat - qv, ...); FP32 would put ~1e-6 of qsat (1e-8 kg/kg) into
//    every saturation adjustment;
//  * latent heats, the moist heat capacity cvm and their ratios, so a
//    transfer limited by the heat that brings a cell to tice or t_wfr is
//    accurate to FP64;
//  * temperature differences (tz - tice, ...) and the limiters that scale a
//    set of sinks to what is available;
//  * where the reference's arithmetic bounds a sink by the species it drains
//    (x = f / (1 + f) q, min(q - qc, ...), x <= qv - qsat, ...), an FP32
//    sink that rounding put just above the FP64 state is bounded to it
//    (`cap`), so FP32 rounding drives no species below zero; sinks the
//    reference leaves unbounded (the subgrid autoconversion) and states the
//    reference itself drove negative are left as they are.
#pragma once

#include "mp_dual.h"
#include "qs_table.h"

namespace shield {

template <class W>
struct CellT {
  W qv, ql, qr, qi, qs, qg, tz, qa;
};
using Cell = CellT<double>;
template <class R>
using CellR = CellT<wide_t<R>>;

template <class R>
struct is_f32 {
  static constexpr bool value = false;
};
template <>
struct is_f32<float> {
  static constexpr bool value = true;
};

// Bound a transfer x out of species q at a site where the reference's
// arithmetic guarantees x <= q (mixed precision only; identity in FP64).  It
// binds only where FP32 rounding put x at most 1e-5 q above a positive q: a
// reference state can hold a negative species (an autoconversion larger than
// the cloud water leaves ql < 0, and qv follows), and there the reference's
// transfer (0, or f q / (1 + f)) stays as it is.
template <class R>
SH_FN wide_t<R> cap(wide_t<R> x, wide_t<R> q) {
  if constexpr (is_f32<R>::value) return (x > q && x <= q * (1.0 + 1.0e-5)) ? q : x;
  return x;
}

// The part f / (1 + f) q of q that a backward-Euler collection step takes
// (f the step's rate factor).  FP64 keeps the reference's expression.  In
// mixed precision the residue q / (1 + f) is formed in FP64 from the FP32 f
// and the transfer is q minus it: the residue then carries the relative
// error of 1 + f, at most that of f, where f / (1 + f) q rounded in FP32
// would leave it that error times (1 + f) when f is large.
template <class R>
SH_FN wide_t<R> implicit_take(R f, wide_t<R> q) {
  if constexpr (is_f32<R>::value)
    return q - fdiv<R>(q, 1.0 + double(f));
  else
    return f / (R(1.0) + f) * q;
}

// exp(a t) for an FP64 temperature difference t in a rate law of type R (a_d
// the FP64 value of the coefficient a).  FP64 and the dual numbers keep
// exp(a * R(t)).  FP32 rate laws reduce the
// argument exactly in FP64 (x = k ln2 + r, |r| <= ln2 / 2) and take the FP32
// exponential of r only: rounding x to FP32 first would cost |x| 2^-24
// relative, 1.7e-6 (14 ulp) for the Bigg freezing's exp(0.66 tc) at tc = 43 K.
template <class R>
SH_FN R exp_lin(R a, double a_d, wide_t<R> t) {
  if constexpr (is_f32<R>::value) {
    const double x = a_d * t;
    const double k = __builtin_rint(x * 1.4426950408889634);
    const float r = float(fma(-k, 0.6931471805599453, x));
    return ldexpf(mexp(r), int(k));
  } else {
    return mexp(a * R(t));
  }
}

// moist heat capacity (FP64 in both modes)
#define SH_CVM(c, s, q_liq, q_sol) ((c).c_air_d + (s).qv * (c).c_vap_d + (q_liq) * C_LIQ + (q_sol) * C_ICE)

// ------------------------------------------------ saturation (f:53-145)
template <class R>
SH_FN R wqs1(R ta, R den) {
  return (R(E00) * mexp((R(DC_VAP) * mlog(ta / R(TICE)) + R(LV0) * (ta - R(TICE)) / (ta * R(TICE))) /
                        R(RVGAS))) /
         (R(RVGAS) * ta * den);
}
template <class R>
SH_FN R wqs2(R ta, R den, R& dqdt) {
  R tmp = wqs1(ta, den);
  dqdt = tmp * (R(DC_VAP) + R(LV0) / ta) / (R(RVGAS) * ta);
  return tmp;
}
template <class R>
SH_FN R iqs1(R ta, R den) {
  R tmp;
  if (ta < R(TICE)) {
    if (ta >= R(T_SAT_MIN)) {
      tmp = (R(E00) * mexp((R(D2ICE) * mlog(ta / R(TICE)) + R(LI2) * (ta - R(TICE)) / (ta * R(TICE))) /
                           R(RVGAS))) /
            (R(RVGAS) * ta * den);
    } else {
      tmp = (R(E00) * mexp((R(D2ICE) * mlog(R(1.0) - R(160.0) / R(TICE)) -
                            R(LI2) * R(160.0) / (R(T_SAT_MIN) * R(TICE))) /
                           R(RVGAS))) /
            (R(RVGAS) * R(T_SAT_MIN) * den);
    }
  } else {
    if (ta <= R(TICE) + R(102.0))
      tmp = wqs1(ta, den);
    else
      tmp = wqs1(R(TICE) + R(102.0), den);
  }
  return tmp;
}
template <class R>
SH_FN R iqs2(R ta, R den, R& dqdt) {
  R tmp = iqs1(ta, den);
  if (ta < R(TICE)) {
    if (ta >= R(T_SAT_MIN))
      dqdt = tmp * (R(D2ICE) + R(LI2) / ta) / (R(RVGAS) * ta);
    else
      dqdt = tmp * (R(D2ICE) + R(LI2) / R(T_SAT_MIN)) / (R(RVGAS) * R(T_SAT_MIN));
  } else {
    if (ta <= R(TICE) + R(102.0))
      dqdt = tmp * (R(DC_VAP) + R(LV0) / ta) / (R(RVGAS) * ta);
    else
      dqdt = tmp * (R(DC_VAP) + R(LV0) / (R(TICE) + R(102.0))) / (R(RVGAS) * (R(TICE) + R(102.0)));
  }
  return tmp;
}

// Saturation mixing ratio and gradient in the mixed-precision kernel: the
// closed forms above in FP64, with e_sat(T) read from the FP64-built table of
// qs_table.h (a degree-7 polynomial on each of 128 intervals of u = 1/T,
// < 2e-12 relative, 16 KB; four 16-byte loads and seven multiply-adds in place
// of an exp and a log; a degree-11 table of 32 intervals, 6 KB, measured
// slower), the divisions by constants folded into products and the
// others replaced by reciprocals.  Below T_SAT_MIN and above TICE + 102 the
// reference's ice branches are the same expression at the clamped
// temperature.  Over water the table spans 100-1000 K (the reference's own
// cells reach 460 K where its autoconversion overdraws cloud water) and a
// temperature outside is read at the nearest end: below 100 K the water
// saturation ratio is under 1e-15 kg/kg either way.  (A closed-form fallback
// for the cells outside, inlined at every call site, made the kernel 4%
// slower through its code size.)
//
// e_sat at u = rt = 1/t from table `ice` (0 water: d = DC_VAP, l = LV0;
// 1 ice: d = D2ICE, l = LI2)
SH_FN double esat_table(double rt, int ice) {
  double x = (rt - QS_TAB_U0[ice]) * QS_TAB_RN[ice];
  x = x < 0.0 ? 0.0 : x > double(QS_TAB_N) ? double(QS_TAB_N) : x;
  const int i = x < double(QS_TAB_N - 1) ? int(x) : QS_TAB_N - 1;
  const double s = 2.0 * (x - double(i)) - 1.0;
  const double* c = QS_TAB[ice][i];
  double p = c[QS_TAB_DEG];
#pragma unroll
  for (int j = QS_TAB_DEG - 1; j >= 0; --j) p = fma(p, s, c[j]);
  return p;
}
SH_FN double qs_fast(double t, int ice, double den, double& dqdt) {
  constexpr double rrv = 1.0 / RVGAS;
  const double d = ice ? D2ICE : DC_VAP, l = ice ? LI2 : LV0;
  const double rt = rcp64(t);
  const double e = esat_table(rt, ice);
  const double q = e * rrv * rt * rcp64(den);
  dqdt = q * (d + l * rt) * rrv * rt;
  return q;
}
SH_FN double iqs_fast(double ta, double den, double& dqdt) {
  const bool ice = ta < TICE;
  const double t = ice ? (ta < T_SAT_MIN ? T_SAT_MIN : ta) : (ta > TICE + 102.0 ? TICE + 102.0 : ta);
  return qs_fast(t, ice ? 1 : 0, den, dqdt);
}
// qs_fast(ta, water table, ...) where `water`, else iqs_fast(ta, ...), as one
// evaluation with the table and the clamp chosen per cell
SH_FN double qs_pick(double ta, double den, bool water) {
  const bool ice = !water && ta < TICE;
  const double t = water ? ta : ice ? (ta < T_SAT_MIN ? T_SAT_MIN : ta) : (ta > TICE + 102.0 ? TICE + 102.0 : ta);
  double g;
  return qs_fast(t, ice ? 1 : 0, den, g);
}
// the reference's forms in FP64 mode (and for the adjoint's dual numbers),
// the folded ones in mixed mode
template <class R>
SH_FN wide_t<R> qs_w(wide_t<R> ta, wide_t<R> den, wide_t<R>& dqdt) {
  if constexpr (is_f32<R>::value) return qs_fast(ta, 0, den, dqdt);
  else return wqs2<wide_t<R>>(ta, den, dqdt);
}
template <class R>
SH_FN wide_t<R> qs_w1(wide_t<R> ta, wide_t<R> den) {
  wide_t<R> g;
  if constexpr (is_f32<R>::value) return qs_fast(ta, 0, den, g);
  else return wqs1<wide_t<R>>(ta, den);
}
template <class R>
SH_FN wide_t<R> qs_i(wide_t<R> ta, wide_t<R> den, wide_t<R>& dqdt) {
  if constexpr (is_f32<R>::value) return iqs_fast(ta, den, dqdt);
  else return iqs2<wide_t<R>>(ta, den, dqdt);
}
template <class R>
SH_FN wide_t<R> qs_i1(wide_t<R> ta, wide_t<R> den) {
  wide_t<R> g;
  if constexpr (is_f32<R>::value) return iqs_fast(ta, den, g);
  else return iqs1<wide_t<R>>(ta, den);
}

// ------------------------------------------------ f:149-179
template <class R>
SH_FN R acr3d(R v1, R v2, R q1, R q2, R c, R cac_ik, R cac_i1k, R cac_i2k, R rho) {
  R t1 = msqrt(q1 * rho);
  R s1 = msqrt(q2 * rho);
  R s2 = msqrt(s1);
  return c * mabs(v1 - v2) * q1 * s2 * (cac_ik * t1 + cac_i1k * msqrt(t1) * s2 + cac_i2k * s1);
}
template <class R>
SH_FN R smlt(const Coef<R>& c, R tc, R dqs, R qsrho, R psacw, R psacr, R rho, R rhofac) {
  const int k = T_CSMLT;
  return (ck(c, k) * tc / rho - ck(c, k + 1) * dqs) *
             (ck(c, k + 2) * msqrt(qsrho) + ck(c, k + 3) * pw<R>(qsrho, 0.65625) * msqrt(rhofac)) +
         ck(c, k + 4) * tc * (psacw + psacr);
}
template <class R>
SH_FN R gmlt(const Coef<R>& c, R tc, R dqs, R qgrho, R pgacw, R pgacr, R rho) {
  const int k = T_CGMLT;
  return (ck(c, k) * tc / rho - ck(c, k + 1) * dqs) *
             (ck(c, k + 2) * msqrt(qgrho) + ck(c, k + 3) * pw<R>(qgrho, 0.6875) / pw<R>(rho, 0.25)) +
         ck(c, k + 4) * tc * (pgacw + pgacr);
}

// ------------------------------------------------ f:183-265
template <class R>
SH_FN void revap_racc(const Coef<R>& c, double t_wfr, R dt, R h_var, CellR<R>& s,
                                           wide_t<R> den_d, R denfac) {
  using W = wide_t<R>;
  if ((s.tz > t_wfr) && (s.qr > QRMIN)) {
    const R den = R(den_d);
    W lhl = c.lv00_d + c.d0_vap_d * s.tz;
    W q_liq = s.ql + s.qr;
    W q_sol = s.qi + s.qs + s.qg;
    W cvm = SH_CVM(c, s, q_liq, q_sol);
    W lcpk = fdiv<R>(lhl, cvm);
    const W tin = s.tz - W(lcpk * R(s.ql));
    const W qpz = s.qv + s.ql;
    W dqsdt;
    const W qsat = qs_w<R>(tin, den_d, dqsdt);
    R dqh = mmax(R(s.ql), h_var * mmax(R(qpz), R(QCMIN)));
    dqh = mmin(dqh, R(0.2) * R(qpz));
    const W dqv = qsat - s.qv;
    const W q_minus = qpz - W(dqh);
    const W q_plus = qpz + W(dqh);
    if ((dqv > QVMIN) && (qsat > q_minus)) {
      R dq;
      if (qsat > q_plus)
        dq = R(qsat - qpz);
      else
        dq = R(0.25) * R((q_minus - qsat) * (q_minus - qsat)) / dqh;
      R qden = R(s.qr) * den;
      R t2 = R(tin) * R(tin);
      R evap = ck(c, T_CREVP + 0) * t2 * dq * (ck(c, T_CREVP + 1) * msqrt(qden) + ck(c, T_CREVP + 2) * pw_el<R>(qden, 0.725)) /
               (ck(c, T_CREVP + 3) * t2 + ck(c, T_CREVP + 4) * R(qsat) * den);
      W ev = mmin(s.qr, mmin(W(dt * evap), fdiv<R>(dqv, 1.0 + lcpk * dqsdt)));
      s.qr = s.qr - ev;
      s.qv = s.qv + ev;
      q_liq = q_liq - ev;
      cvm = SH_CVM(c, s, q_liq, q_sol);
      s.tz = s.tz - fdiv<R>(ev * lhl, cvm);
    }
    if ((s.qr > QRMIN) && (s.ql > 1.0e-6) && (qsat < q_minus)) {
      R sink = dt * denfac * c.cracw * pw_el<R>(R(s.qr) * den, 0.95);
      W sk = implicit_take<R>(sink, s.ql);
      s.ql = s.ql - sk;
      s.qr = s.qr + sk;
    }
  }
}

// ------------------------------------------------ f:269-341
template <class R>
SH_FN void fall_speed(const Coef<R>& c, R qg, R qi, R qs, R tc, R den, R& vtg, R& vti,
                                           R& vts) {
  R rhof = msqrt(mmin(R(10.0), R(SFCRHO) / den));
  if (c.const_vi) {
    vti = c.vi_fac;
  } else {
    R vi0 = R(0.01) * c.vi_fac;
    if (qi < R(THI)) {
      vti = R(VF_MIN);
    } else {
      R v = (R(3.0) + mlog(qi * den) / c.log_10) * (tc * (R(AA) * tc + R(BB)) + R(CC)) + R(DD_FS) * tc + R(EE);
      v = vi0 * mexp(c.log_10 * v) * R(0.8);
      vti = mmin(c.vi_max, mmax(R(VF_MIN), v));
    }
  }
  if (c.const_vs) {
    vts = c.vs_fac;
  } else if (qs < R(THS)) {
    vts = R(VF_MIN);
  } else {
    R v = c.vs_fac * R(VCONS) * rhof * pw_el<R>(qs * den / R(NORMS), 0.0625);
    vts = mmin(c.vs_max, mmax(R(VF_MIN), v));
  }
  if (c.const_vg) {
    vtg = c.vg_fac;
  } else if (qg < R(THG)) {
    vtg = R(VF_MIN);
  } else {
    R v = c.vg_fac * R(VCONG) * rhof * msqrt(msqrt(msqrt(qg * den / R(NORMG))));
    vtg = mmin(c.vg_max, mmax(R(VF_MIN), v));
  }
}

// f:344-372 (the r1 half is handled by the caller)
template <class R>
SH_FN R rain_fspeed(const Coef<R>& c, bool no_fall, R qrz, R den) {
  if (no_fall) return R(VF_MIN);
  if (c.const_vr) return c.vr_fac;
  R qden = qrz * den;
  if (qrz < R(THR)) return R(VR_MIN);
  R v = c.vr_fac * R(VCONR) * msqrt(mmin(R(10.0), R(SFCRHO) / den)) * pw_el<R>(qden / R(NORMR), 0.2);
  return mmin(c.vr_max, mmax(R(VR_MIN), v));
}

// ------------------------------------------------ f:375-428
template <class R>
SH_FN void autoconv_no_subgrid_var(const Coef<R>& c, double t_wfr, CellR<R>& s, R den, R ccn,
                                                        R c_praut) {
  using W = wide_t<R>;
  R qc0 = c.fac_rc * ccn;
  if (s.tz > t_wfr) {
    R qc = c.use_ccn ? qc0 : qc0 / den;
    R dq = R(s.ql - W(qc));
    if (dq > R(0.0)) {
      W sink = mmin(W(dq), W(c.dt_rain * c_praut * den * pw_el<R>(R(s.ql), 7.0 / 3.0)));
      sink = cap<R>(sink, s.ql);
      s.ql = s.ql - sink;
      s.qr = s.qr + sink;
    }
  }
}
template <class R>
SH_FN void autoconv_subgrid_var(const Coef<R>& c, double t_wfr, CellR<R>& s, R den, R ccn,
                                                     R c_praut, R dl) {
  using W = wide_t<R>;
  R qc0 = c.fac_rc * ccn;
  if (s.tz > t_wfr + DT_FR) {
    dl = mmin(mmax(R(1.0e-6), dl), R(0.5) * R(s.ql));
    R qc = c.use_ccn ? qc0 : qc0 / den;
    R dq = R(0.5 * (s.ql + W(dl) - W(qc)));
    if (dq > R(0.0)) {
      // unbounded by ql, as in the reference: a cell with ql ~ 1e-2 and a
      // long rain step converts more than it holds and leaves ql < 0
      R sink = mmin(R(1.0), dq / dl) * c.dt_rain * c_praut * den * pw_el<R>(R(s.ql), 7.0 / 3.0);
      const W sk = W(sink);
      s.ql = s.ql - sk;
      s.qr = s.qr + sk;
    }
  }
}

// ------------------------------------------------ f:431-850
template <class R>
SH_FN void subgrid_z_proc(const Coef<R>& c, double tice_d, CellR<R>& s, R h_var, double rh_adj,
                                               double rh_rain, wide_t<R> den_d, R denfac, wide_t<R> p1) {
  using W = wide_t<R>;
  const R tice = c.tice, t_wfr = c.t_wfr, dts = c.dts;
  const R den = R(den_d);
  W lhl = c.lv00_d + c.d0_vap_d * s.tz;
  W lhi = LI00 + DC_ICE * s.tz;
  W q_liq = s.ql + s.qr;
  W q_sol = s.qi + s.qs + s.qg;
  W cvm = SH_CVM(c, s, q_liq, q_sol);
  W lcpk = fdiv<R>(lhl, cvm);
  W icpk = fdiv<R>(lhi, cvm);
  W tcpk = lcpk + icpk;
  // the share of the heat of fusion in the adjustment, FP64 in both modes
  W tcp3 = lcpk + icpk * mmin(W(1.0), fdiv<R>(dim(tice_d, s.tz), tice_d - c.t_wfr_d));
  if (!(p1 >= P_MIN)) return;
  if (s.tz < T_MIN) {
    W sink = dim(1.0e-7, s.qv);
    s.qv = s.qv - sink;
    s.qi = s.qi + sink;
    q_sol = q_sol + sink;
    cvm = SH_CVM(c, s, q_liq, q_sol);
    s.tz = s.tz + fdiv<R>(sink * (lhl + lhi), cvm);
    if (!c.do_qa) s.qa = s.qa + 1.0;
    return;
  }
  lhl = c.lv00_d + c.d0_vap_d * s.tz;
  lhi = LI00 + DC_ICE * s.tz;
  lcpk = fdiv<R>(lhl, cvm);
  icpk = fdiv<R>(lhi, cvm);
  tcpk = lcpk + icpk;
  tcp3 = lcpk + icpk * mmin(W(1.0), fdiv<R>(dim(tice_d, s.tz), tice_d - c.t_wfr_d));

  // instant evaporation / sublimation of all clouds when rh < rh_adj
  const W qpz = s.qv + s.ql + s.qi;
  const W dtin = fdiv<R>(lhl * (s.ql + s.qi) + lhi * s.qi,
                             c.c_air_d + qpz * c.c_vap_d + s.qr * C_LIQ + (s.qs + s.qg) * C_ICE);
  const W tin = s.tz - dtin;
  const W t_sub6 = c.t_sub_d + 6.0;
  W rh = 0.0;
  if (tin > t_sub6) {
    rh = fdiv<R>(qpz, qs_i1<R>(tin, den_d));
    if (rh < rh_adj) {
      // all vapour, liquid and ice of the cell collapse into vapour
      s.tz = tin;
      s.qv = qpz;
      s.ql = 0.0;
      s.qi = 0.0;
    }
  }
  if (!(((tin > t_sub6) && (rh >= rh_adj)) || (tin <= t_sub6))) return;

  // cloud water <-> vapour adjustment
  W dwsdt;
  const W qsw = qs_w<R>(s.tz, den_d, dwsdt);
  const W dq0 = qsw - s.qv;
  W evap;
  if (dq0 > 0.0) {
    // FP64: when this drains nearly all cloud water the residue inherits the
    // relative error of the factor times (cloud water / residue)
    const W factor = mmin(W(1.0), W(c.fac_l2v) * fdiv<R>(10.0 * dq0, qsw));
    evap = mmin(s.ql, fdiv<R>(factor * dq0, 1.0 + tcp3 * dwsdt));
  } else {
    evap = fdiv<R>(dq0, 1.0 + tcp3 * dwsdt);
  }
  s.qv = s.qv + evap;
  s.ql = s.ql - evap;
  q_liq = q_liq - evap;
  cvm = SH_CVM(c, s, q_liq, q_sol);
  s.tz = s.tz - fdiv<R>(evap * lhl, cvm);
  lhi = LI00 + DC_ICE * s.tz;
  icpk = fdiv<R>(lhi, cvm);

  // complete freezing below t_wfr
  const W dtmp = c.t_wfr_d - s.tz;
  if ((dtmp > 0.0) && (s.ql > QCMIN)) {
    W sink = mmin(s.ql, mmin(W(R(s.ql) * R(dtmp) * R(0.125)), fdiv<R>(dtmp, icpk)));
    s.ql = s.ql - sink;
    s.qi = s.qi + sink;
    q_liq = q_liq - sink;
    q_sol = q_sol + sink;
    cvm = SH_CVM(c, s, q_liq, q_sol);
    s.tz = s.tz + fdiv<R>(sink * lhi, cvm);
  }
  lhi = LI00 + DC_ICE * s.tz;
  icpk = fdiv<R>(lhi, cvm);

  // Bigg mechanism
  R dt_pisub;
  if (c.fast_sat_adj) {
    dt_pisub = R(0.5) * dts;
  } else {
    dt_pisub = dts;
    const W tcd = tice_d - s.tz;
    const R tc = R(tcd);
    if ((s.ql > QRMIN) && (tcd > 0.0)) {
      R sink = R(3.3333e-10) * dts * (exp_lin<R>(R(0.66), 0.66, tcd) - R(1.0)) * den * R(s.ql) * R(s.ql);
      W sk = mmin(s.ql, mmin(fdiv<R>(tcd, icpk), W(sink)));
      s.ql = s.ql - sk;
      s.qi = s.qi + sk;
      q_liq = q_liq - sk;
      q_sol = q_sol + sk;
      cvm = SH_CVM(c, s, q_liq, q_sol);
      s.tz = s.tz + fdiv<R>(sk * lhi, cvm);
    }
  }
  lhl = c.lv00_d + c.d0_vap_d * s.tz;
  lhi = LI00 + DC_ICE * s.tz;
  lcpk = fdiv<R>(lhl, cvm);
  icpk = fdiv<R>(lhi, cvm);
  tcpk = lcpk + icpk;

  // sublimation / deposition of cloud ice
  if (s.tz < tice_d) {
    W dqsdt;
    const W qsi = qs_i<R>(s.tz, den_d, dqsdt);
    const W dq = s.qv - qsi;
    const W sink = fdiv<R>(dq, 1.0 + tcpk * dqsdt);
    R pidep;
    if (s.qi > QRMIN) {
      pidep = dt_pisub * R(dq) * R(349138.78) * pw_el<R>(R(s.qi) * den, 0.875) /
              (R(qsi) * den * R(LAT2) / (R(0.0243) * R(RVGAS) * (R(s.tz) * R(s.tz))) + R(4.42478e4));
    } else {
      pidep = R(0.0);
    }
    W sk;
    if (dq > 0.0) {
      const W tmpd = tice_d - s.tz;
      const R tmp = R(tmpd);
      R qi_crt = c.qi_gen * mmin(c.qi_lim, R(0.1) * tmp) / den;
      sk = mmin(sink, mmin(W(mmax(R(W(qi_crt) - s.qi), pidep)), fdiv<R>(tmpd, tcpk)));
      sk = cap<R>(sk, s.qv);
    } else {
      pidep = pidep * mmin(R(1.0), R(dim(s.tz, c.t_sub_d)) * R(0.2));
      sk = mmax(mmax(W(pidep), sink), -s.qi);
    }
    s.qv = s.qv - sk;
    s.qi = s.qi + sk;
    q_sol = q_sol + sk;
    cvm = SH_CVM(c, s, q_liq, q_sol);
    s.tz = s.tz + fdiv<R>(sk * (lhl + lhi), cvm);
  }
  lhl = c.lv00_d + c.d0_vap_d * s.tz;
  lhi = LI00 + DC_ICE * s.tz;
  lcpk = fdiv<R>(lhl, cvm);
  icpk = fdiv<R>(lhi, cvm);
  tcpk = lcpk + icpk;

  // sublimation / deposition of snow
  if (s.qs > QRMIN) {
    W dqsdt;
    const W qsi = qs_i<R>(s.tz, den_d, dqsdt);
    R qden = R(s.qs) * den;
    R tmp = pw_el<R>(qden, 0.65625);
    R tsq = R(s.tz) * R(s.tz);
    const W dsat_d = qsi - s.qv;
    const R dsat = R(dsat_d);
    const W dq = fdiv<R>(dsat_d, 1.0 + tcpk * dqsdt);
    R pssub = ck(c, T_CSSUB + 0) * tsq * (ck(c, T_CSSUB + 1) * msqrt(qden) + ck(c, T_CSSUB + 2) * tmp * msqrt(denfac)) /
              (ck(c, T_CSSUB + 3) * tsq + ck(c, T_CSSUB + 4) * R(qsi) * den);
    pssub = dsat * dts * pssub;
    W ps;
    if (pssub > R(0.0)) {
      ps = mmin(W(pssub * mmin(R(1.0), R(dim(s.tz, c.t_sub_d)) * R(0.2))), s.qs);
    } else {
      if (s.tz > tice_d)
        ps = 0.0;
      else
        ps = cap<R>(-mmax(W(pssub), mmax(dq, fdiv<R>(s.tz - tice_d, tcpk))), s.qv) * -1.0;
    }
    s.qs = s.qs - ps;
    s.qv = s.qv + ps;
    q_sol = q_sol - ps;
    cvm = SH_CVM(c, s, q_liq, q_sol);
    s.tz = s.tz - fdiv<R>(ps * (lhl + lhi), cvm);
  }
  lhl = c.lv00_d + c.d0_vap_d * s.tz;
  lhi = LI00 + DC_ICE * s.tz;
  lcpk = fdiv<R>(lhl, cvm);
  icpk = fdiv<R>(lhi, cvm);
  tcpk = lcpk + icpk;

  // two-way graupel sublimation / deposition
  if (s.qg > QRMIN) {
    W dqsdt;
    const W qsi = qs_i<R>(s.tz, den_d, dqsdt);
    const W dq = fdiv<R>(s.qv - qsi, 1.0 + tcpk * dqsdt);
    R pgsub = R(fdiv<R>(s.qv, qsi) - 1.0) * R(s.qg);
    W pg;
    if (pgsub > R(0.0)) {
      if (s.tz > tice_d)
        pg = 0.0;
      else
        pg = cap<R>(mmin(mmin(W(c.fac_v2g * pgsub), 0.2 * dq), mmin(s.ql + s.qr, fdiv<R>(tice_d - s.tz, tcpk))),
                    s.qv);
    } else {
      pg = -cap<R>(-(mmax(W(c.fac_g2v * pgsub), dq) * W(mmin(R(1.0), R(dim(s.tz, c.t_sub_d)) * R(0.1)))),
                   s.qg);
    }
    s.qg = s.qg + pg;
    s.qv = s.qv - pg;
    q_sol = q_sol + pg;
    cvm = SH_CVM(c, s, q_liq, q_sol);
    s.tz = s.tz + fdiv<R>(pg * (lhl + lhi), cvm);
  }

  // minimum evaporation of rain in dry air
  lhl = c.lv00_d + c.d0_vap_d * s.tz;
  lcpk = fdiv<R>(lhl, cvm);
  if (s.qr > QCMIN) {
    W dqsdt;
    const W qsw2 = qs_w<R>(s.tz, den_d, dqsdt);
    W sink = mmin(s.qr, fdiv<R>(dim(rh_rain * qsw2, s.qv), 1.0 + lcpk * dqsdt));
    s.qv = s.qv + sink;
    s.qr = s.qr - sink;
    q_liq = q_liq - sink;
    cvm = SH_CVM(c, s, q_liq, q_sol);
    s.tz = s.tz - fdiv<R>(sink * lhl, cvm);
  }
  lhl = c.lv00_d + c.d0_vap_d * s.tz;
  cvm = c.c_air_d + (s.qv + q_liq + q_sol) * c.c_vap_d;
  lcpk = fdiv<R>(lhl, cvm);

  // cloud fraction
  if (!c.do_qa) {
    const W q_sol2 = c.rad_snow ? s.qi + s.qs : s.qi;
    const W q_liq2 = c.rad_rain ? s.ql + s.qr : s.ql;
    const W q_cond = q_liq2 + q_sol2;
    const W qpz2 = s.qv + q_cond;
    const W tin2 = s.tz - (lcpk * q_cond + icpk * q_sol2);
    W qstar;
    if constexpr (is_f32<R>::value) {
      // The same three cases with one saturation evaluation for every cell
      // (over ice below tice, over water above) and a second, over water, for
      // the cells between t_wfr and tice: the wavefront runs two evaluations
      // instead of one per case.
      const bool cold = tin2 <= c.t_wfr_d, warm = !cold && tin2 >= tice_d;
      const W qa = qs_pick(tin2, den_d, warm);
      qstar = qa;
      if (!cold && !warm) {
        W g;
        const W qsw3 = qs_fast(tin2, 0, den_d, g);
        W rqi;
        if (q_cond > 3.0e-6)
          rqi = fdiv<R>(q_sol2, q_cond);
        else
          rqi = fdiv<R>(tice_d - tin2, tice_d - c.t_wfr_d);
        qstar = rqi * qa + (1.0 - rqi) * qsw3;
      }
    } else if (tin2 <= c.t_wfr_d) {
      qstar = qs_i1<R>(tin2, den_d);
    } else if (tin2 >= tice_d) {
      qstar = qs_w1<R>(tin2, den_d);
    } else {
      const W qsi = qs_i1<R>(tin2, den_d);
      const W qsw3 = qs_w1<R>(tin2, den_d);
      W rqi;
      if (q_cond > 3.0e-6)
        rqi = fdiv<R>(q_sol2, q_cond);
      else
        rqi = fdiv<R>(tice_d - tin2, tice_d - c.t_wfr_d);
      qstar = rqi * qsi + (1.0 - rqi) * qsw3;
    }
    if (qpz2 > QRMIN) {
      const W dq = mmax(QCMIN, W(h_var) * qpz2);
      const W q_plus = qpz2 + dq;
      const W q_minus = qpz2 - dq;
      if (qstar < q_minus)
        s.qa = s.qa + 1.0;
      else if ((qstar < q_plus) && (q_cond > c.qc_crt_d))
        s.qa = s.qa + fdiv<R>(q_plus - qstar, dq + dq);
    }
  }
}

// ------------------------------------------------ microphysics.py:702-746
// Melting of cloud ice in a layer at or below the first level warmer than
// tice (the caller selects the layers).
template <class R>
SH_FN void sedi_ice_melt(const Coef<R>& c, double tice_d, double ql_mlt, CellR<R>& s) {
  using W = wide_t<R>;
  const W tcd = s.tz - tice_d;
  if ((s.qi > QCMIN) && (tcd > 0.0)) {
    // latent heat and heat capacity in FP64 (both modes)
    const W lhi = LI00 + DC_ICE * s.tz;
    W q_liq = s.ql + s.qr;
    W q_sol = s.qi + s.qs + s.qg;
    W cvm = c.c_air_d + s.qv * c.c_vap_d + q_liq * C_LIQ + q_sol * C_ICE;
    const W icpk = fdiv<R>(lhi, cvm);
    const W sink = mmin(s.qi, fdiv<R>(W(c.fac_imlt) * tcd, icpk));
    const W tmp = mmin(sink, dim(W(ql_mlt), s.ql));
    s.ql = s.ql + tmp;
    s.qr = s.qr + sink - tmp;
    s.qi = s.qi - sink;
    q_liq = q_liq + sink;
    q_sol = q_sol - sink;
    cvm = c.c_air_d + s.qv * c.c_vap_d + q_liq * C_LIQ + q_sol * C_ICE;
    s.tz = s.tz - fdiv<R>(sink * lhi, cvm);
  }
}

// ------------------------------------------------ microphysics.py:1290-1338
// Instant melting of cloud ice above tice (pimlt) or homogeneous freezing of
// cloud water below t_wfr (pihom); returns the heat-capacity terms the rest of
// icloud continues from.
template <class R>
SH_FN void icloud_melt_freeze(const Coef<R>& c, double tice_d, double t_wfr_d, double ql_mlt,
                                                   CellR<R>& s, wide_t<R> den, wide_t<R>& q_liq, wide_t<R>& q_sol,
                                                   wide_t<R>& cvm) {
  using W = wide_t<R>;
  const W lhi = LI00 + DC_ICE * s.tz;
  q_liq = s.ql + s.qr;
  q_sol = s.qi + s.qs + s.qg;
  cvm = SH_CVM(c, s, q_liq, q_sol);
  const W icpk = fdiv<R>(lhi, cvm);
  if ((s.tz > tice_d) && (s.qi > QCMIN)) {
    const W melt = mmin(s.qi, fdiv<R>(W(c.fac_imlt) * (s.tz - tice_d), icpk));
    const W tmp = mmin(melt, dim(ql_mlt, s.ql));
    s.ql = s.ql + tmp;
    s.qr = s.qr + melt - tmp;
    s.qi = s.qi - melt;
    q_liq = q_liq + melt;
    q_sol = q_sol - melt;
    cvm = SH_CVM(c, s, q_liq, q_sol);
    s.tz = s.tz - fdiv<R>(melt * lhi, cvm);
  } else if ((s.tz < t_wfr_d) && (s.ql > QCMIN)) {
    const W dtmp = t_wfr_d - s.tz;
    const R factor = mmin(R(1.0), R(dtmp) / R(DT_FR));
    const W sink = mmin(s.ql * W(factor), fdiv<R>(dtmp, icpk));
    const R qi_crt = c.qi_gen * mmin(c.qi_lim, R(0.1) * R(tice_d - s.tz)) / R(den);
    const W tmp = mmin(sink, dim(W(qi_crt), s.qi));
    s.ql = s.ql - sink;
    s.qs = s.qs + sink - tmp;
    s.qi = s.qi + tmp;
    q_liq = q_liq - sink;
    q_sol = q_sol + sink;
    cvm = SH_CVM(c, s, q_liq, q_sol);
    s.tz = s.tz + fdiv<R>(sink * lhi, cvm);
  }
}

// ------------------------------------------------ f:853-1260
template <class R>
SH_FN void icloud_main(const Coef<R>& c, double tice_d, double tice0_d, CellR<R>& s, R h_var,
                                            double rh_adj, double rh_rain, wide_t<R> den_d, R denfac, R vtgz, R vtrz,
                                            R vtsz, wide_t<R> p1d, R di, wide_t<R> q_liq, wide_t<R> q_sol,
                                            wide_t<R> cvm) {
  using W = wide_t<R>;
  const R dts = c.dts, rdts = c.rdts;
  const R den = R(den_d);
  W lhi = LI00 + DC_ICE * s.tz;
  W icpk = fdiv<R>(lhi, cvm);
  if (p1d >= P_MIN) {
    R pgacr = R(0.0), pgacw = R(0.0);
    R tc = R(s.tz - tice_d);
    if (s.tz - tice_d >= 0.0) {
      R dqs0 = R(fdiv<R>(W(c.ces0), p1d) - s.qv);
      if (s.qs > QCMIN) {
        R psacw, psacr, pracs;
        if (s.ql > QRMIN) {
          R factor = denfac * c.csacw * pw_el<R>(R(s.qs) * den, 0.8125);
          psacw = factor / (R(1.0) + dts * factor) * R(s.ql);
        } else {
          psacw = R(0.0);
        }
        if (s.qr > QRMIN) {
          psacr = mmin(acr3d(vtsz, vtrz, R(s.qr), R(s.qs), c.csacr, ck(c, T_ACCO + 1), ck(c, T_ACCO + 5), ck(c, T_ACCO + 9), den),
                       R(s.qr) * rdts);
          pracs = acr3d(vtrz, vtsz, R(s.qs), R(s.qr), c.cracs, ck(c, T_ACCO + 0), ck(c, T_ACCO + 4), ck(c, T_ACCO + 8), den);
        } else {
          psacr = R(0.0);
          pracs = R(0.0);
        }
        R psmlt = mmax(R(0.0), smlt(c, tc, dqs0, R(s.qs) * den, psacw, psacr, den, denfac));
        W sink = mmin(s.qs, mmin(W(dts * (psmlt + pracs)), fdiv<R>(s.tz - tice_d, icpk)));
        s.qs = s.qs - sink;
        W tmp = mmin(sink, W(dim(c.qs_mlt, R(s.ql))));
        s.ql = s.ql + tmp;
        s.qr = s.qr + sink - tmp;
        q_liq = q_liq + sink;
        q_sol = q_sol - sink;
        cvm = SH_CVM(c, s, q_liq, q_sol);
        s.tz = s.tz - fdiv<R>(sink * lhi, cvm);
        tc = R(s.tz - tice_d);
      }
      lhi = LI00 + DC_ICE * s.tz;
      icpk = fdiv<R>(lhi, cvm);
      if ((s.qg > QCMIN) && (tc > R(0.0))) {
        if (s.qr > QRMIN)
          pgacr = mmin(acr3d(vtgz, vtrz, R(s.qr), R(s.qg), c.cgacr, ck(c, T_ACCO + 2), ck(c, T_ACCO + 6), ck(c, T_ACCO + 10), den),
                       rdts * R(s.qr));
        R qden = R(s.qg) * den;
        if (s.ql > QRMIN) {
          R factor = c.cgacw * qden / msqrt(den * msqrt(msqrt(qden)));
          pgacw = factor / (R(1.0) + dts * factor) * R(s.ql);
        }
        R pgmlt = dts * gmlt(c, tc, dqs0, qden, pgacw, pgacr, den);
        W pm = mmin(W(mmax(R(0.0), pgmlt)), mmin(s.qg, fdiv<R>(s.tz - tice_d, icpk)));
        s.qg = s.qg - pm;
        s.qr = s.qr + pm;
        q_liq = q_liq + pm;
        q_sol = q_sol - pm;
        cvm = SH_CVM(c, s, q_liq, q_sol);
        s.tz = s.tz - fdiv<R>(pm * lhi, cvm);
      }
    } else {
      if (s.qi > 3.0e-7) {
        W psaci;
        if (s.qs > 1.0e-7) {
          R e;
          if constexpr (is_f32<R>::value)
            e = mexp(R(0.05) * tc) * pw_split(R(s.qs) * den, 0.8125);
          else
            e = mexp(R(0.05) * tc + R(0.8125) * mlog(R(s.qs) * den));
          R factor = dts * denfac * c.csaci * e;
          psaci = implicit_take<R>(factor, s.qi);
        } else {
          psaci = W(0.0);
        }
        R qim = c.qi0_crt / den;
        R tmp = c.const_vi ? c.fac_i2s : c.fac_i2s * mexp(R(0.025) * tc);
        di = mmax(di, R(QRMIN));
        const W q_plus = s.qi + W(di);
        const W qimd = W(qim);
        R psaut;
        if (q_plus > (qimd + QRMIN)) {
          R dq;
          if (qimd > (s.qi - W(di)))
            dq = (R(0.25) * R((q_plus - qimd) * (q_plus - qimd))) / di;
          else
            dq = R(s.qi - qimd);
          psaut = tmp * dq;
        } else {
          psaut = R(0.0);
        }
        W sink = mmin(0.75 * s.qi, psaci + W(psaut));
        s.qi = s.qi - sink;
        s.qs = s.qs + sink;
        if (s.qg > 1.0e-6) {
          R factor = dts * c.cgaci * msqrt(den) * R(s.qg);
          W pgaci = implicit_take<R>(factor, s.qi);
          s.qi = s.qi - pgaci;
          s.qg = s.qg + pgaci;
        }
      }
      tc = R(s.tz - tice_d);
      if ((s.qr > 1e-7) && (s.tz - tice_d < 0.0)) {
        R psacr;
        if (s.qs > 1.0e-7)
          psacr = dts * acr3d(vtsz, vtrz, R(s.qr), R(s.qs), c.csacr, ck(c, T_ACCO + 1), ck(c, T_ACCO + 5), ck(c, T_ACCO + 9), den);
        else
          psacr = R(0.0);
        R pgfr = dts * ck(c, T_CGFR + 0) / den * (exp_lin<R>(-ck(c, T_CGFR + 1), -c.cgfr1_d, s.tz - tice_d) - R(1.0)) *
                 pw_el<R>(R(s.qr) * den, 1.75);
        // the limiter that can drain all rain is applied in FP64, so a fully
        // frozen layer keeps no FP32 residue
        const W sink = W(psacr) + W(pgfr);
        const W factor = fdiv<R>(mmin(sink, mmin(s.qr, fdiv<R>(-(s.tz - tice_d), icpk))), mmax(sink, QRMIN));
        W ps = factor * W(psacr), pg = factor * W(pgfr);
        if constexpr (is_f32<R>::value) {
          W tot = ps + pg;
          if (tot > s.qr) {
            W f = fdiv<R>(s.qr, tot);
            ps *= f;
            pg = s.qr - ps;
          }
        }
        W sk = ps + pg;
        s.qr = s.qr - sk;
        s.qs = s.qs + ps;
        s.qg = s.qg + pg;
        q_liq = q_liq - sk;
        q_sol = q_sol + sk;
        cvm = SH_CVM(c, s, q_liq, q_sol);
        s.tz = s.tz + fdiv<R>(sk * lhi, cvm);
      }
      lhi = LI00 + DC_ICE * s.tz;
      icpk = fdiv<R>(lhi, cvm);
      if (s.qs > 1.0e-7) {
        R sink;
        if (s.qg > QRMIN)
          sink = dts * acr3d(vtgz, vtsz, R(s.qs), R(s.qg), c.cgacs, ck(c, T_ACCO + 3), ck(c, T_ACCO + 7), ck(c, T_ACCO + 11), den);
        else
          sink = R(0.0);
        R qsm = c.qs0_crt / den;
        if (s.qs > W(qsm)) {
          R factor = dts * R(1.0e-3) * exp_lin<R>(R(0.09), 0.09, s.tz - tice_d);
          sink = sink + factor / (R(1.0) + factor) * R(s.qs - W(qsm));
        }
        W sk = mmin(s.qs, W(sink));
        s.qs = s.qs - sk;
        s.qg = s.qg + sk;
      }
      if ((s.qg > 1.0e-7) && (s.tz < tice0_d)) {
        // both collections are bounded by the FP64 species they take from,
        // so a layer whose rain the graupel collects entirely keeps none
        W gw, gr;
        if (s.ql > 1.0e-6) {
          R qden = R(s.qg) * den;
          R factor = dts * c.cgacw * qden / msqrt(den * msqrt(msqrt(qden)));
          gw = implicit_take<R>(factor, s.ql);
        } else {
          gw = W(0.0);
        }
        if (s.qr > 1.0e-6)
          gr = mmin(W(dts * acr3d(vtgz, vtrz, R(s.qr), R(s.qg), c.cgacr, ck(c, T_ACCO + 2), ck(c, T_ACCO + 6),
                                  ck(c, T_ACCO + 10), den)),
                    s.qr);
        else
          gr = W(0.0);
        const W sink = gr + gw;
        const W factor = fdiv<R>(mmin(sink, fdiv<R>(dim(tice_d, s.tz), icpk)), mmax(sink, QRMIN));
        const W pr = cap<R>(factor * gr, s.qr), pl = cap<R>(factor * gw, s.ql);
        W sk = pr + pl;
        s.qg = s.qg + sk;
        s.qr = s.qr - pr;
        s.ql = s.ql - pl;
        q_liq = q_liq - sk;
        q_sol = q_sol + sk;
        cvm = SH_CVM(c, s, q_liq, q_sol);
        s.tz = s.tz + fdiv<R>(sk * lhi, cvm);
      }
    }
  }
  subgrid_z_proc<R>(c, tice_d, s, h_var, rh_adj, rh_rain, den_d, denfac, p1d);
}

}  // namespace shield
