// mp_point_adj.h -- hand-written reverse mode of FP64 point-wise processes of
// one cell: the icloud step -- icloud_melt_freeze (microphysics.py:1290-1338),
// icloud_main (microphysics_funcs.py:853-1260, without subgrid_z_proc) and
// subgrid_z_proc (microphysics_funcs.py:431-850) -- in warm_rain, the
// evaporation / accretion of rain revap_racc (microphysics_funcs.py:183-265),
// the autoconversion of cloud water (microphysics_funcs.py:375-428) and the
// rain fall speed, and in sedimentation the ice melt and the fall speeds.
//
// The icloud step is cut into blocks (one process each: a melt, a collection, a
// deposition, ...).  Every block reads the cell state X at its entry -- the
// eight cell values and the running q_liq, q_sol, cvm -- and derives its latent
// heats, their ratios to cvm, tc and dqs0 from X, so X is the whole interface
// between blocks.  The forward sweep keeps X at the entry of every block; the
// reverse sweep walks the blocks backwards, and each block's transpose first
// re-evaluates the block's forward from its entry state (the same FP64
// expressions, in the same order, as the process code of mp_process.h, so
// every branch is the one the forward took) and then propagates the
// cotangent of its output state to its entry state.
//
// Divisions: the re-evaluated forward values keep the IEEE division of the
// process code, so they and their branches are the forward's bit for bit;
// the cotangent expressions multiply by rcp64 (v_rcp_f64 and one Newton step,
// 2.2e-15 relative), a third of the IEEE sequence's instructions and a shorter
// dependent chain.  The same holds for the column reverses of mp_adjoint.h.
//
// Derivative conventions are those of the dual-number path and of the
// reference's select-based min / max: min(a, b) follows a where a < b and b
// otherwise; max(a, b) follows a where a > b; dim(x, y) = x - y where that is
// positive, 0 otherwise; |x| has slope -1 where x < 0 and +1 otherwise.
#pragma once

#include "mp_process.h"

namespace shield {

// state at a block boundary
struct IcX {
  double qv, ql, qr, qi, qs, qg, tz, qa;
  double q_liq, q_sol, cvm;
};

// the cell-wise inputs besides the state, and their cotangents
struct IcIn {
  double den, denfac, vtg, vtr, vts, p1;
};
struct IcInB {
  double den, denfac, vtg, vtr, vts, p1, qi_gen;
  double dl;  // the ice slope limiter's value (z_slope_ice)
};

struct IcK {
  const Coef<double>& c;
  double tice, t_wfr, ql_mlt, tice0, hv, rh_adj, rh_rain;
  IcIn in;
  // z_slope_ice: the ice autoconversion's width takes the column's slope
  // limiter value dl of the post-melt cloud ice where it is larger
  bool slope = false;
  double dl = 0.0;
};

SH_FN double ic_cvm(const Coef<double>& c, double qv, double q_liq, double q_sol) {
  return c.c_air_d + qv * c.c_vap_d + q_liq * C_LIQ + q_sol * C_ICE;
}
// Transpose of one latent-heat transfer that closes a block:
//   tz1 = tz + sg (a L) / cvm1,  cvm1 = cvm(qv1, q_liq1, q_sol1)
// xb holds the cotangents after the block; on return its tz, qv, q_liq,
// q_sol entries are those of the values before the heating and the cvm
// update (cvm1 replaces the entry cvm, so the entry cvm's cotangent from
// here is zero).  Returns the cotangent of a; adds L's to *lb.
SH_FN double heat_rev(const Coef<double>& c, double sg, double a, double L, double cvm1, IcX& xb, double& lb) {
  const double tb = xb.tz;
  const double dT = (a * L) / cvm1;
  const double ab = sg * tb * L * rcp64(cvm1);
  lb += sg * tb * a * rcp64(cvm1);
  const double cb = xb.cvm - sg * tb * dT * rcp64(cvm1);
  xb.qv += cb * c.c_vap_d;
  xb.q_liq += cb * C_LIQ;
  xb.q_sol += cb * C_ICE;
  xb.cvm = 0.0;
  return ab;
}

// Transpose of the saturation functions at (ta, den): q = wqs1 / iqs1 and, for
// the two-output forms, dqdt of wqs2 / iqs2.  Where iqs clamps ta (below
// T_SAT_MIN, above TICE + 102) q and dqdt do not depend on ta.
SH_FN void sat_rev(bool ice, double ta, double den, double q, double dqdt, double qb, double dqdtb, double& tab,
                   double& denb) {
  double d = DC_VAP, l = LV0;
  bool active = true;
  if (ice) {
    if (ta < TICE) {
      active = ta >= T_SAT_MIN;
      d = D2ICE;
      l = LI2;
    } else {
      active = ta <= TICE + 102.0;
    }
  }
  denb -= (qb * q + dqdtb * dqdt) * rcp64(den);
  if (active) {
    const double g = (d + l / ta) / (RVGAS * ta);
    const double qt = q * (g - 1.0 / ta);
    const double gp = -(d + 2.0 * l / ta) / (RVGAS * ta * ta);
    tab += qb * qt + dqdtb * (qt * g + q * gp);
  }
}

// Transpose of acr3d (microphysics_funcs.py:149-161) with respect to its
// fall speeds, mixing ratios and density.
SH_FN void acr3d_rev(double v1, double v2, double q1, double q2, double cc, double a0, double a1, double a2,
                     double rho, double ob, double& v1b, double& v2b, double& q1b, double& q2b, double& rhob) {
  const double t1 = sqrt(q1 * rho);
  const double s1 = sqrt(q2 * rho);
  const double s2 = sqrt(s1);
  const double st1 = sqrt(t1);
  const double dv = v1 - v2;
  const double sg = dv < 0.0 ? -1.0 : 1.0;
  const double ad = sg * dv;
  const double P = cc * ad * q1 * s2;
  const double Q = a0 * t1 + a1 * st1 * s2 + a2 * s1;
  const double Pb = ob * Q, Qb = ob * P;
  const double adb = Pb * cc * q1 * s2;
  q1b += Pb * cc * ad * s2;
  double s2b = Pb * cc * ad * q1 + Qb * a1 * st1;
  const double t1b = Qb * a0 + Qb * a1 * s2 * 0.5 * rcp64(st1);
  const double s1b = Qb * a2 + s2b * 0.5 * rcp64(s2);
  q1b += t1b * 0.5 * rcp64(t1) * rho;
  rhob += t1b * 0.5 * rcp64(t1) * q1 + s1b * 0.5 * rcp64(s1) * q2;
  q2b += s1b * 0.5 * rcp64(s1) * rho;
  v1b += adb * sg;
  v2b -= adb * sg;
}

// x / y with cotangent ob: adds to xb, yb
SH_FN void div_rev(double x, double y, double ob, double& xb, double& yb) {
  const double q = x / y;
  xb += ob * rcp64(y);
  yb -= ob * q * rcp64(y);
}

// the derived heats of a block entry: lhl, lhi, lcpk = lhl / cvm,
// icpk = lhi / cvm; their transpose onto tz and cvm
struct IcHeat {
  double lhl, lhi, lcpk, icpk;
};
SH_FN IcHeat ic_heat(const Coef<double>& c, const IcX& x) {
  IcHeat h;
  h.lhl = c.lv00_d + c.d0_vap_d * x.tz;
  h.lhi = LI00 + DC_ICE * x.tz;
  h.lcpk = h.lhl / x.cvm;
  h.icpk = h.lhi / x.cvm;
  return h;
}
SH_FN void ic_heat_rev(const Coef<double>& c, const IcX& x, const IcHeat& h, double lhlb, double lhib, double lcpkb,
                       double icpkb, IcX& xb) {
  lhlb += lcpkb * rcp64(x.cvm);
  lhib += icpkb * rcp64(x.cvm);
  xb.cvm -= (lcpkb * h.lcpk + icpkb * h.icpk) * rcp64(x.cvm);
  xb.tz += lhlb * c.d0_vap_d + lhib * DC_ICE;
}

// ------------------------------------------------------------ melt / freeze
// icloud_melt_freeze: reads the cell, writes q_liq, q_sol, cvm
SH_FN int blk_melt_freeze(const IcK& k, IcX& x, bool freeze = true) {
  const Coef<double>& c = k.c;
  const double lhi = LI00 + DC_ICE * x.tz;
  x.q_liq = x.ql + x.qr;
  x.q_sol = x.qi + x.qs + x.qg;
  x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
  const double icpk = lhi / x.cvm;
  if ((x.tz > k.tice) && (x.qi > QCMIN)) {
    const double melt = mmin(x.qi, (c.fac_imlt * (x.tz - k.tice)) / icpk);
    const double tmp = mmin(melt, dim(k.ql_mlt, x.ql));
    x.ql = x.ql + tmp;
    x.qr = x.qr + melt - tmp;
    x.qi = x.qi - melt;
    x.q_liq = x.q_liq + melt;
    x.q_sol = x.q_sol - melt;
    x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
    x.tz = x.tz - (melt * lhi) / x.cvm;
    return 1;
  } else if (freeze && (x.tz < k.t_wfr) && (x.ql > QCMIN)) {
    const double dtmp = k.t_wfr - x.tz;
    const double factor = mmin(1.0, dtmp / DT_FR);
    const double sink = mmin(x.ql * factor, dtmp / icpk);
    const double qi_crt = c.qi_gen * mmin(c.qi_lim, 0.1 * (k.tice - x.tz)) / k.in.den;
    const double tmp = mmin(sink, dim(qi_crt, x.qi));
    x.ql = x.ql - sink;
    x.qs = x.qs + sink - tmp;
    x.qi = x.qi + tmp;
    x.q_liq = x.q_liq - sink;
    x.q_sol = x.q_sol + sink;
    x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
    x.tz = x.tz + (sink * lhi) / x.cvm;
    return 2;
  }
  return 0;
}

SH_FN void blk_melt_freeze_rev(const IcK& k, const IcX& x, IcX& xb, IcInB& ib, bool freeze = true) {
  const Coef<double>& c = k.c;
  IcX y = x;
  const int br = blk_melt_freeze(k, y, freeze);
  const double lhi = LI00 + DC_ICE * x.tz;
  const double q_liq0 = x.ql + x.qr, q_sol0 = x.qi + x.qs + x.qg;
  const double cvm0 = ic_cvm(c, x.qv, q_liq0, q_sol0);
  const double icpk = lhi / cvm0;
  double lhib = 0.0, icpkb = 0.0;
  if (br == 1) {
    const double a = (c.fac_imlt * (x.tz - k.tice)) / icpk;
    const double melt = mmin(x.qi, a);
    const double dl = dim(k.ql_mlt, x.ql);
    double meltb = heat_rev(c, -1.0, melt, lhi, y.cvm, xb, lhib);
    meltb += xb.q_liq - xb.q_sol + xb.qr - xb.qi;
    const double tmpb = xb.ql - xb.qr;
    if (melt < dl)
      meltb += tmpb;
    else if (k.ql_mlt - x.ql > 0.0)
      xb.ql -= tmpb;
    if (x.qi < a) {
      xb.qi += meltb;
    } else {
      xb.tz += meltb * c.fac_imlt * rcp64(icpk);
      icpkb -= meltb * a * rcp64(icpk);
    }
  } else if (br == 2) {
    const double dtmp = k.t_wfr - x.tz;
    const double r8 = dtmp / DT_FR;
    const double factor = mmin(1.0, r8);
    const double s1 = x.ql * factor, s2 = dtmp / icpk;
    const double sink = mmin(s1, s2);
    const double xq = 0.1 * (k.tice - x.tz);
    const double mq = mmin(c.qi_lim, xq);
    const double qi_crt = c.qi_gen * mq / k.in.den;
    const double dd = dim(qi_crt, x.qi);
    double sinkb = heat_rev(c, 1.0, sink, lhi, y.cvm, xb, lhib);
    sinkb += xb.q_sol - xb.q_liq + xb.qs - xb.ql;
    const double tmpb = xb.qi - xb.qs;
    if (sink < dd) {
      sinkb += tmpb;
    } else if (qi_crt - x.qi > 0.0) {
      const double qcb = tmpb;
      xb.qi -= qcb;
      ib.qi_gen += qcb * mq * rcp64(k.in.den);
      ib.den -= qcb * qi_crt * rcp64(k.in.den);
      if (!(c.qi_lim < xq)) xb.tz -= qcb * c.qi_gen * rcp64(k.in.den) * 0.1;
    }
    double dtmpb = 0.0;
    if (s1 < s2) {
      xb.ql += sinkb * factor;
      if (!(1.0 < r8)) dtmpb += sinkb * x.ql * (1.0 / DT_FR);
    } else {
      dtmpb += sinkb * rcp64(icpk);
      icpkb -= sinkb * s2 * rcp64(icpk);
    }
    xb.tz -= dtmpb;
  }
  // icpk = lhi / cvm0; cvm0 = cvm(qv, q_liq0, q_sol0)
  lhib += icpkb * rcp64(cvm0);
  const double cb = xb.cvm - icpkb * icpk * rcp64(cvm0);
  xb.qv += cb * c.c_vap_d;
  const double qlb = xb.q_liq + cb * C_LIQ, qsb = xb.q_sol + cb * C_ICE;
  xb.ql += qlb;
  xb.qr += qlb;
  xb.qi += qsb;
  xb.qs += qsb;
  xb.qg += qsb;
  xb.q_liq = xb.q_sol = xb.cvm = 0.0;
  xb.tz += lhib * DC_ICE;
}
// ------------------------------------------------------- icloud_main, warm
// snow melt (qs > QCMIN)
SH_FN bool blk_w1(const IcK& k, IcX& x) {
  if (!(x.qs > QCMIN)) return false;
  const Coef<double>& c = k.c;
  const double den = k.in.den, denfac = k.in.denfac;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double icpk = lhi / x.cvm;
  const double tc = x.tz - k.tice;
  const double dqs0 = c.ces0 / k.in.p1 - x.qv;
  double psacw = 0.0, psacr = 0.0, pracs = 0.0;
  if (x.ql > QRMIN) {
    const double factor = denfac * c.csacw * pw_el<double>(x.qs * den, 0.8125);
    psacw = factor / (1.0 + c.dts * factor) * x.ql;
  }
  if (x.qr > QRMIN) {
    psacr = mmin(acr3d<double>(k.in.vts, k.in.vtr, x.qr, x.qs, c.csacr, ck(c, T_ACCO + 1), ck(c, T_ACCO + 5),
                               ck(c, T_ACCO + 9), den),
                 x.qr * c.rdts);
    pracs = acr3d<double>(k.in.vtr, k.in.vts, x.qs, x.qr, c.cracs, ck(c, T_ACCO + 0), ck(c, T_ACCO + 4),
                          ck(c, T_ACCO + 8), den);
  }
  const double psmlt = mmax(0.0, smlt<double>(c, tc, dqs0, x.qs * den, psacw, psacr, den, denfac));
  const double sink = mmin(x.qs, mmin(c.dts * (psmlt + pracs), (x.tz - k.tice) / icpk));
  x.qs = x.qs - sink;
  const double tmp = mmin(sink, dim(c.qs_mlt, x.ql));
  x.ql = x.ql + tmp;
  x.qr = x.qr + sink - tmp;
  x.q_liq = x.q_liq + sink;
  x.q_sol = x.q_sol - sink;
  x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
  x.tz = x.tz - (sink * lhi) / x.cvm;
  return true;
}

SH_FN void blk_w1_rev(const IcK& k, const IcX& x, IcX& xb, IcInB& ib) {
  IcX y = x;
  if (!blk_w1(k, y)) return;
  const Coef<double>& c = k.c;
  const double den = k.in.den, denfac = k.in.denfac;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double icpk = lhi / x.cvm;
  const double tc = x.tz - k.tice;
  const double dqs0 = c.ces0 / k.in.p1 - x.qv;
  double psacw = 0.0, psacr = 0.0, pracs = 0.0, a1 = 0.0, a2 = 0.0;
  if (x.ql > QRMIN) {
    const double factor = denfac * c.csacw * pw_el<double>(x.qs * den, 0.8125);
    psacw = factor / (1.0 + c.dts * factor) * x.ql;
  }
  if (x.qr > QRMIN) {
    a1 = acr3d<double>(k.in.vts, k.in.vtr, x.qr, x.qs, c.csacr, ck(c, T_ACCO + 1), ck(c, T_ACCO + 5),
                       ck(c, T_ACCO + 9), den);
    a2 = x.qr * c.rdts;
    psacr = mmin(a1, a2);
    pracs = acr3d<double>(k.in.vtr, k.in.vts, x.qs, x.qr, c.cracs, ck(c, T_ACCO + 0), ck(c, T_ACCO + 4),
                          ck(c, T_ACCO + 8), den);
  }
  const double sm = smlt<double>(c, tc, dqs0, x.qs * den, psacw, psacr, den, denfac);
  const double psmlt = mmax(0.0, sm);
  const double b1 = c.dts * (psmlt + pracs), b2 = (x.tz - k.tice) / icpk;
  const double inner = mmin(b1, b2);
  const double sink = mmin(x.qs, inner);
  const double dl = dim(c.qs_mlt, x.ql);
  double lhib = 0.0, icpkb = 0.0, tcb = 0.0, dqs0b = 0.0;
  double sinkb = heat_rev(c, -1.0, sink, lhi, y.cvm, xb, lhib);
  sinkb += xb.q_liq - xb.q_sol + xb.qr - xb.qs;
  const double tmpb = xb.ql - xb.qr;
  if (sink < dl)
    sinkb += tmpb;
  else if (c.qs_mlt - x.ql > 0.0)
    xb.ql -= tmpb;
  double innerb = 0.0;
  if (x.qs < inner)
    xb.qs += sinkb;
  else
    innerb = sinkb;
  double b1b = 0.0;
  if (b1 < b2) {
    b1b = innerb;
  } else {
    xb.tz += innerb * rcp64(icpk);
    icpkb -= innerb * b2 * rcp64(icpk);
  }
  const double psmltb = b1b * c.dts, pracsb = b1b * c.dts;
  const double smb = (0.0 > sm) ? 0.0 : psmltb;
  // smlt (microphysics_funcs.py:163-170)
  const int kk = T_CSMLT;
  const double qsrho = x.qs * den;
  const double sq = sqrt(qsrho), pwq = pow(qsrho, 0.65625), sdf = sqrt(denfac);
  const double A = ck(c, kk) * tc / den - ck(c, kk + 1) * dqs0;
  const double B = ck(c, kk + 2) * sq + ck(c, kk + 3) * pwq * sdf;
  const double Ab = smb * B, Bb = smb * A;
  tcb += smb * ck(c, kk + 4) * (psacw + psacr);
  const double psacwb = smb * ck(c, kk + 4) * tc, psacrb = psacwb;
  tcb += Ab * ck(c, kk) * rcp64(den);
  ib.den -= Ab * (ck(c, kk) * tc * rcp64(den)) * rcp64(den);
  dqs0b -= Ab * ck(c, kk + 1);
  const double qsrhob = Bb * (ck(c, kk + 2) * 0.5 * rcp64(sq) + ck(c, kk + 3) * 0.65625 * pwq * rcp64(qsrho) * sdf);
  ib.denfac += Bb * ck(c, kk + 3) * pwq * 0.5 * rcp64(sdf);
  xb.qs += qsrhob * den;
  ib.den += qsrhob * x.qs;
  if (x.qr > QRMIN) {
    acr3d_rev(k.in.vtr, k.in.vts, x.qs, x.qr, c.cracs, ck(c, T_ACCO + 0), ck(c, T_ACCO + 4), ck(c, T_ACCO + 8), den,
              pracsb, ib.vtr, ib.vts, xb.qs, xb.qr, ib.den);
    if (a1 < a2)
      acr3d_rev(k.in.vts, k.in.vtr, x.qr, x.qs, c.csacr, ck(c, T_ACCO + 1), ck(c, T_ACCO + 5), ck(c, T_ACCO + 9),
                den, psacrb, ib.vts, ib.vtr, xb.qr, xb.qs, ib.den);
    else
      xb.qr += psacrb * c.rdts;
  }
  if (x.ql > QRMIN) {
    const double pe = pw_el<double>(x.qs * den, 0.8125);
    const double factor = denfac * c.csacw * pe;
    const double D = 1.0 + c.dts * factor;
    const double fb = psacwb * x.ql * rcp64(D * D);
    xb.ql += psacwb * factor * rcp64(D);
    ib.denfac += fb * c.csacw * pe;
    const double qdb = fb * denfac * c.csacw * 0.8125 * pe * rcp64(x.qs * den);
    xb.qs += qdb * den;
    ib.den += qdb * x.qs;
  }
  ib.p1 -= dqs0b * (c.ces0 * rcp64(k.in.p1)) * rcp64(k.in.p1);
  xb.qv -= dqs0b;
  xb.tz += tcb;
  lhib += icpkb * rcp64(x.cvm);
  xb.cvm -= icpkb * icpk * rcp64(x.cvm);
  xb.tz += lhib * DC_ICE;
}

// graupel melt (qg > QCMIN, tc > 0)
SH_FN bool blk_w2(const IcK& k, IcX& x) {
  const double tc = x.tz - k.tice;
  if (!((x.qg > QCMIN) && (tc > 0.0))) return false;
  const Coef<double>& c = k.c;
  const double den = k.in.den;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double icpk = lhi / x.cvm;
  const double dqs0 = c.ces0 / k.in.p1 - x.qv;
  double pgacr = 0.0, pgacw = 0.0;
  if (x.qr > QRMIN)
    pgacr = mmin(acr3d<double>(k.in.vtg, k.in.vtr, x.qr, x.qg, c.cgacr, ck(c, T_ACCO + 2), ck(c, T_ACCO + 6),
                               ck(c, T_ACCO + 10), den),
                 c.rdts * x.qr);
  const double qden = x.qg * den;
  if (x.ql > QRMIN) {
    const double factor = c.cgacw * qden / sqrt(den * sqrt(sqrt(qden)));
    pgacw = factor / (1.0 + c.dts * factor) * x.ql;
  }
  const double pgmlt = c.dts * gmlt<double>(c, tc, dqs0, qden, pgacw, pgacr, den);
  const double pm = mmin(mmax(0.0, pgmlt), mmin(x.qg, (x.tz - k.tice) / icpk));
  x.qg = x.qg - pm;
  x.qr = x.qr + pm;
  x.q_liq = x.q_liq + pm;
  x.q_sol = x.q_sol - pm;
  x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
  x.tz = x.tz - (pm * lhi) / x.cvm;
  return true;
}

SH_FN void blk_w2_rev(const IcK& k, const IcX& x, IcX& xb, IcInB& ib) {
  IcX y = x;
  if (!blk_w2(k, y)) return;
  const Coef<double>& c = k.c;
  const double den = k.in.den;
  const double tc = x.tz - k.tice;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double icpk = lhi / x.cvm;
  const double dqs0 = c.ces0 / k.in.p1 - x.qv;
  double pgacr = 0.0, pgacw = 0.0, a1 = 0.0, a2 = 0.0;
  if (x.qr > QRMIN) {
    a1 = acr3d<double>(k.in.vtg, k.in.vtr, x.qr, x.qg, c.cgacr, ck(c, T_ACCO + 2), ck(c, T_ACCO + 6),
                       ck(c, T_ACCO + 10), den);
    a2 = c.rdts * x.qr;
    pgacr = mmin(a1, a2);
  }
  const double qden = x.qg * den;
  const double sq2 = sqrt(qden), sq4 = sqrt(sq2), dn = sqrt(den * sq4);
  const double factor = c.cgacw * qden / dn;
  if (x.ql > QRMIN) pgacw = factor / (1.0 + c.dts * factor) * x.ql;
  const double pgmlt = c.dts * gmlt<double>(c, tc, dqs0, qden, pgacw, pgacr, den);
  const double m0 = mmax(0.0, pgmlt), b2 = (x.tz - k.tice) / icpk, m1 = mmin(x.qg, b2);
  const double pm = mmin(m0, m1);
  double lhib = 0.0, icpkb = 0.0, tcb = 0.0, dqs0b = 0.0;
  double pmb = heat_rev(c, -1.0, pm, lhi, y.cvm, xb, lhib);
  pmb += xb.q_liq - xb.q_sol + xb.qr - xb.qg;
  double m0b = 0.0, m1b = 0.0;
  if (m0 < m1)
    m0b = pmb;
  else
    m1b = pmb;
  if (x.qg < b2) {
    xb.qg += m1b;
  } else {
    xb.tz += m1b * rcp64(icpk);
    icpkb -= m1b * b2 * rcp64(icpk);
  }
  const double gmb = ((0.0 > pgmlt) ? 0.0 : m0b) * c.dts;
  // gmlt (microphysics_funcs.py:172-179)
  const int kk = T_CGMLT;
  const double pq = pow(qden, 0.6875), pr = pow(den, 0.25);
  const double A = ck(c, kk) * tc / den - ck(c, kk + 1) * dqs0;
  const double B = ck(c, kk + 2) * sq2 + ck(c, kk + 3) * pq / pr;
  const double Ab = gmb * B, Bb = gmb * A;
  tcb += gmb * ck(c, kk + 4) * (pgacw + pgacr);
  const double pgacwb = gmb * ck(c, kk + 4) * tc, pgacrb = pgacwb;
  tcb += Ab * ck(c, kk) * rcp64(den);
  ib.den -= Ab * (ck(c, kk) * tc * rcp64(den)) * rcp64(den);
  dqs0b -= Ab * ck(c, kk + 1);
  double qdenb = Bb * (ck(c, kk + 2) * 0.5 * rcp64(sq2) + ck(c, kk + 3) * 0.6875 * pq * rcp64(qden) * rcp64(pr));
  ib.den -= Bb * ck(c, kk + 3) * (pq * rcp64(pr)) * 0.25 * rcp64(den);
  if (x.ql > QRMIN) {
    const double D = 1.0 + c.dts * factor;
    const double fb = pgacwb * x.ql * rcp64(D * D);
    xb.ql += pgacwb * factor * rcp64(D);
    qdenb += fb * c.cgacw * rcp64(dn);
    const double dnb = -fb * factor * rcp64(dn);
    ib.den += dnb * 0.5 * rcp64(dn) * sq4;
    const double sq4b = dnb * 0.5 * rcp64(dn) * den;
    qdenb += sq4b * 0.5 * rcp64(sq4) * 0.5 * rcp64(sq2);
  }
  xb.qg += qdenb * den;
  ib.den += qdenb * x.qg;
  if (x.qr > QRMIN) {
    if (a1 < a2)
      acr3d_rev(k.in.vtg, k.in.vtr, x.qr, x.qg, c.cgacr, ck(c, T_ACCO + 2), ck(c, T_ACCO + 6), ck(c, T_ACCO + 10),
                den, pgacrb, ib.vtg, ib.vtr, xb.qr, xb.qg, ib.den);
    else
      xb.qr += pgacrb * c.rdts;
  }
  ib.p1 -= dqs0b * (c.ces0 * rcp64(k.in.p1)) * rcp64(k.in.p1);
  xb.qv -= dqs0b;
  xb.tz += tcb;
  lhib += icpkb * rcp64(x.cvm);
  xb.cvm -= icpkb * icpk * rcp64(x.cvm);
  xb.tz += lhib * DC_ICE;
}

// ------------------------------------------------------- icloud_main, cold
// ice -> snow (accretion psaci, autoconversion psaut) and ice -> graupel
// (pgaci), for qi > 3e-7
SH_FN bool blk_c1(const IcK& k, IcX& x) {
  if (!(x.qi > 3.0e-7)) return false;
  const Coef<double>& c = k.c;
  const double den = k.in.den, denfac = k.in.denfac;
  const double tc = x.tz - k.tice;
  double psaci = 0.0;
  if (x.qs > 1.0e-7) {
    const double e = exp(0.05 * tc + 0.8125 * log(x.qs * den));
    const double factor = c.dts * denfac * c.csaci * e;
    psaci = factor / (1.0 + factor) * x.qi;
  }
  const double qim = c.qi0_crt / den;
  const double tmp = c.const_vi ? c.fac_i2s : c.fac_i2s * exp(0.025 * tc);
  const double dm = mmax(QVMIN, k.hv * x.qi);
  const double di = mmax(k.slope ? mmax(k.dl, dm) : dm, QRMIN);
  const double q_plus = x.qi + di;
  double psaut = 0.0;
  if (q_plus > (qim + QRMIN)) {
    double dq;
    if (qim > (x.qi - di))
      dq = (0.25 * ((q_plus - qim) * (q_plus - qim))) / di;
    else
      dq = x.qi - qim;
    psaut = tmp * dq;
  }
  const double sink = mmin(0.75 * x.qi, psaci + psaut);
  x.qi = x.qi - sink;
  x.qs = x.qs + sink;
  if (x.qg > 1.0e-6) {
    const double factor = c.dts * c.cgaci * sqrt(den) * x.qg;
    const double pgaci = factor / (1.0 + factor) * x.qi;
    x.qi = x.qi - pgaci;
    x.qg = x.qg + pgaci;
  }
  return true;
}

SH_FN void blk_c1_rev(const IcK& k, const IcX& x, IcX& xb, IcInB& ib) {
  if (!(x.qi > 3.0e-7)) return;
  const Coef<double>& c = k.c;
  const double den = k.in.den, denfac = k.in.denfac;
  const double tc = x.tz - k.tice;
  double psaci = 0.0, e = 0.0, f1 = 0.0;
  if (x.qs > 1.0e-7) {
    e = exp(0.05 * tc + 0.8125 * log(x.qs * den));
    f1 = c.dts * denfac * c.csaci * e;
    psaci = f1 / (1.0 + f1) * x.qi;
  }
  const double qim = c.qi0_crt / den;
  const double tmp = c.const_vi ? c.fac_i2s : c.fac_i2s * exp(0.025 * tc);
  const double dm = mmax(QVMIN, k.hv * x.qi);
  const double di0 = k.slope ? mmax(k.dl, dm) : dm;
  const double di = mmax(di0, QRMIN);
  const double q_plus = x.qi + di;
  double psaut = 0.0, dq = 0.0;
  const bool aut = q_plus > (qim + QRMIN);
  const bool quad = qim > (x.qi - di);
  if (aut) {
    dq = quad ? (0.25 * ((q_plus - qim) * (q_plus - qim))) / di : x.qi - qim;
    psaut = tmp * dq;
  }
  const double sink = mmin(0.75 * x.qi, psaci + psaut);
  double qi1b = xb.qi;
  if (x.qg > 1.0e-6) {
    const double qi1 = x.qi - sink;
    const double sd = sqrt(den);
    const double f2 = c.dts * c.cgaci * sd * x.qg;
    const double D = 1.0 + f2;
    const double pgacib = xb.qg - xb.qi;
    qi1b += pgacib * f2 * rcp64(D);
    const double f2b = pgacib * qi1 * rcp64(D * D);
    xb.qg += f2b * (c.dts * c.cgaci * sd);
    ib.den += f2b * (c.dts * c.cgaci) * x.qg * 0.5 * rcp64(sd);
  }
  const double sinkb = xb.qs - qi1b;
  double qib = qi1b, psacib = 0.0, psautb = 0.0;
  if (0.75 * x.qi < psaci + psaut) {
    qib += sinkb * 0.75;
  } else {
    psacib = sinkb;
    psautb = sinkb;
  }
  double tcb = 0.0, dib = 0.0;
  if (aut) {
    const double tmpb = psautb * dq, dqb = psautb * tmp;
    double qimb;
    if (quad) {
      const double dd = q_plus - qim;
      const double ddb = dqb * 0.5 * dd * rcp64(di);
      dib += ddb - dqb * dq * rcp64(di);
      qib += ddb;
      qimb = -ddb;
    } else {
      qib += dqb;
      qimb = -dqb;
    }
    ib.den -= qimb * qim * rcp64(den);
    if (!c.const_vi) tcb += tmpb * 0.025 * tmp;
  }
  if (di0 > QRMIN) {
    if (k.slope && k.dl > dm)
      ib.dl += dib;
    else if (!(QVMIN > k.hv * x.qi))
      qib += dib * k.hv;
  }
  if (x.qs > 1.0e-7) {
    const double D = 1.0 + f1;
    const double fb = psacib * x.qi * rcp64(D * D);
    qib += psacib * f1 * rcp64(D);
    ib.denfac += fb * c.dts * c.csaci * e;
    const double argb = fb * (c.dts * denfac * c.csaci) * e;
    tcb += argb * 0.05;
    const double qdb = argb * 0.8125 * rcp64(x.qs * den);
    xb.qs += qdb * den;
    ib.den += qdb * x.qs;
  }
  xb.qi = qib;
  xb.tz += tcb;
}

// rain -> snow (psacr) and rain -> graupel (pgfr), qr > 1e-7 below tice
SH_FN bool blk_c2(const IcK& k, IcX& x) {
  if (!((x.qr > 1.0e-7) && (x.tz - k.tice < 0.0))) return false;
  const Coef<double>& c = k.c;
  const double den = k.in.den;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double icpk = lhi / x.cvm;
  const double tc = x.tz - k.tice;
  double psacr = 0.0;
  if (x.qs > 1.0e-7)
    psacr = c.dts * acr3d<double>(k.in.vts, k.in.vtr, x.qr, x.qs, c.csacr, ck(c, T_ACCO + 1), ck(c, T_ACCO + 5),
                                  ck(c, T_ACCO + 9), den);
  const double pgfr = c.dts * ck(c, T_CGFR + 0) / den * (exp(-ck(c, T_CGFR + 1) * tc) - 1.0) *
                      pw_el<double>(x.qr * den, 1.75);
  const double sink = psacr + pgfr;
  const double factor = mmin(sink, mmin(x.qr, (-(x.tz - k.tice)) / icpk)) / mmax(sink, QRMIN);
  const double ps = factor * psacr, pg = factor * pgfr;
  const double sk = ps + pg;
  x.qr = x.qr - sk;
  x.qs = x.qs + ps;
  x.qg = x.qg + pg;
  x.q_liq = x.q_liq - sk;
  x.q_sol = x.q_sol + sk;
  x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
  x.tz = x.tz + (sk * lhi) / x.cvm;
  return true;
}

SH_FN void blk_c2_rev(const IcK& k, const IcX& x, IcX& xb, IcInB& ib) {
  IcX y = x;
  if (!blk_c2(k, y)) return;
  const Coef<double>& c = k.c;
  const double den = k.in.den;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double icpk = lhi / x.cvm;
  const double tc = x.tz - k.tice;
  double psacr = 0.0;
  if (x.qs > 1.0e-7)
    psacr = c.dts * acr3d<double>(k.in.vts, k.in.vtr, x.qr, x.qs, c.csacr, ck(c, T_ACCO + 1), ck(c, T_ACCO + 5),
                                  ck(c, T_ACCO + 9), den);
  const double g0 = c.dts * ck(c, T_CGFR + 0) / den;
  const double ex = exp(-ck(c, T_CGFR + 1) * tc);
  const double pq = pw_el<double>(x.qr * den, 1.75);
  const double pgfr = g0 * (ex - 1.0) * pq;
  const double sink = psacr + pgfr;
  const double lim = (-(x.tz - k.tice)) / icpk;
  const double m1 = mmin(x.qr, lim);
  const double d2 = mmax(sink, QRMIN);
  const double factor = mmin(sink, m1) / d2;
  const double ps = factor * psacr, pg = factor * pgfr;
  const double sk = ps + pg;
  double lhib = 0.0, icpkb = 0.0;
  double skb = heat_rev(c, 1.0, sk, lhi, y.cvm, xb, lhib);
  skb += xb.q_sol - xb.q_liq - xb.qr;
  const double psb = xb.qs + skb, pgb = xb.qg + skb;
  const double factorb = psb * psacr + pgb * pgfr;
  double psacrb = psb * factor, pgfrb = pgb * factor;
  const double numb = factorb * rcp64(d2), d2b = -factorb * factor * rcp64(d2);
  double sinkb = (sink > QRMIN) ? d2b : 0.0;
  double m1b = 0.0;
  if (sink < m1)
    sinkb += numb;
  else
    m1b = numb;
  if (x.qr < lim) {
    xb.qr += m1b;
  } else {
    xb.tz -= m1b * rcp64(icpk);
    icpkb -= m1b * lim * rcp64(icpk);
  }
  psacrb += sinkb;
  pgfrb += sinkb;
  ib.den -= pgfrb * (ex - 1.0) * pq * g0 * rcp64(den);
  const double tcb = pgfrb * g0 * pq * ex * (-ck(c, T_CGFR + 1));
  const double qrdb = pgfrb * g0 * (ex - 1.0) * 1.75 * pq * rcp64(x.qr * den);
  xb.qr += qrdb * den;
  ib.den += qrdb * x.qr;
  if (x.qs > 1.0e-7)
    acr3d_rev(k.in.vts, k.in.vtr, x.qr, x.qs, c.csacr, ck(c, T_ACCO + 1), ck(c, T_ACCO + 5), ck(c, T_ACCO + 9), den,
              psacrb * c.dts, ib.vts, ib.vtr, xb.qr, xb.qs, ib.den);
  xb.tz += tcb;
  lhib += icpkb * rcp64(x.cvm);
  xb.cvm -= icpkb * icpk * rcp64(x.cvm);
  xb.tz += lhib * DC_ICE;
}
// snow -> graupel (accretion pgacs, autoconversion psaut), qs > 1e-7
SH_FN bool blk_c3(const IcK& k, IcX& x) {
  if (!(x.qs > 1.0e-7)) return false;
  const Coef<double>& c = k.c;
  const double den = k.in.den;
  double sink = 0.0;
  if (x.qg > QRMIN)
    sink = c.dts * acr3d<double>(k.in.vtg, k.in.vts, x.qs, x.qg, c.cgacs, ck(c, T_ACCO + 3), ck(c, T_ACCO + 7),
                                 ck(c, T_ACCO + 11), den);
  const double qsm = c.qs0_crt / den;
  if (x.qs > qsm) {
    const double factor = c.dts * 1.0e-3 * exp(0.09 * (x.tz - k.tice));
    sink = sink + factor / (1.0 + factor) * (x.qs - qsm);
  }
  const double sk = mmin(x.qs, sink);
  x.qs = x.qs - sk;
  x.qg = x.qg + sk;
  return true;
}

SH_FN void blk_c3_rev(const IcK& k, const IcX& x, IcX& xb, IcInB& ib) {
  if (!(x.qs > 1.0e-7)) return;
  const Coef<double>& c = k.c;
  const double den = k.in.den;
  double sink = 0.0;
  if (x.qg > QRMIN)
    sink = c.dts * acr3d<double>(k.in.vtg, k.in.vts, x.qs, x.qg, c.cgacs, ck(c, T_ACCO + 3), ck(c, T_ACCO + 7),
                                 ck(c, T_ACCO + 11), den);
  const double qsm = c.qs0_crt / den;
  const bool aut = x.qs > qsm;
  double f = 0.0;
  if (aut) {
    f = c.dts * 1.0e-3 * exp(0.09 * (x.tz - k.tice));
    sink = sink + f / (1.0 + f) * (x.qs - qsm);
  }
  const double skb = xb.qg - xb.qs;
  double sinkb = 0.0;
  if (x.qs < sink)
    xb.qs += skb;
  else
    sinkb = skb;
  if (aut) {
    const double D = 1.0 + f;
    const double t = f / D;
    const double fb = sinkb * (x.qs - qsm) * rcp64(D * D);
    xb.qs += sinkb * t;
    ib.den += sinkb * t * qsm * rcp64(den);
    xb.tz += fb * f * 0.09;
  }
  if (x.qg > QRMIN)
    acr3d_rev(k.in.vtg, k.in.vts, x.qs, x.qg, c.cgacs, ck(c, T_ACCO + 3), ck(c, T_ACCO + 7), ck(c, T_ACCO + 11), den,
              sinkb * c.dts, ib.vtg, ib.vts, xb.qs, xb.qg, ib.den);
}

// cloud water and rain -> graupel (pgacw, pgacr), qg > 1e-7 below tice0
SH_FN bool blk_c4(const IcK& k, IcX& x) {
  if (!((x.qg > 1.0e-7) && (x.tz < k.tice0))) return false;
  const Coef<double>& c = k.c;
  const double den = k.in.den;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double icpk = lhi / x.cvm;
  double pgacw = 0.0, pgacr = 0.0;
  if (x.ql > 1.0e-6) {
    const double qden = x.qg * den;
    const double factor = c.dts * c.cgacw * qden / sqrt(den * sqrt(sqrt(qden)));
    pgacw = factor / (1.0 + factor) * x.ql;
  }
  if (x.qr > 1.0e-6)
    pgacr = mmin(c.dts * acr3d<double>(k.in.vtg, k.in.vtr, x.qr, x.qg, c.cgacr, ck(c, T_ACCO + 2),
                                       ck(c, T_ACCO + 6), ck(c, T_ACCO + 10), den),
                 x.qr);
  const double sink = pgacr + pgacw;
  const double factor = mmin(sink, dim(k.tice, x.tz) / icpk) / mmax(sink, QRMIN);
  const double pr = factor * pgacr, pl = factor * pgacw;
  const double sk = pr + pl;
  x.qg = x.qg + sk;
  x.qr = x.qr - pr;
  x.ql = x.ql - pl;
  x.q_liq = x.q_liq - sk;
  x.q_sol = x.q_sol + sk;
  x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
  x.tz = x.tz + (sk * lhi) / x.cvm;
  return true;
}

SH_FN void blk_c4_rev(const IcK& k, const IcX& x, IcX& xb, IcInB& ib) {
  IcX y = x;
  if (!blk_c4(k, y)) return;
  const Coef<double>& c = k.c;
  const double den = k.in.den;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double icpk = lhi / x.cvm;
  const double qden = x.qg * den;
  const double sq2 = sqrt(qden), sq4 = sqrt(sq2), dn = sqrt(den * sq4);
  const double fw = c.dts * c.cgacw * qden / dn;
  double pgacw = 0.0, pgacr = 0.0, a = 0.0;
  if (x.ql > 1.0e-6) pgacw = fw / (1.0 + fw) * x.ql;
  if (x.qr > 1.0e-6) {
    a = c.dts * acr3d<double>(k.in.vtg, k.in.vtr, x.qr, x.qg, c.cgacr, ck(c, T_ACCO + 2), ck(c, T_ACCO + 6),
                              ck(c, T_ACCO + 10), den);
    pgacr = mmin(a, x.qr);
  }
  const double sink = pgacr + pgacw;
  const double lim = dim(k.tice, x.tz) / icpk;
  const double d2 = mmax(sink, QRMIN);
  const double factor = mmin(sink, lim) / d2;
  const double pr = factor * pgacr, pl = factor * pgacw;
  const double sk = pr + pl;
  double lhib = 0.0, icpkb = 0.0;
  double skb = heat_rev(c, 1.0, sk, lhi, y.cvm, xb, lhib);
  skb += xb.q_sol - xb.q_liq + xb.qg;
  const double prb = skb - xb.qr, plb = skb - xb.ql;
  const double factorb = prb * pgacr + plb * pgacw;
  double pgacrb = prb * factor, pgacwb = plb * factor;
  const double numb = factorb * rcp64(d2), d2b = -factorb * factor * rcp64(d2);
  double sinkb = (sink > QRMIN) ? d2b : 0.0;
  if (sink < lim) {
    sinkb += numb;
  } else {
    if (k.tice - x.tz > 0.0) xb.tz -= numb * rcp64(icpk);
    icpkb -= numb * lim * rcp64(icpk);
  }
  pgacrb += sinkb;
  pgacwb += sinkb;
  if (x.qr > 1.0e-6) {
    if (a < x.qr)
      acr3d_rev(k.in.vtg, k.in.vtr, x.qr, x.qg, c.cgacr, ck(c, T_ACCO + 2), ck(c, T_ACCO + 6), ck(c, T_ACCO + 10),
                den, pgacrb * c.dts, ib.vtg, ib.vtr, xb.qr, xb.qg, ib.den);
    else
      xb.qr += pgacrb;
  }
  if (x.ql > 1.0e-6) {
    const double D = 1.0 + fw;
    const double fb = pgacwb * x.ql * rcp64(D * D);
    xb.ql += pgacwb * fw * rcp64(D);
    double qdenb = fb * (c.dts * c.cgacw) * rcp64(dn);
    const double dnb = -fb * fw * rcp64(dn);
    ib.den += dnb * 0.5 * rcp64(dn) * sq4;
    const double sq4b = dnb * 0.5 * rcp64(dn) * den;
    qdenb += sq4b * 0.5 * rcp64(sq4) * 0.5 * rcp64(sq2);
    xb.qg += qdenb * den;
    ib.den += qdenb * x.qg;
  }
  lhib += icpkb * rcp64(x.cvm);
  xb.cvm -= icpkb * icpk * rcp64(x.cvm);
  xb.tz += lhib * DC_ICE;
}

// -------------------------------------------------------- subgrid_z_proc
// q_liq, q_sol, cvm restart from the cell
SH_FN void blk_sg_reset(const IcK& k, IcX& x) {
  x.q_liq = x.ql + x.qr;
  x.q_sol = x.qi + x.qs + x.qg;
  x.cvm = ic_cvm(k.c, x.qv, x.q_liq, x.q_sol);
}
SH_FN void blk_sg_reset_rev(const IcK& k, IcX& xb) {
  const double cb = xb.cvm;
  xb.qv += cb * k.c.c_vap_d;
  const double lb = xb.q_liq + cb * C_LIQ, sb = xb.q_sol + cb * C_ICE;
  xb.ql += lb;
  xb.qr += lb;
  xb.qi += sb;
  xb.qs += sb;
  xb.qg += sb;
  xb.q_liq = xb.q_sol = xb.cvm = 0.0;
}

// below T_MIN: vapour above 1e-7 deposits as ice, then the step ends
SH_FN bool blk_tmin(const IcK& k, IcX& x) {
  if (!(x.tz < T_MIN)) return false;
  const Coef<double>& c = k.c;
  const double lhl = c.lv00_d + c.d0_vap_d * x.tz;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double sink = dim(1.0e-7, x.qv);
  x.qv = x.qv - sink;
  x.qi = x.qi + sink;
  x.q_sol = x.q_sol + sink;
  x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
  x.tz = x.tz + (sink * (lhl + lhi)) / x.cvm;
  if (!c.do_qa) x.qa = x.qa + 1.0;
  return true;
}
SH_FN void blk_tmin_rev(const IcK& k, const IcX& x, IcX& xb) {
  IcX y = x;
  if (!blk_tmin(k, y)) return;
  const Coef<double>& c = k.c;
  const double L = (c.lv00_d + c.d0_vap_d * x.tz) + (LI00 + DC_ICE * x.tz);
  const double sink = dim(1.0e-7, x.qv);
  double Lb = 0.0;
  double sinkb = heat_rev(c, 1.0, sink, L, y.cvm, xb, Lb);
  sinkb += xb.q_sol + xb.qi - xb.qv;
  if (1.0e-7 - x.qv > 0.0) xb.qv -= sinkb;
  xb.tz += Lb * (c.d0_vap_d + DC_ICE);
}

// instant evaporation / sublimation of all clouds where rh < rh_adj, after
// which the step ends; otherwise no change
SH_FN bool blk_coll(const IcK& k, IcX& x) {
  const Coef<double>& c = k.c;
  const double lhl = c.lv00_d + c.d0_vap_d * x.tz;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double qpz = x.qv + x.ql + x.qi;
  const double dtin =
      (lhl * (x.ql + x.qi) + lhi * x.qi) / (c.c_air_d + qpz * c.c_vap_d + x.qr * C_LIQ + (x.qs + x.qg) * C_ICE);
  const double tin = x.tz - dtin;
  if (tin > c.t_sub_d + 6.0) {
    const double rh = qpz / iqs1<double>(tin, k.in.den);
    if (rh < k.rh_adj) {
      x.tz = tin;
      x.qv = qpz;
      x.ql = 0.0;
      x.qi = 0.0;
      return true;
    }
  }
  return false;
}
SH_FN void blk_coll_rev(const IcK& k, const IcX& x, IcX& xb) {
  const Coef<double>& c = k.c;
  const double lhl = c.lv00_d + c.d0_vap_d * x.tz;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double qpz = x.qv + x.ql + x.qi;
  const double N = lhl * (x.ql + x.qi) + lhi * x.qi;
  const double Dd = c.c_air_d + qpz * c.c_vap_d + x.qr * C_LIQ + (x.qs + x.qg) * C_ICE;
  const double dtin = N / Dd;
  const double tinb = xb.tz;
  double qpzb = xb.qv;
  const double Nb = -tinb * rcp64(Dd), Ddb = tinb * dtin * rcp64(Dd);
  xb.tz = tinb + (Nb * (x.ql + x.qi)) * c.d0_vap_d + (Nb * x.qi) * DC_ICE;
  qpzb += Ddb * c.c_vap_d;
  xb.qr += Ddb * C_LIQ;
  xb.qs += Ddb * C_ICE;
  xb.qg += Ddb * C_ICE;
  xb.qv = qpzb;
  xb.ql = qpzb + Nb * lhl;
  xb.qi = qpzb + Nb * lhl + Nb * lhi;
}

// cloud water <-> vapour adjustment
SH_FN void blk_evap(const IcK& k, IcX& x) {
  const Coef<double>& c = k.c;
  const IcHeat h = ic_heat(c, x);
  const double tcp3 = h.lcpk + h.icpk * mmin(1.0, dim(k.tice, x.tz) / (c.tice - c.t_wfr));
  double dwsdt;
  const double qsw = wqs2<double>(x.tz, k.in.den, dwsdt);
  const double dq0 = qsw - x.qv;
  double evap;
  if (dq0 > 0.0) {
    const double factor = mmin(1.0, c.fac_l2v * ((10.0 * dq0) / qsw));
    evap = mmin(x.ql, (factor * dq0) / (1.0 + tcp3 * dwsdt));
  } else {
    evap = dq0 / (1.0 + tcp3 * dwsdt);
  }
  x.qv = x.qv + evap;
  x.ql = x.ql - evap;
  x.q_liq = x.q_liq - evap;
  x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
  x.tz = x.tz - (evap * h.lhl) / x.cvm;
}
SH_FN void blk_evap_rev(const IcK& k, const IcX& x, IcX& xb, IcInB& ib) {
  const Coef<double>& c = k.c;
  IcX y = x;
  blk_evap(k, y);
  const IcHeat h = ic_heat(c, x);
  const double ratio = dim(k.tice, x.tz) / (c.tice - c.t_wfr);
  const double mn = mmin(1.0, ratio);
  const double tcp3 = h.lcpk + h.icpk * mn;
  double dwsdt;
  const double qsw = wqs2<double>(x.tz, k.in.den, dwsdt);
  const double dq0 = qsw - x.qv;
  const double Dn = 1.0 + tcp3 * dwsdt;
  double evap, e2 = 0.0, factor = 0.0, fm = 0.0, r10 = 0.0;
  if (dq0 > 0.0) {
    r10 = (10.0 * dq0) / qsw;
    fm = c.fac_l2v * r10;
    factor = mmin(1.0, fm);
    e2 = (factor * dq0) / Dn;
    evap = mmin(x.ql, e2);
  } else {
    evap = dq0 / Dn;
  }
  double lhlb = 0.0;
  double evapb = heat_rev(c, -1.0, evap, h.lhl, y.cvm, xb, lhlb);
  evapb += xb.qv - xb.ql - xb.q_liq;
  double dq0b = 0.0, Dnb = 0.0, qswb = 0.0;
  if (dq0 > 0.0) {
    double e2b = 0.0;
    if (x.ql < e2)
      xb.ql += evapb;
    else
      e2b = evapb;
    const double fb = e2b * dq0 * rcp64(Dn);
    dq0b += e2b * factor * rcp64(Dn);
    Dnb -= e2b * e2 * rcp64(Dn);
    if (!(1.0 < fm)) {
      const double r10b = fb * c.fac_l2v;
      dq0b += r10b * 10.0 * rcp64(qsw);
      qswb -= r10b * r10 * rcp64(qsw);
    }
  } else {
    dq0b += evapb * rcp64(Dn);
    Dnb -= evapb * evap * rcp64(Dn);
  }
  const double tcp3b = Dnb * dwsdt, dwsdtb = Dnb * tcp3;
  qswb += dq0b;
  xb.qv -= dq0b;
  sat_rev(false, x.tz, k.in.den, qsw, dwsdt, qswb, dwsdtb, xb.tz, ib.den);
  if (!(1.0 < ratio) && (k.tice - x.tz > 0.0)) xb.tz -= tcp3b * h.icpk * rcp64(c.tice - c.t_wfr);
  ic_heat_rev(c, x, h, lhlb, 0.0, tcp3b, tcp3b * mn, xb);
}

// complete freezing of cloud water below t_wfr
SH_FN bool blk_frz(const IcK& k, IcX& x) {
  const Coef<double>& c = k.c;
  const double dtmp = c.t_wfr_d - x.tz;
  if (!((dtmp > 0.0) && (x.ql > QCMIN))) return false;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double icpk = lhi / x.cvm;
  const double sink = mmin(x.ql, mmin(x.ql * dtmp * 0.125, dtmp / icpk));
  x.ql = x.ql - sink;
  x.qi = x.qi + sink;
  x.q_liq = x.q_liq - sink;
  x.q_sol = x.q_sol + sink;
  x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
  x.tz = x.tz + (sink * lhi) / x.cvm;
  return true;
}
SH_FN void blk_frz_rev(const IcK& k, const IcX& x, IcX& xb) {
  IcX y = x;
  if (!blk_frz(k, y)) return;
  const Coef<double>& c = k.c;
  const double dtmp = c.t_wfr_d - x.tz;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double icpk = lhi / x.cvm;
  const double a1 = x.ql * dtmp * 0.125, a2 = dtmp / icpk;
  const double m = mmin(a1, a2);
  const double sink = mmin(x.ql, m);
  double lhib = 0.0, icpkb = 0.0;
  double sinkb = heat_rev(c, 1.0, sink, lhi, y.cvm, xb, lhib);
  sinkb += xb.q_sol - xb.q_liq + xb.qi - xb.ql;
  double mb = 0.0;
  if (x.ql < m)
    xb.ql += sinkb;
  else
    mb = sinkb;
  double dtmpb = 0.0;
  if (a1 < a2) {
    xb.ql += mb * dtmp * 0.125;
    dtmpb += mb * x.ql * 0.125;
  } else {
    dtmpb += mb * rcp64(icpk);
    icpkb -= mb * a2 * rcp64(icpk);
  }
  xb.tz -= dtmpb;
  lhib += icpkb * rcp64(x.cvm);
  xb.cvm -= icpkb * icpk * rcp64(x.cvm);
  xb.tz += lhib * DC_ICE;
}

// Bigg freezing of cloud water below tice
SH_FN bool blk_bigg(const IcK& k, IcX& x) {
  const Coef<double>& c = k.c;
  if (c.fast_sat_adj) return false;
  const double tcd = k.tice - x.tz;
  if (!((x.ql > QRMIN) && (tcd > 0.0))) return false;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double icpk = lhi / x.cvm;
  const double sink = 3.3333e-10 * c.dts * (exp(0.66 * tcd) - 1.0) * k.in.den * x.ql * x.ql;
  const double sk = mmin(x.ql, mmin(tcd / icpk, sink));
  x.ql = x.ql - sk;
  x.qi = x.qi + sk;
  x.q_liq = x.q_liq - sk;
  x.q_sol = x.q_sol + sk;
  x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
  x.tz = x.tz + (sk * lhi) / x.cvm;
  return true;
}
SH_FN void blk_bigg_rev(const IcK& k, const IcX& x, IcX& xb, IcInB& ib) {
  IcX y = x;
  if (!blk_bigg(k, y)) return;
  const Coef<double>& c = k.c;
  const double den = k.in.den;
  const double tcd = k.tice - x.tz;
  const double lhi = LI00 + DC_ICE * x.tz;
  const double icpk = lhi / x.cvm;
  const double C0 = 3.3333e-10 * c.dts;
  const double ex = exp(0.66 * tcd);
  const double sink = C0 * (ex - 1.0) * den * x.ql * x.ql;
  const double a2 = tcd / icpk;
  const double m = mmin(a2, sink);
  const double sk = mmin(x.ql, m);
  double lhib = 0.0, icpkb = 0.0;
  double skb = heat_rev(c, 1.0, sk, lhi, y.cvm, xb, lhib);
  skb += xb.q_sol - xb.q_liq + xb.qi - xb.ql;
  double mb = 0.0;
  if (x.ql < m)
    xb.ql += skb;
  else
    mb = skb;
  double tcdb = 0.0;
  if (a2 < sink) {
    tcdb += mb * rcp64(icpk);
    icpkb -= mb * a2 * rcp64(icpk);
  } else {
    tcdb += mb * C0 * ex * 0.66 * den * x.ql * x.ql;
    ib.den += mb * C0 * (ex - 1.0) * x.ql * x.ql;
    xb.ql += mb * 2.0 * C0 * (ex - 1.0) * den * x.ql;
  }
  xb.tz -= tcdb;
  lhib += icpkb * rcp64(x.cvm);
  xb.cvm -= icpkb * icpk * rcp64(x.cvm);
  xb.tz += lhib * DC_ICE;
}
// sublimation / deposition of cloud ice below tice
SH_FN bool blk_idep(const IcK& k, IcX& x, double dt_pisub, double* amt = nullptr) {
  if (!(x.tz < k.tice)) return false;
  const Coef<double>& c = k.c;
  const double den = k.in.den;
  const IcHeat h = ic_heat(c, x);
  const double tcpk = h.lcpk + h.icpk;
  double dqsdt;
  const double qsi = iqs2<double>(x.tz, den, dqsdt);
  const double dq = x.qv - qsi;
  const double sink = dq / (1.0 + tcpk * dqsdt);
  double pidep = 0.0;
  if (x.qi > QRMIN)
    pidep = dt_pisub * dq * 349138.78 * pw_el<double>(x.qi * den, 0.875) /
            (qsi * den * LAT2 / (0.0243 * RVGAS * (x.tz * x.tz)) + 4.42478e4);
  double sk;
  if (dq > 0.0) {
    const double tmp = k.tice - x.tz;
    const double qi_crt = c.qi_gen * mmin(c.qi_lim, 0.1 * tmp) / den;
    sk = mmin(sink, mmin(mmax(qi_crt - x.qi, pidep), tmp / tcpk));
  } else {
    pidep = pidep * mmin(1.0, dim(x.tz, c.t_sub_d) * 0.2);
    sk = mmax(mmax(pidep, sink), -x.qi);
  }
  if (amt) *amt = sk;
  x.qv = x.qv - sk;
  x.qi = x.qi + sk;
  x.q_sol = x.q_sol + sk;
  x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
  x.tz = x.tz + (sk * (h.lhl + h.lhi)) / x.cvm;
  return true;
}
SH_FN void blk_idep_rev(const IcK& k, const IcX& x, double dt_pisub, IcX& xb, IcInB& ib) {
  IcX y = x;
  double sk;
  if (!blk_idep(k, y, dt_pisub, &sk)) return;
  const Coef<double>& c = k.c;
  const double den = k.in.den;
  const IcHeat h = ic_heat(c, x);
  const double tcpk = h.lcpk + h.icpk;
  double dqsdt;
  const double qsi = iqs2<double>(x.tz, den, dqsdt);
  const double dq = x.qv - qsi;
  const double Dn = 1.0 + tcpk * dqsdt;
  const double sink = dq / Dn;
  const bool hasqi = x.qi > QRMIN;
  const double T2 = x.tz * x.tz, K2 = 0.0243 * RVGAS;
  const double dd0 = qsi * den * LAT2 / (K2 * T2);
  const double dd = dd0 + 4.42478e4;
  const double pq = hasqi ? pw_el<double>(x.qi * den, 0.875) : 0.0;
  const double num = dt_pisub * dq * 349138.78 * pq;
  const double pidep = hasqi ? num / dd : 0.0;
  const double L = h.lhl + h.lhi;
  double Lb = 0.0;
  const double skb = heat_rev(c, 1.0, sk, L, y.cvm, xb, Lb) + xb.q_sol + xb.qi - xb.qv;
  double sinkb = 0.0, pidepb = 0.0, tcpkb = 0.0;
  if (dq > 0.0) {
    const double tmp = k.tice - x.tz;
    const double xq = 0.1 * tmp;
    const double mq = mmin(c.qi_lim, xq);
    const double qi_crt = c.qi_gen * mq / den;
    const double x1 = mmax(qi_crt - x.qi, pidep), x2 = tmp / tcpk;
    const double m = mmin(x1, x2);
    double mb = 0.0;
    if (sink < m)
      sinkb += skb;
    else
      mb = skb;
    double tmpb = 0.0;
    if (x1 < x2) {
      if (qi_crt - x.qi > pidep) {
        xb.qi -= mb;
        ib.qi_gen += mb * mq * rcp64(den);
        ib.den -= mb * qi_crt * rcp64(den);
        if (!(c.qi_lim < xq)) tmpb += mb * c.qi_gen * rcp64(den) * 0.1;
      } else {
        pidepb += mb;
      }
    } else {
      tmpb += mb * rcp64(tcpk);
      tcpkb -= mb * x2 * rcp64(tcpk);
    }
    xb.tz -= tmpb;
  } else {
    const double dt = dim(x.tz, c.t_sub_d);
    const double fm = mmin(1.0, dt * 0.2);
    const double p2 = pidep * fm;
    const double xm = mmax(p2, sink);
    double xmb = 0.0;
    if (xm > -x.qi)
      xmb = skb;
    else
      xb.qi -= skb;
    double p2b = 0.0;
    if (p2 > sink)
      p2b = xmb;
    else
      sinkb += xmb;
    pidepb += p2b * fm;
    if (!(1.0 < dt * 0.2) && (x.tz - c.t_sub_d > 0.0)) xb.tz += p2b * pidep * 0.2;
  }
  double dqb = 0.0, qsib = 0.0;
  if (hasqi) {
    const double numb = pidepb * rcp64(dd), ddb = -pidepb * pidep * rcp64(dd);
    dqb += numb * dt_pisub * 349138.78 * pq;
    const double pqb = numb * dt_pisub * dq * 349138.78;
    const double qidb = pqb * 0.875 * pq * rcp64(x.qi * den);
    xb.qi += qidb * den;
    ib.den += qidb * x.qi;
    qsib += ddb * den * LAT2 * rcp64(K2 * T2);
    ib.den += ddb * qsi * LAT2 * rcp64(K2 * T2);
    xb.tz -= ddb * 2.0 * dd0 * rcp64(x.tz);
  }
  dqb += sinkb * rcp64(Dn);
  const double Dnb = -sinkb * sink * rcp64(Dn);
  tcpkb += Dnb * dqsdt;
  const double dqsdtb = Dnb * tcpk;
  xb.qv += dqb;
  qsib -= dqb;
  sat_rev(true, x.tz, den, qsi, dqsdt, qsib, dqsdtb, xb.tz, ib.den);
  ic_heat_rev(c, x, h, Lb, Lb, tcpkb, tcpkb, xb);
}

// sublimation / deposition of snow
SH_FN bool blk_sdep(const IcK& k, IcX& x, double* amt = nullptr) {
  if (!(x.qs > QRMIN)) return false;
  const Coef<double>& c = k.c;
  const double den = k.in.den;
  const IcHeat h = ic_heat(c, x);
  const double tcpk = h.lcpk + h.icpk;
  double dqsdt;
  const double qsi = iqs2<double>(x.tz, den, dqsdt);
  const double qden = x.qs * den;
  const double tmp = pw_el<double>(qden, 0.65625);
  const double tsq = x.tz * x.tz;
  const double dsat = qsi - x.qv;
  const double dq = dsat / (1.0 + tcpk * dqsdt);
  double pssub = ck(c, T_CSSUB + 0) * tsq *
                 (ck(c, T_CSSUB + 1) * sqrt(qden) + ck(c, T_CSSUB + 2) * tmp * sqrt(k.in.denfac)) /
                 (ck(c, T_CSSUB + 3) * tsq + ck(c, T_CSSUB + 4) * qsi * den);
  pssub = dsat * c.dts * pssub;
  double ps;
  if (pssub > 0.0) {
    ps = mmin(pssub * mmin(1.0, dim(x.tz, c.t_sub_d) * 0.2), x.qs);
  } else {
    if (x.tz > k.tice)
      ps = 0.0;
    else
      ps = mmax(pssub, mmax(dq, (x.tz - k.tice) / tcpk));
  }
  if (amt) *amt = ps;
  x.qs = x.qs - ps;
  x.qv = x.qv + ps;
  x.q_sol = x.q_sol - ps;
  x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
  x.tz = x.tz - (ps * (h.lhl + h.lhi)) / x.cvm;
  return true;
}
SH_FN void blk_sdep_rev(const IcK& k, const IcX& x, IcX& xb, IcInB& ib) {
  IcX y = x;
  double ps;
  if (!blk_sdep(k, y, &ps)) return;
  const Coef<double>& c = k.c;
  const double den = k.in.den, denfac = k.in.denfac;
  const IcHeat h = ic_heat(c, x);
  const double tcpk = h.lcpk + h.icpk;
  double dqsdt;
  const double qsi = iqs2<double>(x.tz, den, dqsdt);
  const double qden = x.qs * den;
  const double tmp = pw_el<double>(qden, 0.65625);
  const double tsq = x.tz * x.tz;
  const double dsat = qsi - x.qv;
  const double Dn = 1.0 + tcpk * dqsdt;
  const double dq = dsat / Dn;
  const double sqd = sqrt(qden), sdf = sqrt(denfac);
  const double Bq = ck(c, T_CSSUB + 1) * sqd + ck(c, T_CSSUB + 2) * tmp * sdf;
  const double num = ck(c, T_CSSUB + 0) * tsq * Bq;
  const double dd = ck(c, T_CSSUB + 3) * tsq + ck(c, T_CSSUB + 4) * qsi * den;
  const double pss0 = num / dd;
  const double pssub = dsat * c.dts * pss0;
  const double L = h.lhl + h.lhi;
  double Lb = 0.0;
  double psb = heat_rev(c, -1.0, ps, L, y.cvm, xb, Lb);
  psb += -xb.qs + xb.qv - xb.q_sol;
  double pssubb = 0.0, dqb = 0.0, tcpkb = 0.0;
  if (pssub > 0.0) {
    const double dt = dim(x.tz, c.t_sub_d);
    const double fm = mmin(1.0, dt * 0.2);
    const double a = pssub * fm;
    double ab = 0.0;
    if (a < x.qs)
      ab = psb;
    else
      xb.qs += psb;
    pssubb += ab * fm;
    if (!(1.0 < dt * 0.2) && (x.tz - c.t_sub_d > 0.0)) xb.tz += ab * pssub * 0.2;
  } else if (!(x.tz > k.tice)) {
    const double x3 = (x.tz - k.tice) / tcpk;
    const double m = mmax(dq, x3);
    double mb = 0.0;
    if (pssub > m)
      pssubb += psb;
    else
      mb = psb;
    if (dq > x3) {
      dqb += mb;
    } else {
      xb.tz += mb * rcp64(tcpk);
      tcpkb -= mb * x3 * rcp64(tcpk);
    }
  }
  double dsatb = pssubb * c.dts * pss0;
  const double pss0b = pssubb * (dsat * c.dts);
  const double numb = pss0b * rcp64(dd), ddb = -pss0b * pss0 * rcp64(dd);
  double tsqb = numb * ck(c, T_CSSUB + 0) * Bq + ddb * ck(c, T_CSSUB + 3);
  const double Bqb = numb * ck(c, T_CSSUB + 0) * tsq;
  double qdenb = Bqb * ck(c, T_CSSUB + 1) * 0.5 * rcp64(sqd);
  const double tmpb = Bqb * ck(c, T_CSSUB + 2) * sdf;
  ib.denfac += Bqb * ck(c, T_CSSUB + 2) * tmp * 0.5 * rcp64(sdf);
  double qsib = ddb * ck(c, T_CSSUB + 4) * den;
  ib.den += ddb * ck(c, T_CSSUB + 4) * qsi;
  dsatb += dqb * rcp64(Dn);
  const double Dnb = -dqb * dq * rcp64(Dn);
  tcpkb += Dnb * dqsdt;
  const double dqsdtb = Dnb * tcpk;
  qsib += dsatb;
  xb.qv -= dsatb;
  xb.tz += tsqb * 2.0 * x.tz;
  qdenb += tmpb * 0.65625 * tmp * rcp64(qden);
  xb.qs += qdenb * den;
  ib.den += qdenb * x.qs;
  sat_rev(true, x.tz, den, qsi, dqsdt, qsib, dqsdtb, xb.tz, ib.den);
  ic_heat_rev(c, x, h, Lb, Lb, tcpkb, tcpkb, xb);
}

// two-way graupel sublimation / deposition
SH_FN bool blk_gdep(const IcK& k, IcX& x, double* amt = nullptr) {
  if (!(x.qg > QRMIN)) return false;
  const Coef<double>& c = k.c;
  const double den = k.in.den;
  const IcHeat h = ic_heat(c, x);
  const double tcpk = h.lcpk + h.icpk;
  double dqsdt;
  const double qsi = iqs2<double>(x.tz, den, dqsdt);
  const double dq = (x.qv - qsi) / (1.0 + tcpk * dqsdt);
  const double pgsub = (x.qv / qsi - 1.0) * x.qg;
  double pg;
  if (pgsub > 0.0) {
    if (x.tz > k.tice)
      pg = 0.0;
    else
      pg = mmin(mmin(c.fac_v2g * pgsub, 0.2 * dq), mmin(x.ql + x.qr, (k.tice - x.tz) / tcpk));
  } else {
    pg = mmax(c.fac_g2v * pgsub, dq) * mmin(1.0, dim(x.tz, c.t_sub_d) * 0.1);
  }
  if (amt) *amt = pg;
  x.qg = x.qg + pg;
  x.qv = x.qv - pg;
  x.q_sol = x.q_sol + pg;
  x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
  x.tz = x.tz + (pg * (h.lhl + h.lhi)) / x.cvm;
  return true;
}
// icpk_ext: the cotangent of lhi / cvm at this block's entry from its use in
// the cloud fraction
SH_FN void blk_gdep_rev(const IcK& k, const IcX& x, double icpk_ext, IcX& xb, IcInB& ib) {
  const Coef<double>& c = k.c;
  const IcHeat h = ic_heat(c, x);
  IcX y = x;
  double pg;
  if (!blk_gdep(k, y, &pg)) {
    ic_heat_rev(c, x, h, 0.0, 0.0, 0.0, icpk_ext, xb);
    return;
  }
  const double den = k.in.den;
  const double tcpk = h.lcpk + h.icpk;
  double dqsdt;
  const double qsi = iqs2<double>(x.tz, den, dqsdt);
  const double Dn = 1.0 + tcpk * dqsdt;
  const double dq = (x.qv - qsi) / Dn;
  const double rq = x.qv / qsi;
  const double pgsub = (rq - 1.0) * x.qg;
  const double L = h.lhl + h.lhi;
  double Lb = 0.0;
  double pgb = heat_rev(c, 1.0, pg, L, y.cvm, xb, Lb);
  pgb += xb.qg - xb.qv + xb.q_sol;
  double pgsubb = 0.0, dqb = 0.0, tcpkb = 0.0;
  if (pgsub > 0.0) {
    if (!(x.tz > k.tice)) {
      const double a = c.fac_v2g * pgsub, b = 0.2 * dq;
      const double m1 = mmin(a, b);
      const double cc = x.ql + x.qr, d = (k.tice - x.tz) / tcpk;
      const double m2 = mmin(cc, d);
      double m1b = 0.0, m2b = 0.0;
      if (m1 < m2)
        m1b = pgb;
      else
        m2b = pgb;
      if (a < b)
        pgsubb += m1b * c.fac_v2g;
      else
        dqb += m1b * 0.2;
      if (cc < d) {
        xb.ql += m2b;
        xb.qr += m2b;
      } else {
        xb.tz -= m2b * rcp64(tcpk);
        tcpkb -= m2b * d * rcp64(tcpk);
      }
    }
  } else {
    const double a = c.fac_g2v * pgsub;
    const double mx = mmax(a, dq);
    const double dt = dim(x.tz, c.t_sub_d);
    const double fm = mmin(1.0, dt * 0.1);
    const double mxb = pgb * fm;
    if (a > dq)
      pgsubb += mxb * c.fac_g2v;
    else
      dqb += mxb;
    if (!(1.0 < dt * 0.1) && (x.tz - c.t_sub_d > 0.0)) xb.tz += pgb * mx * 0.1;
  }
  const double rqb = pgsubb * x.qg;
  xb.qg += pgsubb * (rq - 1.0);
  xb.qv += rqb * rcp64(qsi);
  double qsib = -rqb * rq * rcp64(qsi);
  const double nb = dqb * rcp64(Dn), Dnb = -dqb * dq * rcp64(Dn);
  xb.qv += nb;
  qsib -= nb;
  tcpkb += Dnb * dqsdt;
  const double dqsdtb = Dnb * tcpk;
  sat_rev(true, x.tz, den, qsi, dqsdt, qsib, dqsdtb, xb.tz, ib.den);
  ic_heat_rev(c, x, h, Lb, Lb, tcpkb, tcpkb + icpk_ext, xb);
}

// minimum evaporation of rain in dry air
SH_FN bool blk_revp(const IcK& k, IcX& x) {
  if (!(x.qr > QCMIN)) return false;
  const Coef<double>& c = k.c;
  const double lhl = c.lv00_d + c.d0_vap_d * x.tz;
  const double lcpk = lhl / x.cvm;
  double dqsdt;
  const double qsw2 = wqs2<double>(x.tz, k.in.den, dqsdt);
  const double sink = mmin(x.qr, dim(k.rh_rain * qsw2, x.qv) / (1.0 + lcpk * dqsdt));
  x.qv = x.qv + sink;
  x.qr = x.qr - sink;
  x.q_liq = x.q_liq - sink;
  x.cvm = ic_cvm(c, x.qv, x.q_liq, x.q_sol);
  x.tz = x.tz - (sink * lhl) / x.cvm;
  return true;
}
SH_FN void blk_revp_rev(const IcK& k, const IcX& x, IcX& xb, IcInB& ib) {
  IcX y = x;
  if (!blk_revp(k, y)) return;
  const Coef<double>& c = k.c;
  const double lhl = c.lv00_d + c.d0_vap_d * x.tz;
  const double lcpk = lhl / x.cvm;
  double dqsdt;
  const double qsw2 = wqs2<double>(x.tz, k.in.den, dqsdt);
  const double a = k.rh_rain * qsw2;
  const double Dn = 1.0 + lcpk * dqsdt;
  const double b = dim(a, x.qv) / Dn;
  const double sink = mmin(x.qr, b);
  double lhlb = 0.0;
  double sinkb = heat_rev(c, -1.0, sink, lhl, y.cvm, xb, lhlb);
  sinkb += xb.qv - xb.qr - xb.q_liq;
  double bb = 0.0;
  if (x.qr < b)
    xb.qr += sinkb;
  else
    bb = sinkb;
  const double ddb = bb * rcp64(Dn), Dnb = -bb * b * rcp64(Dn);
  double qswb = 0.0;
  if (a - x.qv > 0.0) {
    qswb = ddb * k.rh_rain;
    xb.qv -= ddb;
  }
  const double lcpkb = Dnb * dqsdt, dqsdtb = Dnb * lcpk;
  sat_rev(false, x.tz, k.in.den, qsw2, dqsdt, qswb, dqsdtb, xb.tz, ib.den);
  lhlb += lcpkb * rcp64(x.cvm);
  xb.cvm -= lcpkb * lcpk * rcp64(x.cvm);
  xb.tz += lhlb * c.d0_vap_d;
}

// cloud fraction from the subgrid distribution of total water (do_qa off);
// icpk7 is lhi / cvm of the graupel block's entry
SH_FN bool blk_cf(const IcK& k, IcX& x, double icpk7) {
  const Coef<double>& c = k.c;
  if (c.do_qa) return false;
  const double den = k.in.den;
  const double lhl = c.lv00_d + c.d0_vap_d * x.tz;
  const double cvm = c.c_air_d + (x.qv + x.q_liq + x.q_sol) * c.c_vap_d;
  const double lcpk = lhl / cvm;
  const double q_sol2 = c.rad_snow ? x.qi + x.qs : x.qi;
  const double q_liq2 = c.rad_rain ? x.ql + x.qr : x.ql;
  const double q_cond = q_liq2 + q_sol2;
  const double qpz2 = x.qv + q_cond;
  const double tin2 = x.tz - (lcpk * q_cond + icpk7 * q_sol2);
  double qstar;
  if (tin2 <= c.t_wfr_d) {
    qstar = iqs1<double>(tin2, den);
  } else if (tin2 >= k.tice) {
    qstar = wqs1<double>(tin2, den);
  } else {
    const double qsi = iqs1<double>(tin2, den);
    const double qsw3 = wqs1<double>(tin2, den);
    const double rqi = (q_cond > 3.0e-6) ? q_sol2 / q_cond : (k.tice - tin2) / (k.tice - c.t_wfr_d);
    qstar = rqi * qsi + (1.0 - rqi) * qsw3;
  }
  if (qpz2 > QRMIN) {
    const double dq = mmax(QCMIN, k.hv * qpz2);
    const double q_plus = qpz2 + dq;
    const double q_minus = qpz2 - dq;
    if (qstar < q_minus)
      x.qa = x.qa + 1.0;
    else if ((qstar < q_plus) && (q_cond > c.qc_crt_d))
      x.qa = x.qa + (q_plus - qstar) / (dq + dq);
  }
  x.cvm = cvm;
  return true;
}
// returns the cotangent of icpk7
SH_FN double blk_cf_rev(const IcK& k, const IcX& x, double icpk7, IcX& xb, IcInB& ib) {
  const Coef<double>& c = k.c;
  if (c.do_qa) return 0.0;
  const double den = k.in.den;
  const double lhl = c.lv00_d + c.d0_vap_d * x.tz;
  const double cvm = c.c_air_d + (x.qv + x.q_liq + x.q_sol) * c.c_vap_d;
  const double lcpk = lhl / cvm;
  const double q_sol2 = c.rad_snow ? x.qi + x.qs : x.qi;
  const double q_liq2 = c.rad_rain ? x.ql + x.qr : x.ql;
  const double q_cond = q_liq2 + q_sol2;
  const double qpz2 = x.qv + q_cond;
  const double tin2 = x.tz - (lcpk * q_cond + icpk7 * q_sol2);
  if (!(qpz2 > QRMIN)) return 0.0;
  double qstar, qsi = 0.0, qsw3 = 0.0, rqi = 0.0;
  const int br = (tin2 <= c.t_wfr_d) ? 0 : (tin2 >= k.tice) ? 1 : 2;
  if (br == 0) {
    qstar = iqs1<double>(tin2, den);
  } else if (br == 1) {
    qstar = wqs1<double>(tin2, den);
  } else {
    qsi = iqs1<double>(tin2, den);
    qsw3 = wqs1<double>(tin2, den);
    rqi = (q_cond > 3.0e-6) ? q_sol2 / q_cond : (k.tice - tin2) / (k.tice - c.t_wfr_d);
    qstar = rqi * qsi + (1.0 - rqi) * qsw3;
  }
  const double dq = mmax(QCMIN, k.hv * qpz2);
  const double q_plus = qpz2 + dq, q_minus = qpz2 - dq;
  if (qstar < q_minus) return 0.0;
  if (!((qstar < q_plus) && (q_cond > c.qc_crt_d))) return 0.0;
  const double d2 = dq + dq;
  const double f = (q_plus - qstar) / d2;
  const double nb = xb.qa * rcp64(d2), d2b = -xb.qa * f * rcp64(d2);
  const double qstarb = -nb;
  double qpz2b = nb;
  const double dqb = 2.0 * d2b + nb;
  if (!(QCMIN > k.hv * qpz2)) qpz2b += dqb * k.hv;
  double tin2b = 0.0, q_sol2b = 0.0, q_condb = 0.0;
  if (br == 0) {
    sat_rev(true, tin2, den, qstar, 0.0, qstarb, 0.0, tin2b, ib.den);
  } else if (br == 1) {
    sat_rev(false, tin2, den, qstar, 0.0, qstarb, 0.0, tin2b, ib.den);
  } else {
    const double rqib = qstarb * (qsi - qsw3);
    sat_rev(true, tin2, den, qsi, 0.0, qstarb * rqi, 0.0, tin2b, ib.den);
    sat_rev(false, tin2, den, qsw3, 0.0, qstarb * (1.0 - rqi), 0.0, tin2b, ib.den);
    if (q_cond > 3.0e-6) {
      q_sol2b += rqib * rcp64(q_cond);
      q_condb -= rqib * rqi * rcp64(q_cond);
    } else {
      tin2b -= rqib * rcp64(k.tice - c.t_wfr_d);
    }
  }
  xb.tz += tin2b;
  const double lcpkb = -tin2b * q_cond;
  q_condb -= tin2b * lcpk;
  const double icpk7b = -tin2b * q_sol2;
  q_sol2b -= tin2b * icpk7;
  xb.qv += qpz2b;
  q_condb += qpz2b;
  q_sol2b += q_condb;
  xb.ql += q_condb;
  if (c.rad_rain) xb.qr += q_condb;
  xb.qi += q_sol2b;
  if (c.rad_snow) xb.qs += q_sol2b;
  const double cvmb = -lcpkb * lcpk * rcp64(cvm);
  xb.qv += cvmb * c.c_vap_d;
  xb.q_liq += cvmb * c.c_vap_d;
  xb.q_sol += cvmb * c.c_vap_d;
  xb.tz += lcpkb * rcp64(cvm) * c.d0_vap_d;
  return icpk7b;
}

// ------------------------------------------------------------------ driver
// Vector-Jacobian product of icloud for one cell.  s: the cell at icloud's
// entry; yb: in, the cotangents of the cell after icloud (qv, ql, qr, qi, qs,
// qg, tz, qa); out, those of the cell before.  ib accumulates the input
// cotangents.  out (optional) receives the cell after icloud.  With MID, once
// the forward sweep has formed the cell after icloud, mid(cell, yb) is called
// before the reverse sweep reads yb (so the cotangents of icloud's outputs can
// be formed from its outputs in the same pass).
struct IcNoMid {
  SH_FN void operator()(const Cell&, double (&)[8]) const {}
};
// With xb_mf, the reverse stops before the melt / freeze block: xb_mf
// receives the cotangents of the state after it (the cell and the running
// q_liq, q_sol, cvm), and yb is left as it was after mid.
template <bool MID = false, class Mid = IcNoMid>
SH_FN void icloud_vjp(const IcK& k, const Cell& s, double (&yb)[8], IcInB& ib, Cell* out = nullptr,
                      const Mid& mid = Mid(), IcX* xb_mf = nullptr) {
  const Coef<double>& c = k.c;
  const IcX x0{s.qv, s.ql, s.qr, s.qi, s.qs, s.qg, s.tz, s.qa, 0.0, 0.0, 0.0};
  IcX a = x0;
  blk_melt_freeze(k, a);
  const bool im = k.in.p1 >= P_MIN;
  const bool warm = im && (a.tz - k.tice >= 0.0);
  const bool cold = im && !warm;
  IcX b = a;
  if (warm) blk_w1(k, b);
  if (cold) blk_c1(k, b);
  IcX d = b;
  if (warm) blk_w2(k, d);
  if (cold) blk_c2(k, d);
  IcX e = d;
  if (cold) blk_c3(k, e);
  IcX g = e;
  if (cold) blk_c4(k, g);
  blk_sg_reset(k, g);
  const double dt_pisub = c.fast_sat_adj ? 0.5 * c.dts : c.dts;
  IcX t = g;
  bool tmin = false, coll = false;
  if (im) tmin = blk_tmin(k, t);
  if (im && !tmin) coll = blk_coll(k, t);
  const bool run = im && !tmin && !coll;
  IcX i1 = g, i2, i3, i4, i5, i6, i7;
  if (run) blk_evap(k, i1);
  i2 = i1;
  if (run) blk_frz(k, i2);
  i3 = i2;
  if (run) blk_bigg(k, i3);
  i4 = i3;
  if (run) blk_idep(k, i4, dt_pisub);
  i5 = i4;
  if (run) blk_sdep(k, i5);
  const double icpk7 = (LI00 + DC_ICE * i5.tz) / i5.cvm;
  i6 = i5;
  if (run) blk_gdep(k, i6);
  i7 = i6;
  if (run) blk_revp(k, i7);
  if (MID || out) {
    // member by member: a select of the whole struct is compiled as a
    // per-lane pointer into a private copy of both (scratch memory)
    IcX o;
    o.qv = run ? i7.qv : t.qv;
    o.ql = run ? i7.ql : t.ql;
    o.qr = run ? i7.qr : t.qr;
    o.qi = run ? i7.qi : t.qi;
    o.qs = run ? i7.qs : t.qs;
    o.qg = run ? i7.qg : t.qg;
    o.tz = run ? i7.tz : t.tz;
    o.qa = run ? i7.qa : t.qa;
    o.q_liq = run ? i7.q_liq : t.q_liq;
    o.q_sol = run ? i7.q_sol : t.q_sol;
    o.cvm = run ? i7.cvm : t.cvm;
    if (run) blk_cf(k, o, icpk7);
    const Cell oc{o.qv, o.ql, o.qr, o.qi, o.qs, o.qg, o.tz, o.qa};
    if (out) *out = oc;
    if constexpr (MID) mid(oc, yb);
  }
  // reverse
  IcX xb{yb[0], yb[1], yb[2], yb[3], yb[4], yb[5], yb[6], yb[7], 0.0, 0.0, 0.0};
  if (run) {
    const double icpk7b = blk_cf_rev(k, i7, icpk7, xb, ib);
    blk_revp_rev(k, i6, xb, ib);
    blk_gdep_rev(k, i5, icpk7b, xb, ib);
    blk_sdep_rev(k, i4, xb, ib);
    blk_idep_rev(k, i3, dt_pisub, xb, ib);
    blk_bigg_rev(k, i2, xb, ib);
    blk_frz_rev(k, i1, xb);
    blk_evap_rev(k, g, xb, ib);
  } else if (coll) {
    blk_coll_rev(k, g, xb);
  } else if (tmin) {
    blk_tmin_rev(k, g, xb);
  }
  blk_sg_reset_rev(k, xb);
  if (cold) {
    blk_c4_rev(k, e, xb, ib);
    blk_c3_rev(k, d, xb, ib);
    blk_c2_rev(k, b, xb, ib);
    blk_c1_rev(k, a, xb, ib);
  } else if (warm) {
    blk_w2_rev(k, b, xb, ib);
    blk_w1_rev(k, a, xb, ib);
  }
  if (xb_mf) {
    *xb_mf = xb;
    return;
  }
  blk_melt_freeze_rev(k, x0, xb, ib);
  yb[0] = xb.qv;
  yb[1] = xb.ql;
  yb[2] = xb.qr;
  yb[3] = xb.qi;
  yb[4] = xb.qs;
  yb[5] = xb.qg;
  yb[6] = xb.tz;
  yb[7] = xb.qa;
}
// ---------------------------------------------------------------- warm_rain
// Vector-Jacobian product of revap_racc (evaporation of rain below
// saturation, accretion of cloud water by rain) for one cell.  s: the cell
// before the call; xb: in, the cotangents of the cell after it (qv .. tz;
// q_liq, q_sol, cvm zero); out, before it.  Adds den's and denfac's
// cotangents.
SH_FN void revap_vjp(const Coef<double>& c, double t_wfr, double dt, double h_var, const Cell& s, double den,
                     double denfac, IcX& xb, double& denb, double& denfacb) {
  if (!((s.tz > t_wfr) && (s.qr > QRMIN))) return;
  const double lhl = c.lv00_d + c.d0_vap_d * s.tz;
  const double q_liq = s.ql + s.qr;
  const double q_sol = s.qi + s.qs + s.qg;
  const double cvm = ic_cvm(c, s.qv, q_liq, q_sol);
  const double lcpk = lhl / cvm;
  const double tin = s.tz - lcpk * s.ql;
  const double qpz = s.qv + s.ql;
  double dqsdt;
  const double qsat = wqs2<double>(tin, den, dqsdt);
  const double mq = mmax(qpz, QCMIN);
  const double dqh0 = mmax(s.ql, h_var * mq);
  const double dqh = mmin(dqh0, 0.2 * qpz);
  const double dqv = qsat - s.qv;
  const double q_minus = qpz - dqh;
  const double q_plus = qpz + dqh;
  const bool ev_on = (dqv > QVMIN) && (qsat > q_minus);
  double dq = 0.0, qden = 0.0, t2 = 0.0, Bq = 0.0, Dd = 0.0, evap = 0.0, a = 0.0, Dn = 0.0, b = 0.0, ev = 0.0;
  double pe7 = 0.0, cvm1 = 0.0;
  double qr1 = s.qr;
  if (ev_on) {
    if (qsat > q_plus)
      dq = qsat - qpz;
    else
      dq = 0.25 * ((q_minus - qsat) * (q_minus - qsat)) / dqh;
    qden = s.qr * den;
    t2 = tin * tin;
    pe7 = pw_el<double>(qden, 0.725);
    Bq = ck(c, T_CREVP + 1) * sqrt(qden) + ck(c, T_CREVP + 2) * pe7;
    Dd = ck(c, T_CREVP + 3) * t2 + ck(c, T_CREVP + 4) * qsat * den;
    evap = ck(c, T_CREVP + 0) * t2 * dq * Bq / Dd;
    a = dt * evap;
    Dn = 1.0 + lcpk * dqsdt;
    b = dqv / Dn;
    ev = mmin(s.qr, mmin(a, b));
    qr1 = s.qr - ev;
    cvm1 = ic_cvm(c, s.qv + ev, q_liq - ev, q_sol);
  }
  // accretion of cloud water by rain
  if ((qr1 > QRMIN) && (s.ql > 1.0e-6) && (qsat < q_minus)) {
    const double pe = pw_el<double>(qr1 * den, 0.95);
    const double sink = dt * denfac * c.cracw * pe;
    const double D = 1.0 + sink;
    const double skb = xb.qr - xb.ql;
    xb.ql += skb * sink * rcp64(D);
    const double sinkb = skb * s.ql * rcp64(D * D);
    denfacb += sinkb * dt * c.cracw * pe;
    const double qdb = sinkb * dt * denfac * c.cracw * 0.95 * pe * rcp64(qr1 * den);
    xb.qr += qdb * den;
    denb += qdb * qr1;
  }
  double lcpkb = 0.0, lhlb = 0.0, tinb = 0.0, qpzb = 0.0, dqhb = 0.0, qsatb = 0.0, dqsdtb = 0.0;
  if (ev_on) {
    double evb = heat_rev(c, -1.0, ev, lhl, cvm1, xb, lhlb);
    evb += xb.qv - xb.qr - xb.q_liq;
    const double m = mmin(a, b);
    double mb = 0.0;
    if (s.qr < m)
      xb.qr += evb;
    else
      mb = evb;
    double evapb = 0.0, dqvb = 0.0;
    if (a < b) {
      evapb = mb * dt;
    } else {
      dqvb += mb * rcp64(Dn);
      const double Dnb = -mb * b * rcp64(Dn);
      lcpkb += Dnb * dqsdt;
      dqsdtb += Dnb * lcpk;
    }
    const double Nb = evapb * rcp64(Dd), Ddb = -evapb * evap * rcp64(Dd);
    const double c0 = ck(c, T_CREVP + 0);
    const double t2b = Nb * c0 * dq * Bq + Ddb * ck(c, T_CREVP + 3);
    const double dqb = Nb * c0 * t2 * Bq;
    const double Bqb = Nb * c0 * t2 * dq;
    qsatb += Ddb * ck(c, T_CREVP + 4) * den;
    denb += Ddb * ck(c, T_CREVP + 4) * qsat;
    const double qdenb = Bqb * (ck(c, T_CREVP + 1) * 0.5 * rcp64(sqrt(qden)) + ck(c, T_CREVP + 2) * 0.725 * pe7 * rcp64(qden));
    xb.qr += qdenb * den;
    denb += qdenb * s.qr;
    tinb += t2b * 2.0 * tin;
    if (qsat > q_plus) {
      qsatb += dqb;
      qpzb -= dqb;
    } else {
      const double e = q_minus - qsat;
      const double eb = dqb * 0.5 * e * rcp64(dqh);
      dqhb -= dqb * dq * rcp64(dqh);
      qpzb += eb;  // q_minus = qpz - dqh
      dqhb -= eb;
      qsatb -= eb;
    }
    qsatb += dqvb;
    xb.qv -= dqvb;
  }
  // dqh = min(max(ql, h_var max(qpz, QCMIN)), 0.2 qpz)
  if (dqh0 < 0.2 * qpz) {
    if (s.ql > h_var * mq)
      xb.ql += dqhb;
    else if (qpz > QCMIN)
      qpzb += dqhb * h_var;
  } else {
    qpzb += 0.2 * dqhb;
  }
  sat_rev(false, tin, den, qsat, dqsdt, qsatb, dqsdtb, tinb, denb);
  xb.qv += qpzb;
  xb.ql += qpzb;
  // tin = tz - lcpk ql
  xb.tz += tinb;
  lcpkb -= tinb * s.ql;
  xb.ql -= tinb * lcpk;
  lhlb += lcpkb * rcp64(cvm);
  const double cb = xb.cvm - lcpkb * lcpk * rcp64(cvm);
  xb.qv += cb * c.c_vap_d;
  const double lb = xb.q_liq + cb * C_LIQ, sb = xb.q_sol + cb * C_ICE;
  xb.ql += lb;
  xb.qr += lb;
  xb.qi += sb;
  xb.qs += sb;
  xb.qg += sb;
  xb.q_liq = xb.q_sol = xb.cvm = 0.0;
  xb.tz += lhlb * c.d0_vap_d;
}

// Cotangents of the autoconversion's inputs besides the cell
struct AcB {
  double den, ccn, c_praut, dl, fac_rc;
};

// Vector-Jacobian product of autoconv_no_subgrid_var (subgrid = false) or
// autoconv_subgrid_var with the width dl (subgrid = true) for one cell.
// qlb, qrb: in, the cotangents of ql and qr after the call; out, before.
SH_FN void autoconv_vjp(const Coef<double>& c, double t_wfr, bool subgrid, double tz, double ql, double den,
                        double ccn, double c_praut, double dl, double& qlb, double& qrb, AcB& ab) {
  const double qc0 = c.fac_rc * ccn;
  const double qc = c.use_ccn ? qc0 : qc0 / den;
  double qcb = 0.0;
  if (!subgrid) {
    if (!(tz > t_wfr)) return;
    const double dq = ql - qc;
    if (!(dq > 0.0)) return;
    const double pe = pw_el<double>(ql, 7.0 / 3.0);
    const double a2 = c.dt_rain * c_praut * den * pe;
    const double sinkb = qrb - qlb;
    if (dq < a2) {
      qlb += sinkb;
      qcb -= sinkb;
    } else {
      ab.c_praut += sinkb * c.dt_rain * den * pe;
      ab.den += sinkb * c.dt_rain * c_praut * pe;
      qlb += sinkb * c.dt_rain * c_praut * den * (7.0 / 3.0) * pe * rcp64(ql);
    }
  } else {
    if (!(tz > t_wfr + DT_FR)) return;
    const double mx = mmax(1.0e-6, dl);
    const double dl1 = mmin(mx, 0.5 * ql);
    const double dq = 0.5 * (ql + dl1 - qc);
    if (!(dq > 0.0)) return;
    const double pe = pw_el<double>(ql, 7.0 / 3.0);
    const double r = dq / dl1;
    const double m = mmin(1.0, r);
    const double sinkb = qrb - qlb;
    const double K = c.dt_rain * c_praut * den * pe;
    const double mb = sinkb * K;
    ab.c_praut += sinkb * m * c.dt_rain * den * pe;
    ab.den += sinkb * m * c.dt_rain * c_praut * pe;
    qlb += sinkb * m * c.dt_rain * c_praut * den * (7.0 / 3.0) * pe * rcp64(ql);
    double dqb = 0.0, dl1b = 0.0;
    if (!(1.0 < r)) {
      dqb += mb * rcp64(dl1);
      dl1b -= mb * r * rcp64(dl1);
    }
    qlb += 0.5 * dqb;
    dl1b += 0.5 * dqb;
    qcb -= 0.5 * dqb;
    if (mx < 0.5 * ql) {
      if (!(1.0e-6 > dl)) ab.dl += dl1b;
    } else {
      qlb += 0.5 * dl1b;
    }
  }
  double qc0b = qcb;
  if (!c.use_ccn) {
    qc0b = qcb * rcp64(den);
    ab.den -= qcb * qc * rcp64(den);
  }
  ab.fac_rc += qc0b * ccn;
  ab.ccn += qc0b * c.fac_rc;
}
// ------------------------------------------------------------ sedimentation
// Vector-Jacobian product of sedi_ice_melt (microphysics.py:702-746), the
// melt branch of icloud_melt_freeze.  xb: in, the cotangents of the cell after
// the melt (q_liq, q_sol, cvm zero); out, before it.
SH_FN void sedi_ice_melt_vjp(const Coef<double>& c, double tice, double ql_mlt, const Cell& s, IcX& xb) {
  const IcK k{c, tice, 0.0, ql_mlt, 0.0, 0.0, 0.0, 0.0, {1.0, 1.0, 0.0, 0.0, 0.0, 0.0}};
  IcInB ib{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const IcX x{s.qv, s.ql, s.qr, s.qi, s.qs, s.qg, s.tz, s.qa, 0.0, 0.0, 0.0};
  blk_melt_freeze_rev(k, x, xb, ib, false);
}

// Cotangents of the fall speeds' inputs (fall_speed, microphysics_funcs.py:269-341)
struct FsB {
  double qg, qi, qs, tc, den, vs_fac, vg_fac, vi_fac;
};

// sqrt(min(10, SFCRHO / den)) and its cotangent onto den
SH_FN double rhof_of(double den) { return sqrt(mmin(10.0, SFCRHO / den)); }
SH_FN void rhof_rev(double den, double rb, double& denb) {
  const double x = SFCRHO / den;
  if (!(10.0 < x)) denb -= rb * 0.5 * rcp64(sqrt(x)) * x * rcp64(den);
}

// Vector-Jacobian product of fall_speed with cotangents bg, bi, bs of vtg,
// vti, vts
SH_FN void fall_speed_vjp(const Coef<double>& c, double qg, double qi, double qs, double tc, double den, double bg,
                          double bi, double bs, FsB& fb) {
  const double rhof = rhof_of(den);
  double rhofb = 0.0;
  if (c.const_vi) {
    fb.vi_fac += bi;
  } else if (!(qi < THI)) {
    const double vi0 = 0.01 * c.vi_fac;
    const double lg = log(qi * den);
    const double poly = tc * (AA * tc + BB) + CC;
    const double v1 = (3.0 + lg / c.log_10) * poly + DD_FS * tc + EE;
    const double e = exp(c.log_10 * v1);
    const double v2 = vi0 * e * 0.8;
    const double mx = mmax(VF_MIN, v2);
    if (!(c.vi_max < mx) && !(VF_MIN > v2)) {
      fb.vi_fac += bi * 0.01 * e * 0.8;
      const double v1b = bi * vi0 * 0.8 * e * c.log_10;
      const double lgb = v1b * poly * rcp64(c.log_10);
      fb.tc += v1b * (3.0 + lg * rcp64(c.log_10)) * (2.0 * AA * tc + BB) + v1b * DD_FS;
      fb.qi += lgb * rcp64(qi);
      fb.den += lgb * rcp64(den);
    }
  }
  if (c.const_vs) {
    fb.vs_fac += bs;
  } else if (!(qs < THS)) {
    const double x = qs * den / NORMS;
    const double p = pw_el<double>(x, 0.0625);
    const double v = c.vs_fac * VCONS * rhof * p;
    if (!(c.vs_max < mmax(VF_MIN, v)) && !(VF_MIN > v)) {
      fb.vs_fac += bs * VCONS * rhof * p;
      rhofb += bs * c.vs_fac * VCONS * p;
      const double xb = bs * c.vs_fac * VCONS * rhof * 0.0625 * p * rcp64(x);
      fb.qs += xb * den * (1.0 / NORMS);
      fb.den += xb * qs * (1.0 / NORMS);
    }
  }
  if (c.const_vg) {
    fb.vg_fac += bg;
  } else if (!(qg < THG)) {
    const double x = qg * den / NORMG;
    const double r8 = sqrt(sqrt(sqrt(x)));
    const double v = c.vg_fac * VCONG * rhof * r8;
    if (!(c.vg_max < mmax(VF_MIN, v)) && !(VF_MIN > v)) {
      fb.vg_fac += bg * VCONG * rhof * r8;
      rhofb += bg * c.vg_fac * VCONG * r8;
      const double xb = bg * c.vg_fac * VCONG * rhof * 0.125 * r8 * rcp64(x);
      fb.qg += xb * den * (1.0 / NORMG);
      fb.den += xb * qg * (1.0 / NORMG);
    }
  }
  rhof_rev(den, rhofb, fb.den);
}

// Vector-Jacobian product of rain_fspeed (microphysics_funcs.py:344-372) with
// cotangent vb: adds to qrb, denb, vr_facb
SH_FN void rain_fspeed_vjp(const Coef<double>& c, bool no_fall, double qr, double den, double vb, double& qrb,
                           double& denb, double& vr_facb) {
  if (no_fall) return;
  if (c.const_vr) {
    vr_facb += vb;
    return;
  }
  if (qr < THR) return;
  const double rhof = rhof_of(den);
  const double x = qr * den / NORMR;
  const double p = pw_el<double>(x, 0.2);
  const double v = c.vr_fac * VCONR * rhof * p;
  if (c.vr_max < mmax(VR_MIN, v) || VR_MIN > v) return;
  vr_facb += vb * VCONR * rhof * p;
  rhof_rev(den, vb * c.vr_fac * VCONR * p, denb);
  const double xb = vb * c.vr_fac * VCONR * rhof * 0.2 * p * rcp64(x);
  qrb += xb * den * (1.0 / NORMR);
  denb += xb * qr * (1.0 / NORMR);
}
}  // namespace shield
