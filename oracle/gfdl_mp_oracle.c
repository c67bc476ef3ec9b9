/*
 * gfdl_mp_oracle.c -- TEST INFRASTRUCTURE, not product code.
 *
 * FP64 CPU restatement ("the oracle") of pySHiELD's GFDL single-moment cloud
 * microphysics, Microphysics.__call__ (pySHiELD/stencils/microphysics.py
 * :1981-2263) with its process library (pySHiELD/functions/
 * microphysics_funcs.py).  Each GT4Py `computation` block is restated as a
 * loop over k in the block's own order (PARALLEL = any order, FORWARD =
 * top->bottom, BACKWARD = bottom->top) so that recurrences read exactly the
 * values the reference reads, and every arithmetic expression keeps the
 * reference's operation order.  One column at a time, OpenMP over columns.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this file, and only as the checker / CPU baseline.
 *
 * Rejected namelist switches (as the product does): use_ppm and prog_ccn,
 * both broken upstream (microphysics.py:413-443 computes no Lagrangian fall;
 * :167-170 reads a never-filled qn).
 *
 * Parity status of each part is recorded in oracle/README.md.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mp_oracle.h"

/* ------------------------------------------------------------------------
 * ndsl.constants, GFS set (values restated; NDSL itself is not available).
 * ---------------------------------------------------------------------- */
#define GRAV 9.80665
#define RDGAS 287.05
#define RVGAS 461.50
#define HLV 2.5e6
#define HLF 3.3358e5
#define CP_AIR 1004.6
#define TICE 273.16
#define C_LIQ 4185.5
#define C_ICE 1972.0
#define E00 611.21
#define T_MIN 178.0
static const double RGRAV = 1.0 / GRAV;
static const double CV_AIR = CP_AIR - RDGAS;
static const double CV_VAP = 3.0 * RVGAS;
static const double CP_VAP = 4.0 * RVGAS;
static const double EPS = RDGAS / RVGAS;
static const double DC_VAP = 4.0 * RVGAS - C_LIQ;
static const double DC_ICE = C_LIQ - C_ICE;
static const double D2ICE = (4.0 * RVGAS - C_LIQ) + (C_LIQ - C_ICE);
static const double LV0 = HLV - (4.0 * RVGAS - C_LIQ) * TICE;
static const double LI00 = HLF - (C_LIQ - C_ICE) * TICE;
static const double LI2 = (HLV - (4.0 * RVGAS - C_LIQ) * TICE) + (HLF - (C_LIQ - C_ICE) * TICE);
static const double T_SAT_MIN = TICE - 160.0;
static const double LAT2 = (HLV + HLF) * (HLV + HLF);

/* microphysics_funcs.py:7-42 */
#define VCONS 6.6280504
#define VCONG 87.2382675
#define NORMS 942477796.076938
#define NORMG 5026548245.74367
#define VCONR 2503.23638966667
#define NORMR 25132741228.7183
#define THR 1.0e-8
#define THI 1.0e-8
#define THG 1.0e-8
#define THS 1.0e-8
#define AA (-4.14122e-5)
#define BB (-0.00538922)
#define CC (-0.0516344)
#define DD_FS 0.00216078
#define EE 1.9714
#define VR_MIN 1.0e-3
#define VF_MIN 1.0e-5
#define P_MIN 100.0
#define DT_FR 8.0
#define SFCRHO 1.2
#define RHOS 1.0e2
#define RHOG 4.0e2
#define RHOR 1.0e3
#define QCMIN 1.0e-12
#define QRMIN 1.0e-8
#define QVMIN 1.0e-20

/* ------------------------------------------------------------------------
 * Derived scalars (Microphysics.__init__, setupm, _set_timestep).
 * ---------------------------------------------------------------------- */
typedef struct {
  oracle_nml n;
  double c_air, c_vap, d0_vap, lv00, cpaut, so3, zs, log_10, tice0, t_wfr;
  int use_ccn, do_sedi_w;
  double fac_rc, cracs, csacr, cgacr, cgacs, acco[3][4], csacw, csaci, cgacw, cgaci, cracw;
  double cssub[5], cgsub[5], crevp[5], cgfr[2], csmlt[5], cgmlt[5], ces0;
  double mpdt, rdt, dts, dt_rain, rdts, dt_evap;
  double fac_i2s, fac_g2v, fac_v2g, fac_imlt, fac_l2v;
  int ntimes;
} P_t;

static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }
/* microphysics_funcs.py:45-49 */
static double dim(double x, double y) {
  double diff = x - y;
  return diff > 0.0 ? diff : 0.0;
}

/* microphysics.py:1964-1979; Python round() is round-half-even, as nearbyint
 * under the default rounding mode. */
static void set_timestep(P_t* p, double timestep) {
  p->mpdt = dmin(timestep, p->n.mp_time);
  p->rdt = 1.0 / timestep;
  p->ntimes = (int)nearbyint(timestep / p->mpdt);
  p->dts = timestep / p->ntimes;
  p->dt_rain = p->dts * 0.5;
  p->rdts = 1.0 / p->dts;
  p->dt_evap = p->n.fast_sat_adj ? 0.5 * p->dts : p->dts;
  p->fac_i2s = 1.0 - exp(-p->dts / p->n.tau_i2s);
  p->fac_g2v = 1.0 - exp(-p->dts / p->n.tau_g2v);
  p->fac_v2g = 1.0 - exp(-p->dts / p->n.tau_v2g);
  p->fac_imlt = 1.0 - exp(-0.5 * p->dts / p->n.tau_imlt);
  p->fac_l2v = 1.0 - exp(-p->dt_evap / p->n.tau_l2v);
}

/* microphysics.py:1643-1659, 1712-1713, 1806-1954 */
static void setup(P_t* p, const oracle_nml* n, double timestep) {
  memset(p, 0, sizeof(*p));
  p->n = *n;
  p->c_air = CP_AIR;
  p->c_vap = CP_VAP;
  p->d0_vap = p->c_vap - C_LIQ;
  p->lv00 = HLV - p->d0_vap * TICE;
  p->do_sedi_w = n->hydrostatic ? 0 : n->do_sedi_w;
  p->cpaut = n->c_paut * 0.104 * GRAV / 1.717e-5;
  p->use_ccn = n->prog_ccn ? 0 : 1;
  p->so3 = 7.0 / 3.0;
  p->zs = 0.0;
  p->log_10 = log(10.0);
  p->tice0 = n->tice - 0.01;
  p->t_wfr = n->tice - 40.0;

  const double gam263 = 1.456943, gam275 = 1.608355, gam290 = 1.827363;
  const double gam325 = 2.54925, gam350 = 3.323363, gam380 = 4.694155;
  const double rnzs = 3.0e6, rnzr = 8.0e6, rnzg = 4.0e6;
  const double acc[3] = {5.0, 2.0, 0.5};
  const double pie = 4.0 * atan(1.0);
  const double vdifu = 2.11e-5, tcond = 2.36e-2, visk = 1.259e-5;
  const double hlts = 2.8336e6, hltc = 2.5e6, hltf = 3.336e5, ch2o = 4.1855e3;
  p->fac_rc = (4.0 / 3.0) * pie * RHOR * pow(n->rthresh, 3.0);
  const double pisq = pie * pie;
  const double scm3 = pow(visk / vdifu, 1.0 / 3.0);
  p->cracs = pisq * rnzr * rnzs * RHOS;
  p->csacr = pisq * rnzr * rnzs * RHOR;
  p->cgacr = pisq * rnzr * rnzg * RHOR;
  p->cgacs = pisq * rnzg * rnzs * RHOS;
  p->cgacs = p->cgacs * n->c_pgacs;
  double act[8];
  act[0] = pie * rnzs * RHOS;
  act[1] = pie * rnzr * RHOR;
  act[5] = pie * rnzg * RHOG;
  act[2] = act[1];
  act[3] = act[0];
  act[4] = act[1];
  act[6] = act[0];
  act[7] = act[5];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 4; ++k)
      p->acco[i][k] = acc[i] / (pow(act[2 * k], (6 - i) * 0.25) * pow(act[2 * k + 1], (i + 1) * 0.25));
  const double gcon = 40.74 * sqrt(SFCRHO);
  p->csacw = pie * rnzs * n->clin * gam325 / (4.0 * pow(act[0], 0.8125));
  const double craci = pie * rnzr * n->alin * gam380 / (4.0 * pow(act[1], 0.95));
  p->csaci = p->csacw * n->c_psaci;
  p->cgacw = pie * rnzg * gam350 * gcon / (4.0 * pow(act[5], 0.875));
  p->cgaci = p->cgacw * 0.05;
  p->cracw = n->c_cracw * craci;
  p->cssub[0] = 2.0 * pie * vdifu * tcond * RVGAS * rnzs;
  p->cssub[1] = 0.78 / sqrt(act[0]);
  p->cssub[2] = 0.31 * scm3 * gam263 * sqrt(n->clin / visk) / pow(act[0], 0.65625);
  p->cssub[3] = tcond * RVGAS;
  p->cssub[4] = (hlts * hlts) * vdifu;
  p->cgsub[0] = 2.0 * pie * vdifu * tcond * RVGAS * rnzg;
  p->cgsub[1] = 0.78 / sqrt(act[5]);
  p->cgsub[2] = 0.31 * scm3 * gam275 * sqrt(gcon / visk) / pow(act[5], 0.6875);
  p->cgsub[3] = p->cssub[3];
  p->cgsub[4] = p->cssub[4];
  p->crevp[0] = 2.0 * pie * vdifu * tcond * RVGAS * rnzr;
  p->crevp[1] = 0.78 / sqrt(act[1]);
  p->crevp[2] = 0.31 * scm3 * gam290 * sqrt(n->alin / visk) / pow(act[1], 0.725);
  p->crevp[3] = p->cssub[3];
  p->crevp[4] = hltc * hltc * vdifu;
  p->cgfr[0] = 20.0e2 * pisq * rnzr * RHOR / pow(act[1], 1.75);
  p->cgfr[1] = 0.66;
  p->csmlt[0] = 2.0 * pie * tcond * rnzs / hltf;
  p->csmlt[1] = 2.0 * pie * vdifu * rnzs * hltc / hltf;
  p->csmlt[2] = p->cssub[1];
  p->csmlt[3] = p->cssub[2];
  p->csmlt[4] = ch2o / hltf;
  p->cgmlt[0] = 2.0 * pie * tcond * rnzg / hltf;
  p->cgmlt[1] = 2.0 * pie * vdifu * rnzg * hltc / hltf;
  p->cgmlt[2] = p->cgsub[1];
  p->cgmlt[3] = p->cgsub[2];
  p->cgmlt[4] = ch2o / hltf;
  const double es0 = 6.107799961e2;
  p->ces0 = EPS * es0;
  set_timestep(p, timestep);
}

/* ------------------------------------------------------------------------
 * Point-wise process functions (microphysics_funcs.py).
 * ---------------------------------------------------------------------- */

/* :53-64 */
static double wqs1(double ta, double den) {
  return (E00 * exp((DC_VAP * log(ta / TICE) + LV0 * (ta - TICE) / (ta * TICE)) / RVGAS)) /
         (RVGAS * ta * den);
}
/* :68-72 */
static double wqs2(double ta, double den, double* dqdt) {
  double tmp = wqs1(ta, den);
  *dqdt = tmp * (DC_VAP + LV0 / ta) / (RVGAS * ta);
  return tmp;
}
/* :76-111 */
static double iqs1(double ta, double den) {
  double tmp;
  if (ta < TICE) {
    if (ta >= T_SAT_MIN) {
      tmp = (E00 * exp((D2ICE * log(ta / TICE) + LI2 * (ta - TICE) / (ta * TICE)) / RVGAS)) /
            (RVGAS * ta * den);
    } else {
      tmp = (E00 * exp((D2ICE * log(1.0 - 160.0 / TICE) - LI2 * 160.0 / (T_SAT_MIN * TICE)) / RVGAS)) /
            (RVGAS * T_SAT_MIN * den);
    }
  } else {
    if (ta <= TICE + 102.0)
      tmp = wqs1(ta, den);
    else
      tmp = wqs1(TICE + 102.0, den);
  }
  return tmp;
}
/* :115-145 */
static double iqs2(double ta, double den, double* dqdt) {
  double tmp = iqs1(ta, den), dtmp;
  if (ta < TICE) {
    if (ta >= T_SAT_MIN)
      dtmp = tmp * (D2ICE + LI2 / ta) / (RVGAS * ta);
    else
      dtmp = tmp * (D2ICE + LI2 / T_SAT_MIN) / (RVGAS * T_SAT_MIN);
  } else {
    if (ta <= TICE + 102.0)
      dtmp = tmp * (DC_VAP + LV0 / ta) / (RVGAS * ta);
    else
      dtmp = tmp * (DC_VAP + LV0 / (TICE + 102.0)) / (RVGAS * (TICE + 102.0));
  }
  *dqdt = dtmp;
  return tmp;
}
/* :149-161 */
static double acr3d(double v1, double v2, double q1, double q2, double c, double cac_ik,
                    double cac_i1k, double cac_i2k, double rho) {
  double t1 = sqrt(q1 * rho);
  double s1 = sqrt(q2 * rho);
  double s2 = sqrt(s1);
  return c * fabs(v1 - v2) * q1 * s2 * (cac_ik * t1 + cac_i1k * sqrt(t1) * s2 + cac_i2k * s1);
}
/* :166-170 */
static double smlt(double tc, double dqs, double qsrho, double psacw, double psacr, const double* c,
                   double rho, double rhofac) {
  return (c[0] * tc / rho - c[1] * dqs) * (c[2] * sqrt(qsrho) + c[3] * pow(qsrho, 0.65625) * sqrt(rhofac)) +
         c[4] * tc * (psacw + psacr);
}
/* :175-179 */
static double gmlt(double tc, double dqs, double qgrho, double pgacw, double pgacr, const double* c,
                   double rho) {
  return (c[0] * tc / rho - c[1] * dqs) * (c[2] * sqrt(qgrho) + c[3] * pow(qgrho, 0.6875) / pow(rho, 0.25)) +
         c[4] * tc * (pgacw + pgacr);
}

/* :183-265 */
static void revap_racc(const P_t* p, double dt, double h_var, double qg, double qi, double* ql,
                       double* qr, double qs, double* qv, double* tz, double den, double denfac) {
  if ((*tz > p->t_wfr) && (*qr > QRMIN)) {
    double lhl = p->lv00 + p->d0_vap * *tz;
    double q_liq = *ql + *qr;
    double q_sol = qi + qs + qg;
    double cvm = p->c_air + *qv * p->c_vap + q_liq * C_LIQ + q_sol * C_ICE;
    double lcpk = lhl / cvm;
    double tin = *tz - lcpk * *ql;
    double qpz = *qv + *ql;
    double dqsdt;
    double qsat = wqs2(tin, den, &dqsdt);
    double dqh = dmax(*ql, h_var * dmax(qpz, QCMIN));
    dqh = dmin(dqh, 0.2 * qpz);
    double dqv = qsat - *qv;
    double q_minus = qpz - dqh;
    double q_plus = qpz + dqh;
    if ((dqv > QVMIN) && (qsat > q_minus)) {
      double dq;
      if (qsat > q_plus)
        dq = qsat - qpz;
      else
        dq = 0.25 * ((q_minus - qsat) * (q_minus - qsat)) / dqh;
      double qden = *qr * den;
      double t2 = tin * tin;
      double evap = p->crevp[0] * t2 * dq * (p->crevp[1] * sqrt(qden) + p->crevp[2] * exp(0.725 * log(qden))) /
                    (p->crevp[3] * t2 + p->crevp[4] * qsat * den);
      evap = dmin(*qr, dmin(dt * evap, dqv / (1.0 + lcpk * dqsdt)));
      *qr = *qr - evap;
      *qv = *qv + evap;
      q_liq = q_liq - evap;
      cvm = p->c_air + *qv * p->c_vap + q_liq * C_LIQ + q_sol * C_ICE;
      *tz = *tz - evap * lhl / cvm;
    }
    if ((*qr > QRMIN) && (*ql > 1.0e-6) && (qsat < q_minus)) {
      double sink = dt * denfac * p->cracw * exp(0.95 * log(*qr * den));
      sink = sink / (1.0 + sink) * *ql;
      *ql = *ql - sink;
      *qr = *qr + sink;
    }
  }
}

/* :269-341 */
static void fall_speed(const P_t* p, double qg, double qi, double qs, double tk, double den, double* vtg,
                       double* vti, double* vts) {
  const oracle_nml* n = &p->n;
  double rhof = sqrt(dmin(10.0, SFCRHO / den));
  if (n->const_vi) {
    *vti = n->vi_fac;
  } else {
    double vi0 = 0.01 * n->vi_fac;
    if (qi < THI) {
      *vti = VF_MIN;
    } else {
      double tc = tk - n->tice;
      double v = (3.0 + log(qi * den) / p->log_10) * (tc * (AA * tc + BB) + CC) + DD_FS * tc + EE;
      v = vi0 * exp(p->log_10 * v) * 0.8;
      *vti = dmin(n->vi_max, dmax(VF_MIN, v));
    }
  }
  if (n->const_vs) {
    *vts = n->vs_fac;
  } else if (qs < THS) {
    *vts = VF_MIN;
  } else {
    double v = n->vs_fac * VCONS * rhof * exp(0.0625 * log(qs * den / NORMS));
    *vts = dmin(n->vs_max, dmax(VF_MIN, v));
  }
  if (n->const_vg) {
    *vtg = n->vg_fac;
  } else if (qg < THG) {
    *vtg = VF_MIN;
  } else {
    double v = n->vg_fac * VCONG * rhof * sqrt(sqrt(sqrt(qg * den / NORMG)));
    *vtg = dmin(n->vg_max, dmax(VF_MIN, v));
  }
}

/* :344-372 (r1 is handled by the caller) */
static double rain_fspeed(const P_t* p, int no_fall, double qrz, double den) {
  const oracle_nml* n = &p->n;
  if (no_fall == 1) return VF_MIN;
  if (n->const_vr) return n->vr_fac;
  double qden = qrz * den;
  if (qrz < THR) return VR_MIN;
  double v = n->vr_fac * VCONR * sqrt(dmin(10.0, SFCRHO / den)) * exp(0.2 * log(qden / NORMR));
  return dmin(n->vr_max, dmax(VR_MIN, v));
}

/* :375-397 */
static void autoconv_no_subgrid_var(const P_t* p, double* qlz, double* qrz, double tz, double den,
                                    double ccn, double c_praut) {
  double qc0 = p->fac_rc * ccn;
  if (tz > p->t_wfr) {
    double qc = p->use_ccn ? qc0 : qc0 / den;
    double dq = *qlz - qc;
    if (dq > 0.0) {
      double sink = dmin(dq, p->dt_rain * c_praut * den * exp(p->so3 * log(*qlz)));
      *qlz = *qlz - sink;
      *qrz = *qrz + sink;
    }
  }
}

/* :400-428 */
static void autoconv_subgrid_var(const P_t* p, double* qlz, double* qrz, double tz, double den,
                                 double ccn, double c_praut, double dl) {
  double qc0 = p->fac_rc * ccn;
  if (tz > p->t_wfr + DT_FR) {
    dl = dmin(dmax(1.0e-6, dl), 0.5 * *qlz);
    double qc = p->use_ccn ? qc0 : qc0 / den;
    double dq = 0.5 * (*qlz + dl - qc);
    if (dq > 0.0) {
      double sink = dmin(1.0, dq / dl) * p->dt_rain * c_praut * den * exp(p->so3 * log(*qlz));
      *qlz = *qlz - sink;
      *qrz = *qrz + sink;
    }
  }
}

typedef struct {
  double qa, qg, qi, ql, qr, qs, qv, tz;
} cell_t;

#define CVM(c, ql_, qs_) (p->c_air + (c)->qv * p->c_vap + (ql_) * C_LIQ + (qs_) * C_ICE)

/* :431-850 */
static void subgrid_z_proc(const P_t* p, cell_t* c, double h_var, double rh_adj, double rh_rain,
                           double den, double denfac, double p1) {
  const oracle_nml* n = &p->n;
  const double tice = n->tice, t_wfr = p->t_wfr, t_sub = n->t_sub, dts = p->dts;
  double lhl = p->lv00 + p->d0_vap * c->tz;
  double lhi = LI00 + DC_ICE * c->tz;
  double q_liq = c->ql + c->qr;
  double q_sol = c->qi + c->qs + c->qg;
  double cvm = CVM(c, q_liq, q_sol);
  double lcpk = lhl / cvm;
  double icpk = lhi / cvm;
  double tcpk = lcpk + icpk;
  double tcp3 = lcpk + icpk * dmin(1.0, dim(tice, c->tz) / (tice - t_wfr));
  (void)tcpk;
  (void)denfac;
  if (!(p1 >= P_MIN)) return;
  if (c->tz < T_MIN) {
    double sink = dim(1.0e-7, c->qv);
    c->qv = c->qv - sink;
    c->qi = c->qi + sink;
    q_sol = q_sol + sink;
    cvm = CVM(c, q_liq, q_sol);
    c->tz = c->tz + sink * (lhl + lhi) / cvm;
    if (!n->do_qa) c->qa = c->qa + 1.0;
    return;
  }
  lhl = p->lv00 + p->d0_vap * c->tz;
  lhi = LI00 + DC_ICE * c->tz;
  lcpk = lhl / cvm;
  icpk = lhi / cvm;
  tcpk = lcpk + icpk;
  tcp3 = lcpk + icpk * dmin(1.0, dim(tice, c->tz) / (tice - t_wfr));

  double qpz = c->qv + c->ql + c->qi;
  double tin = c->tz - (lhl * (c->ql + c->qi) + lhi * c->qi) /
                           (p->c_air + qpz * p->c_vap + c->qr * C_LIQ + (c->qs + c->qg) * C_ICE);
  double rh = 0.0;
  if (tin > t_sub + 6.0) {
    rh = qpz / iqs1(tin, den);
    if (rh < rh_adj) {
      c->tz = tin;
      c->qv = qpz;
      c->ql = 0.0;
      c->qi = 0.0;
    }
  }
  if (!(((tin > t_sub + 6.0) && (rh >= rh_adj)) || (tin <= t_sub + 6.0))) return;

  /* cloud water <-> vapour adjustment */
  double dwsdt;
  double qsw = wqs2(c->tz, den, &dwsdt);
  double dq0 = qsw - c->qv;
  double evap;
  if (dq0 > 0.0) {
    double factor = dmin(1.0, p->fac_l2v * (10.0 * dq0 / qsw));
    evap = dmin(c->ql, factor * dq0 / (1.0 + tcp3 * dwsdt));
  } else {
    evap = dq0 / (1.0 + tcp3 * dwsdt);
  }
  c->qv = c->qv + evap;
  c->ql = c->ql - evap;
  q_liq = q_liq - evap;
  cvm = CVM(c, q_liq, q_sol);
  c->tz = c->tz - evap * lhl / cvm;
  lhi = LI00 + DC_ICE * c->tz;
  icpk = lhi / cvm;

  /* enforce complete freezing below t_wfr */
  double dtmp = t_wfr - c->tz;
  if ((dtmp > 0.0) && (c->ql > QCMIN)) {
    double sink = dmin(c->ql, dmin(c->ql * dtmp * 0.125, dtmp / icpk));
    c->ql = c->ql - sink;
    c->qi = c->qi + sink;
    q_liq = q_liq - sink;
    q_sol = q_sol + sink;
    cvm = CVM(c, q_liq, q_sol);
    c->tz = c->tz + sink * lhi / cvm;
  }
  lhi = LI00 + DC_ICE * c->tz;
  icpk = lhi / cvm;

  /* Bigg mechanism */
  double dt_pisub;
  if (n->fast_sat_adj) {
    dt_pisub = 0.5 * dts;
  } else {
    dt_pisub = dts;
    double tc = tice - c->tz;
    if ((c->ql > QRMIN) && (tc > 0.0)) {
      double sink = 3.3333e-10 * dts * (exp(0.66 * tc) - 1.0) * den * c->ql * c->ql;
      sink = dmin(c->ql, dmin(tc / icpk, sink));
      c->ql = c->ql - sink;
      c->qi = c->qi + sink;
      q_liq = q_liq - sink;
      q_sol = q_sol + sink;
      cvm = CVM(c, q_liq, q_sol);
      c->tz = c->tz + sink * lhi / cvm;
    }
  }
  lhl = p->lv00 + p->d0_vap * c->tz;
  lhi = LI00 + DC_ICE * c->tz;
  lcpk = lhl / cvm;
  icpk = lhi / cvm;
  tcpk = lcpk + icpk;

  /* sublimation / deposition of ice */
  if (c->tz < tice) {
    double dqsdt;
    double qsi = iqs2(c->tz, den, &dqsdt);
    double dq = c->qv - qsi;
    double sink = dq / (1.0 + tcpk * dqsdt);
    double pidep;
    if (c->qi > QRMIN) {
      pidep = dt_pisub * dq * 349138.78 * exp(0.875 * log(c->qi * den)) /
              (qsi * den * LAT2 / (0.0243 * RVGAS * (c->tz * c->tz)) + 4.42478e4);
    } else {
      pidep = 0.0;
    }
    if (dq > 0.0) {
      double tmp = tice - c->tz;
      double qi_crt = n->qi_gen * dmin(n->qi_lim, 0.1 * tmp) / den;
      sink = dmin(sink, dmin(dmax(qi_crt - c->qi, pidep), tmp / tcpk));
    } else {
      pidep = pidep * dmin(1.0, dim(c->tz, t_sub) * 0.2);
      sink = dmax(pidep, dmax(sink, -c->qi));
    }
    c->qv = c->qv - sink;
    c->qi = c->qi + sink;
    q_sol = q_sol + sink;
    cvm = CVM(c, q_liq, q_sol);
    c->tz = c->tz + sink * (lhl + lhi) / cvm;
  }
  lhl = p->lv00 + p->d0_vap * c->tz;
  lhi = LI00 + DC_ICE * c->tz;
  lcpk = lhl / cvm;
  icpk = lhi / cvm;
  tcpk = lcpk + icpk;

  /* sublimation / deposition of snow */
  if (c->qs > QRMIN) {
    double dqsdt;
    double qsi = iqs2(c->tz, den, &dqsdt);
    double qden = c->qs * den;
    double tmp = exp(0.65625 * log(qden));
    double tsq = c->tz * c->tz;
    double dq = (qsi - c->qv) / (1.0 + tcpk * dqsdt);
    double pssub = p->cssub[0] * tsq * (p->cssub[1] * sqrt(qden) + p->cssub[2] * tmp * sqrt(denfac)) /
                   (p->cssub[3] * tsq + p->cssub[4] * qsi * den);
    pssub = (qsi - c->qv) * dts * pssub;
    if (pssub > 0.0) {
      pssub = dmin(pssub * dmin(1.0, dim(c->tz, t_sub) * 0.2), c->qs);
    } else {
      if (c->tz > tice)
        pssub = 0.0;
      else
        pssub = dmax(pssub, dmax(dq, (c->tz - tice) / tcpk));
    }
    c->qs = c->qs - pssub;
    c->qv = c->qv + pssub;
    q_sol = q_sol - pssub;
    cvm = CVM(c, q_liq, q_sol);
    c->tz = c->tz - pssub * (lhl + lhi) / cvm;
  }
  lhl = p->lv00 + p->d0_vap * c->tz;
  lhi = LI00 + DC_ICE * c->tz;
  lcpk = lhl / cvm;
  icpk = lhi / cvm;
  tcpk = lcpk + icpk;

  /* graupel sublimation / deposition */
  if (c->qg > QRMIN) {
    double dqsdt;
    double qsi = iqs2(c->tz, den, &dqsdt);
    double dq = (c->qv - qsi) / (1.0 + tcpk * dqsdt);
    double pgsub = (c->qv / qsi - 1.0) * c->qg;
    if (pgsub > 0.0) {
      if (c->tz > tice)
        pgsub = 0.0;
      else
        pgsub = dmin(dmin(p->fac_v2g * pgsub, 0.2 * dq), dmin(c->ql + c->qr, (tice - c->tz) / tcpk));
    } else {
      pgsub = dmax(p->fac_g2v * pgsub, dq) * dmin(1.0, dim(c->tz, t_sub) * 0.1);
    }
    c->qg = c->qg + pgsub;
    c->qv = c->qv - pgsub;
    q_sol = q_sol + pgsub;
    cvm = CVM(c, q_liq, q_sol);
    c->tz = c->tz + pgsub * (lhl + lhi) / cvm;
  }

  /* minimum evaporation of rain in dry air */
  lhl = p->lv00 + p->d0_vap * c->tz;
  lcpk = lhl / cvm;
  if (c->qr > QCMIN) {
    double dqsdt;
    double qsw2 = wqs2(c->tz, den, &dqsdt);
    double sink = dmin(c->qr, dim(rh_rain * qsw2, c->qv) / (1.0 + lcpk * dqsdt));
    c->qv = c->qv + sink;
    c->qr = c->qr - sink;
    q_liq = q_liq - sink;
    cvm = CVM(c, q_liq, q_sol);
    c->tz = c->tz - sink * lhl / cvm;
  }
  lhl = p->lv00 + p->d0_vap * c->tz;
  cvm = p->c_air + (c->qv + q_liq + q_sol) * p->c_vap;
  lcpk = lhl / cvm;

  /* cloud fraction */
  if (!n->do_qa) {
    q_sol = n->rad_snow ? c->qi + c->qs : c->qi;
    q_liq = n->rad_rain ? c->ql + c->qr : c->ql;
    double q_cond = q_liq + q_sol;
    double qpz2 = c->qv + q_cond;
    double tin2 = c->tz - (lcpk * q_cond + icpk * q_sol);
    double qstar;
    if (tin2 <= t_wfr) {
      qstar = iqs1(tin2, den);
    } else if (tin2 >= tice) {
      qstar = wqs1(tin2, den);
    } else {
      double qsi = iqs1(tin2, den);
      double qsw3 = wqs1(tin2, den);
      double rqi;
      if (q_cond > 3.0e-6)
        rqi = q_sol / q_cond;
      else
        rqi = (tice - tin2) / (tice - t_wfr);
      qstar = rqi * qsi + (1.0 - rqi) * qsw3;
    }
    if (qpz2 > QRMIN) {
      double dq = dmax(QCMIN, h_var * qpz2);
      double q_plus = qpz2 + dq;
      double q_minus = qpz2 - dq;
      if (qstar < q_minus)
        c->qa = c->qa + 1.0;
      else if ((qstar < q_plus) && (q_cond > n->qc_crt))
        c->qa = c->qa + (q_plus - qstar) / (dq + dq);
    }
  }
}

/* :853-1260 */
static void icloud_main(const P_t* p, cell_t* c, double h_var, double rh_adj, double rh_rain,
                        double den, double denfac, double vtgz, double vtrz, double vtsz, double p1,
                        double di, double q_liq, double q_sol, double cvm) {
  const oracle_nml* n = &p->n;
  const double tice = n->tice, dts = p->dts, rdts = p->rdts;
  double lhi = LI00 + DC_ICE * c->tz;
  double icpk = lhi / cvm;
  if (p1 >= P_MIN) {
    double pgacr = 0.0, pgacw = 0.0;
    double tc = c->tz - tice;
    if (tc >= 0.0) {
      double dqs0 = p->ces0 / p1 - c->qv;
      if (c->qs > QCMIN) {
        double psacw, psacr, pracs;
        if (c->ql > QRMIN) {
          double factor = denfac * p->csacw * exp(0.8125 * log(c->qs * den));
          psacw = factor / (1.0 + dts * factor) * c->ql;
        } else {
          psacw = 0.0;
        }
        if (c->qr > QRMIN) {
          psacr = dmin(acr3d(vtsz, vtrz, c->qr, c->qs, p->csacr, p->acco[0][1], p->acco[1][1], p->acco[2][1], den),
                       c->qr * rdts);
          pracs = acr3d(vtrz, vtsz, c->qs, c->qr, p->cracs, p->acco[0][0], p->acco[1][0], p->acco[2][0], den);
        } else {
          psacr = 0.0;
          pracs = 0.0;
        }
        double psmlt = dmax(0.0, smlt(tc, dqs0, c->qs * den, psacw, psacr, p->csmlt, den, denfac));
        double sink = dmin(c->qs, dmin(dts * (psmlt + pracs), tc / icpk));
        c->qs = c->qs - sink;
        double tmp = dmin(sink, dim(n->qs_mlt, c->ql));
        c->ql = c->ql + tmp;
        c->qr = c->qr + sink - tmp;
        q_liq = q_liq + sink;
        q_sol = q_sol - sink;
        cvm = CVM(c, q_liq, q_sol);
        c->tz = c->tz - sink * lhi / cvm;
        tc = c->tz - tice;
      }
      lhi = LI00 + DC_ICE * c->tz;
      icpk = lhi / cvm;
      if ((c->qg > QCMIN) && (tc > 0.0)) {
        if (c->qr > QRMIN)
          pgacr = dmin(acr3d(vtgz, vtrz, c->qr, c->qg, p->cgacr, p->acco[0][2], p->acco[1][2], p->acco[2][2], den),
                       rdts * c->qr);
        double qden = c->qg * den;
        if (c->ql > QRMIN) {
          double factor = p->cgacw * qden / sqrt(den * sqrt(sqrt(qden)));
          pgacw = factor / (1.0 + dts * factor) * c->ql;
        }
        double pgmlt = dts * gmlt(tc, dqs0, qden, pgacw, pgacr, p->cgmlt, den);
        pgmlt = dmin(dmax(0.0, pgmlt), dmin(c->qg, tc / icpk));
        c->qg = c->qg - pgmlt;
        c->qr = c->qr + pgmlt;
        q_liq = q_liq + pgmlt;
        q_sol = q_sol - pgmlt;
        cvm = CVM(c, q_liq, q_sol);
        c->tz = c->tz - pgmlt * lhi / cvm;
      }
    } else {
      if (c->qi > 3.0e-7) {
        double psaci;
        if (c->qs > 1.0e-7) {
          double factor = dts * denfac * p->csaci * exp(0.05 * tc + 0.8125 * log(c->qs * den));
          psaci = factor / (1.0 + factor) * c->qi;
        } else {
          psaci = 0.0;
        }
        double qim = n->qi0_crt / den;
        double tmp = n->const_vi ? p->fac_i2s : p->fac_i2s * exp(0.025 * tc);
        di = dmax(di, QRMIN);
        double q_plus = c->qi + di;
        double psaut;
        if (q_plus > (qim + QRMIN)) {
          double dq;
          if (qim > (c->qi - di))
            dq = (0.25 * ((q_plus - qim) * (q_plus - qim))) / di;
          else
            dq = c->qi - qim;
          psaut = tmp * dq;
        } else {
          psaut = 0.0;
        }
        double sink = dmin(0.75 * c->qi, psaci + psaut);
        c->qi = c->qi - sink;
        c->qs = c->qs + sink;
        if (c->qg > 1.0e-6) {
          double factor = dts * p->cgaci * sqrt(den) * c->qg;
          double pgaci = factor / (1.0 + factor) * c->qi;
          c->qi = c->qi - pgaci;
          c->qg = c->qg + pgaci;
        }
      }
      tc = c->tz - tice;
      if ((c->qr > 1e-7) && (tc < 0.0)) {
        double psacr;
        if (c->qs > 1.0e-7)
          psacr = dts * acr3d(vtsz, vtrz, c->qr, c->qs, p->csacr, p->acco[0][1], p->acco[1][1], p->acco[2][1], den);
        else
          psacr = 0.0;
        double pgfr = dts * p->cgfr[0] / den * (exp(-p->cgfr[1] * tc) - 1.0) * exp(1.75 * log(c->qr * den));
        double sink = psacr + pgfr;
        double factor = dmin(sink, dmin(c->qr, -tc / icpk)) / dmax(sink, QRMIN);
        psacr = factor * psacr;
        pgfr = factor * pgfr;
        sink = psacr + pgfr;
        c->qr = c->qr - sink;
        c->qs = c->qs + psacr;
        c->qg = c->qg + pgfr;
        q_liq = q_liq - sink;
        q_sol = q_sol + sink;
        cvm = CVM(c, q_liq, q_sol);
        c->tz = c->tz + sink * lhi / cvm;
      }
      lhi = LI00 + DC_ICE * c->tz;
      icpk = lhi / cvm;
      if (c->qs > 1.0e-7) {
        double sink;
        if (c->qg > QRMIN)
          sink = dts * acr3d(vtgz, vtsz, c->qs, c->qg, p->cgacs, p->acco[0][3], p->acco[1][3], p->acco[2][3], den);
        else
          sink = 0.0;
        double qsm = n->qs0_crt / den;
        if (c->qs > qsm) {
          double factor = dts * 1.0e-3 * exp(0.09 * (c->tz - tice));
          sink = sink + factor / (1.0 + factor) * (c->qs - qsm);
        }
        sink = dmin(c->qs, sink);
        c->qs = c->qs - sink;
        c->qg = c->qg + sink;
      }
      if ((c->qg > 1.0e-7) && (c->tz < p->tice0)) {
        if (c->ql > 1.0e-6) {
          double qden = c->qg * den;
          double factor = dts * p->cgacw * qden / sqrt(den * sqrt(sqrt(qden)));
          pgacw = factor / (1.0 + factor) * c->ql;
        } else {
          pgacw = 0.0;
        }
        if (c->qr > 1.0e-6)
          pgacr = dmin(dts * acr3d(vtgz, vtrz, c->qr, c->qg, p->cgacr, p->acco[0][2], p->acco[1][2], p->acco[2][2], den),
                       c->qr);
        else
          pgacr = 0.0;
        double sink = pgacr + pgacw;
        double factor = dmin(sink, dim(tice, c->tz) / icpk) / dmax(sink, QRMIN);
        pgacr = factor * pgacr;
        pgacw = factor * pgacw;
        sink = pgacr + pgacw;
        c->qg = c->qg + sink;
        c->qr = c->qr - pgacr;
        c->ql = c->ql - pgacw;
        q_liq = q_liq - sink;
        q_sol = q_sol + sink;
        cvm = CVM(c, q_liq, q_sol);
        c->tz = c->tz + sink * lhi / cvm;
      }
    }
  }
  subgrid_z_proc(p, c, h_var, rh_adj, rh_rain, den, denfac, p1);
}

/* ------------------------------------------------------------------------
 * One column.
 * ---------------------------------------------------------------------- */
#define NZMAX ORACLE_NZMAX

typedef struct {
  /* prognostic (dry mixing ratios) */
  double qv[NZMAX], ql[NZMAX], qr[NZMAX], qi[NZMAX], qs[NZMAX], qg[NZMAX], tz[NZMAX];
  double qa0[NZMAX], qv0[NZMAX], ql0[NZMAX], qr0[NZMAX], qi0[NZMAX], qs0[NZMAX], qg0[NZMAX];
  double t0[NZMAX], dp0[NZMAX], den0[NZMAX], dz0[NZMAX], u0[NZMAX], v0[NZMAX];
  double dp1[NZMAX], p1[NZMAX], m1[NZMAX], u1[NZMAX], v1[NZMAX];
  double ccn[NZMAX], c_praut[NZMAX], h_var[NZMAX], rh_adj[NZMAX], rh_rain[NZMAX];
  double den[NZMAX], denfac[NZMAX], dz1[NZMAX];
  double vtrz[NZMAX], vtgz[NZMAX], vtsz[NZMAX], m1_sol[NZMAX];
  double rain[NZMAX], snow[NZMAX], ice[NZMAX], graupel[NZMAX];
  double w[NZMAX], qaz[NZMAX];
  /* stencil temporaries */
  double ze[NZMAX], dz[NZMAX], dd[NZMAX], qm[NZMAX], mflux[NZMAX], r1[NZMAX], dm[NZMAX];
  double dgz[NZMAX], cvn[NZMAX], dq[NZMAX], dl[NZMAX], vti[NZMAX], q_liq[NZMAX], q_sol[NZMAX];
  double cvm[NZMAX], icpk[NZMAX], lhi[NZMAX];
  int no_fall[NZMAX], stop_k[NZMAX];
} col_t;

/* Column "no_fall" flag, FORWARD + BACKWARD pair (microphysics.py:349-368):
 * 0 everywhere if any level has q > QRMIN, else 1 everywhere. */
static void column_no_fall(int nz, const double* q, int* no_fall) {
  for (int k = 0; k < nz; ++k) {
    if (k == 0)
      no_fall[k] = (q[k] > QRMIN) ? 0 : 1;
    else if (no_fall[k - 1] == 1)
      no_fall[k] = (q[k] > QRMIN) ? 0 : 1;
    else
      no_fall[k] = 0;
  }
  for (int k = nz - 2; k >= 0; --k)
    if (no_fall[k + 1] == 0) no_fall[k] = no_fall[k + 1];
}

/* Implicit upwind fall of one species (the shared pattern of
 * microphysics.py:445-515 and :830-900, :961-1035, :1094-1168):
 * q is a dry mixing ratio on entry and exit; flux[k] gets the downward mass
 * flux through the bottom of level k; returns the surface flux. The w
 * transport is applied when do_sedi_w. */
static double implicit_fall(col_t* C, int nz, double dt, const double* vt, double* q, double* flux,
                            int do_sedi_w, double zs) {
  for (int k = 0; k < nz; ++k)
    if (do_sedi_w)
      C->dm[k] = C->dp1[k] * (1.0 + C->qv[k] + C->ql[k] + C->qr[k] + C->qi[k] + C->qs[k] + C->qg[k]);
  for (int k = 0; k < nz - 1; ++k) C->dz[k] = C->ze[k] - C->ze[k + 1];
  C->dz[nz - 1] = C->ze[nz - 1] - zs;
  for (int k = 0; k < nz; ++k) {
    C->dd[k] = dt * vt[k];
    q[k] = q[k] * C->dp1[k];
  }
  C->qm[0] = q[0] / (C->dz[0] + C->dd[0]);
  for (int k = 1; k < nz; ++k) C->qm[k] = (q[k] + C->dd[k - 1] * C->qm[k - 1]) / (C->dz[k] + C->dd[k]);
  for (int k = 0; k < nz; ++k) C->qm[k] = C->qm[k] * C->dz[k];
  flux[0] = q[0] - C->qm[0];
  for (int k = 1; k < nz; ++k) flux[k] = flux[k - 1] + q[k] - C->qm[k];
  double surface = flux[nz - 1];
  for (int k = 0; k < nz; ++k) {
    q[k] = C->qm[k] / C->dp1[k];
    if (do_sedi_w) {
      if (k == 0)
        C->w[k] = (C->dm[k] * C->w[k] + flux[k] * vt[k]) / (C->dm[k] - flux[k]);
      else
        C->w[k] = (C->dm[k] * C->w[k] - flux[k - 1] * vt[k - 1] + flux[k] * vt[k]) /
                  (C->dm[k] + flux[k - 1] - flux[k]);
    }
  }
  return surface;
}

/* Sedimentation heating (microphysics.py:516-557 with C_LIQ, :1176-1217 with
 * C_ICE): explicit at the top level, implicit FORWARD below. */
static void sedi_heat(col_t* C, int nz, const double* flux, double cq) {
  for (int k = 0; k < nz; ++k) {
    C->dgz[k] = -0.5 * GRAV * C->dz1[k];
    C->cvn[k] = C->dp1[k] * (CV_AIR + C->qv[k] * CV_VAP + (C->qr[k] + C->ql[k]) * C_LIQ +
                             (C->qi[k] + C->qs[k] + C->qg[k]) * C_ICE);
  }
  {
    double tmp = C->cvn[0] + flux[0] * cq;
    C->tz[0] = C->tz[0] + flux[0] * C->dgz[0] / tmp;
  }
  for (int k = 1; k < nz; ++k)
    C->tz[k] = ((C->cvn[k] + cq * (flux[k] - flux[k - 1])) * C->tz[k] + flux[k - 1] * cq * C->tz[k - 1] +
                C->dgz[k] * (flux[k - 1] + flux[k])) /
               (C->cvn[k] + cq * flux[k]);
}

/* microphysics.py:279-653 */
static void warm_rain(const P_t* p, col_t* C, int nz, int is_first) {
  const oracle_nml* n = &p->n;
  double* mflux = C->mflux; /* m1_rain */
  for (int k = 0; k < nz; ++k) {
    if (is_first) {
      C->dz1[k] = C->dz0[k] * C->tz[k] / C->t0[k];
      C->den[k] = C->den0[k] * C->dz0[k] / C->dz1[k];
      C->denfac[k] = sqrt(SFCRHO / C->den[k]);
    }
    mflux[k] = 0.0;
  }
  const double dt5 = 0.5 * p->dt_rain;
  int* no_fall = C->no_fall;
  column_no_fall(nz, C->qr, no_fall);
  for (int k = 0; k < nz; ++k) {
    C->vtrz[k] = rain_fspeed(p, no_fall[k], C->qr[k], C->den[k]);
    C->r1[k] = 0.0;
  }
  const int falls = (no_fall[0] == 0);
  if (falls) {
    C->ze[nz - 1] = p->zs - C->dz1[nz - 1];
    for (int k = nz - 2; k >= 0; --k) C->ze[k] = C->ze[k + 1] - C->dz1[k];
    for (int k = 0; k < nz; ++k)
      revap_racc(p, dt5, C->h_var[k], C->qg[k], C->qi[k], &C->ql[k], &C->qr[k], C->qs[k], &C->qv[k], &C->tz[k],
                 C->den[k], C->denfac[k]);
    double r1 = implicit_fall(C, nz, p->dt_rain, C->vtrz, C->qr, mflux, p->do_sedi_w, p->zs);
    for (int k = 0; k < nz; ++k) C->r1[k] = r1;
    if (n->do_sedi_heat) sedi_heat(C, nz, mflux, C_LIQ);
    for (int k = 0; k < nz; ++k)
      revap_racc(p, dt5, C->h_var[k], C->qg[k], C->qi[k], &C->ql[k], &C->qr[k], C->qs[k], &C->qv[k], &C->tz[k],
                 C->den[k], C->denfac[k]);
  }
  if (n->irain_f != 0) {
    for (int k = 0; k < nz; ++k)
      autoconv_no_subgrid_var(p, &C->ql[k], &C->qr[k], C->tz[k], C->den[k], C->ccn[k], C->c_praut[k]);
  } else {
    if (n->z_slope_liq) {
      C->dl[0] = 0.0;
      for (int k = 1; k < nz; ++k) C->dq[k] = 0.5 * (C->ql[k] - C->ql[k - 1]);
      for (int k = 1; k < nz - 1; ++k) {
        C->dl[k] = 0.5 * dmin(fabs(C->dq[k] + C->dq[k + 1]), 0.5 * C->ql[k]);
        if (C->dq[k] * C->dq[k + 1] <= 0.0) {
          if (C->dq[k] > 0.0)
            C->dl[k] = dmin(C->dl[k], dmin(C->dq[k], -C->dq[k + 1]));
          else
            C->dl[k] = 0.0;
        }
      }
      C->dl[nz - 1] = 0.0;
    }
    for (int k = 0; k < nz; ++k) {
      double dl;
      if (n->z_slope_liq)
        dl = dmax(C->dl[k], dmax(QVMIN, C->h_var[k] * C->ql[k]));
      else
        dl = dmax(QVMIN, C->h_var[k] * C->ql[k]);
      autoconv_subgrid_var(p, &C->ql[k], &C->qr[k], C->tz[k], C->den[k], C->ccn[k], C->c_praut[k], dl);
    }
  }
  for (int k = 0; k < nz; ++k) {
    C->rain[k] = C->rain[k] + C->r1[k];
    if (is_first)
      C->m1[k] = C->m1[k] + mflux[k];
    else
      C->m1[k] = C->m1[k] + mflux[k] + C->m1_sol[k];
  }
}

/* microphysics.py:656-1217 */
static void sedimentation(const P_t* p, col_t* C, int nz) {
  const oracle_nml* n = &p->n;
  const double tice = n->tice;
  for (int k = 0; k < nz; ++k) {
    fall_speed(p, C->qg[k], C->qi[k], C->qs[k], C->tz[k], C->den[k], &C->vtgz[k], &C->vti[k], &C->vtsz[k]);
    C->m1_sol[k] = 0.0;
    C->lhi[k] = LI00 + DC_ICE * C->tz[k];
    C->q_liq[k] = C->ql[k] + C->qr[k];
    C->q_sol[k] = C->qi[k] + C->qs[k] + C->qg[k];
    C->cvm[k] = p->c_air + C->qv[k] * p->c_vap + C->q_liq[k] * C_LIQ + C->q_sol[k] * C_ICE;
    C->icpk[k] = C->lhi[k] / C->cvm[k];
  }
  /* melting level (:706-724) */
  for (int k = 0; k < nz; ++k) {
    if (k == nz - 1)
      C->stop_k[k] = 1;
    else if (k == 0)
      C->stop_k[k] = (C->tz[k] > tice) ? 1 : 0;
    else if (C->stop_k[k - 1] == 0)
      C->stop_k[k] = (C->tz[k] > tice) ? 1 : 0;
    else
      C->stop_k[k] = 1;
  }
  /* melting of cloud ice before fall (:726-746) */
  for (int k = 0; k < nz; ++k) {
    if (C->stop_k[k] == 1) {
      double tc = C->tz[k] - tice;
      if ((C->qi[k] > QCMIN) && (tc > 0.0)) {
        double sink = dmin(C->qi[k], p->fac_imlt * tc / C->icpk[k]);
        double tmp = dmin(sink, dim(n->ql_mlt, C->ql[k]));
        C->ql[k] = C->ql[k] + tmp;
        C->qr[k] = C->qr[k] + sink - tmp;
        C->qi[k] = C->qi[k] - sink;
        C->q_liq[k] = C->q_liq[k] + sink;
        C->q_sol[k] = C->q_sol[k] - sink;
        C->cvm[k] = p->c_air + C->qv[k] * p->c_vap + C->q_liq[k] * C_LIQ + C->q_sol[k] * C_ICE;
        C->tz[k] = C->tz[k] - sink * C->lhi[k] / C->cvm[k];
      }
    }
  }
  /* heights of the layer bottoms (:756-765) */
  C->ze[nz - 1] = p->zs - C->dz1[nz - 1];
  for (int k = nz - 2; k >= 0; --k) C->ze[k] = C->ze[k + 1] - C->dz1[k];

  double i1 = 0.0, s1 = 0.0, g1 = 0.0;
  int* no_fall = C->no_fall;
  /* cloud ice (:774-900) */
  column_no_fall(nz, C->qi, no_fall);
  if (!(n->vi_fac < 1.0e-5) && no_fall[0] == 0)
    i1 = implicit_fall(C, nz, p->dts, C->vti, C->qi, C->m1_sol, p->do_sedi_w, p->zs);
  /* snow (:902-1035) */
  column_no_fall(nz, C->qs, no_fall);
  if (no_fall[0] == 0) {
    s1 = implicit_fall(C, nz, p->dts, C->vtsz, C->qs, C->mflux, p->do_sedi_w, p->zs);
    for (int k = 0; k < nz; ++k) C->m1_sol[k] = C->m1_sol[k] + C->mflux[k];
  }
  /* graupel (:1037-1168) */
  column_no_fall(nz, C->qg, no_fall);
  if (no_fall[0] == 0) {
    g1 = implicit_fall(C, nz, p->dts, C->vtgz, C->qg, C->mflux, p->do_sedi_w, p->zs);
    for (int k = 0; k < nz; ++k) C->m1_sol[k] = C->m1_sol[k] + C->mflux[k];
  }
  for (int k = 0; k < nz; ++k) {
    C->rain[k] = C->rain[k] + 0.0; /* r1 = 0 (:924-925, :1171) */
    C->snow[k] = C->snow[k] + s1;
    C->graupel[k] = C->graupel[k] + g1;
    C->ice[k] = C->ice[k] + i1;
  }
  if (n->do_sedi_heat) sedi_heat(C, nz, C->m1_sol, C_ICE);
}

/* microphysics.py:1220-1449 */
static void icloud(const P_t* p, col_t* C, int nz) {
  const oracle_nml* n = &p->n;
  const double tice = n->tice, t_wfr = p->t_wfr;
  for (int k = 0; k < nz; ++k) {
    double lhi = LI00 + DC_ICE * C->tz[k];
    C->q_liq[k] = C->ql[k] + C->qr[k];
    C->q_sol[k] = C->qi[k] + C->qs[k] + C->qg[k];
    C->cvm[k] = p->c_air + C->qv[k] * p->c_vap + C->q_liq[k] * C_LIQ + C->q_sol[k] * C_ICE;
    double icpk = lhi / C->cvm[k];
    if ((C->tz[k] > tice) && (C->qi[k] > QCMIN)) {
      double melt = dmin(C->qi[k], p->fac_imlt * (C->tz[k] - tice) / icpk);
      double tmp = dmin(melt, dim(n->ql_mlt, C->ql[k]));
      C->ql[k] = C->ql[k] + tmp;
      C->qr[k] = C->qr[k] + melt - tmp;
      C->qi[k] = C->qi[k] - melt;
      C->q_liq[k] = C->q_liq[k] + melt;
      C->q_sol[k] = C->q_sol[k] - melt;
      C->cvm[k] = p->c_air + C->qv[k] * p->c_vap + C->q_liq[k] * C_LIQ + C->q_sol[k] * C_ICE;
      C->tz[k] = C->tz[k] - melt * lhi / C->cvm[k];
    } else if ((C->tz[k] < t_wfr) && (C->ql[k] > QCMIN)) {
      double dtmp = t_wfr - C->tz[k];
      double factor = dmin(1.0, dtmp / DT_FR);
      double sink = dmin(C->ql[k] * factor, dtmp / icpk);
      double qi_crt = n->qi_gen * dmin(n->qi_lim, 0.1 * (tice - C->tz[k])) / C->den[k];
      double tmp = dmin(sink, dim(qi_crt, C->qi[k]));
      C->ql[k] = C->ql[k] - sink;
      C->qs[k] = C->qs[k] + sink - tmp;
      C->qi[k] = C->qi[k] + tmp;
      C->q_liq[k] = C->q_liq[k] - sink;
      C->q_sol[k] = C->q_sol[k] + sink;
      C->cvm[k] = p->c_air + C->qv[k] * p->c_vap + C->q_liq[k] * C_LIQ + C->q_sol[k] * C_ICE;
      C->tz[k] = C->tz[k] + sink * lhi / C->cvm[k];
    }
  }
  /* vertical subgrid variability of ice (:1342-1376); dl holds di */
  if (n->z_slope_ice) {
    C->dl[0] = 0.0;
    for (int k = 1; k < nz; ++k) C->dq[k] = 0.5 * (C->qi[k] - C->qi[k - 1]);
    for (int k = 1; k < nz - 1; ++k) {
      C->dl[k] = 0.5 * dmin(fabs(C->dq[k] + C->dq[k + 1]), 0.5 * C->qi[k]);
      if (C->dq[k] * C->dq[k + 1] <= 0.0) {
        if (C->dq[k] > 0.0)
          C->dl[k] = dmin(C->dl[k], dmin(C->dq[k], -C->dq[k + 1]));
        else
          C->dl[k] = 0.0;
      }
    }
    C->dl[nz - 1] = 0.0;
  }
  for (int k = 0; k < nz; ++k) {
    double di;
    if (n->z_slope_ice)
      di = dmax(C->dl[k], dmax(QVMIN, C->h_var[k] * C->qi[k]));
    else
      di = dmax(QVMIN, C->h_var[k] * C->qi[k]);
    cell_t c = {C->qaz[k], C->qg[k], C->qi[k], C->ql[k], C->qr[k], C->qs[k], C->qv[k], C->tz[k]};
    icloud_main(p, &c, C->h_var[k], C->rh_adj[k], C->rh_rain[k], C->den[k], C->denfac[k], C->vtgz[k],
                C->vtrz[k], C->vtsz[k], C->p1[k], di, C->q_liq[k], C->q_sol[k], C->cvm[k]);
    C->qaz[k] = c.qa;
    C->qg[k] = c.qg;
    C->qi[k] = c.qi;
    C->ql[k] = c.ql;
    C->qr[k] = c.qr;
    C->qs[k] = c.qs;
    C->qv[k] = c.qv;
    C->tz[k] = c.tz;
  }
}

/* fields_init (microphysics.py:27-276) */
static void fields_init(const P_t* p, col_t* C, int nz, const oracle_fields* f, size_t base, size_t col,
                        double timestep) {
  const oracle_nml* n = &p->n;
  const double land = f->land[col], area = f->area[col];
  for (int k = 0; k < nz; ++k) {
    size_t i = base + k;
    C->graupel[k] = C->rain[k] = C->snow[k] = C->ice[k] = 0.0;
    double qi = f->qi[i], qs = f->qs[i];
    if (n->de_ice) {
      double qio = qi - timestep * f->qi_dt[i];
      double qin = dmax(qio, n->qi0_max);
      if (qi > qin) {
        qs = qs + qi - qin;
        qi = qin;
        double dqi = (qin - qio) * p->rdt;
        f->qs_dt[i] = f->qs_dt[i] + f->qi_dt[i] - dqi;
        f->qi_dt[i] = dqi;
        f->qs[i] = qs; /* the reference updates state.qsnow / state.qice in place */
        f->qi[i] = qi;
      }
    }
    C->qi[k] = qi;
    C->qs[k] = qs;
    C->t0[k] = f->pt[i];
    C->tz[k] = C->t0[k];
    C->dp1[k] = f->delp[i];
    C->dp0[k] = C->dp1[k];
    C->qv[k] = f->qv[i];
    C->ql[k] = f->ql[i];
    C->qr[k] = f->qr[i];
    C->qg[k] = f->qg[i];
    C->dp1[k] = C->dp1[k] * (1.0 - C->qv[k]);
    double omq = C->dp0[k] / C->dp1[k];
    C->qv[k] = C->qv[k] * omq;
    C->ql[k] = C->ql[k] * omq;
    C->qr[k] = C->qr[k] * omq;
    C->qi[k] = C->qi[k] * omq;
    C->qs[k] = C->qs[k] * omq;
    C->qg[k] = C->qg[k] * omq;
    C->qa0[k] = f->qa[i];
    C->dz0[k] = f->delz[i];
    C->den0[k] = -C->dp1[k] / (GRAV * C->dz0[k]);
    C->p1[k] = C->den0[k] * RDGAS * C->t0[k];
    C->qv0[k] = C->qv[k];
    C->ql0[k] = C->ql[k];
    C->qr0[k] = C->qr[k];
    C->qi0[k] = C->qi[k];
    C->qs0[k] = C->qs[k];
    C->qg0[k] = C->qg[k];
    C->m1[k] = 0.0;
    C->u0[k] = f->ua[i];
    C->v0[k] = f->va[i];
    C->u1[k] = C->u0[k];
    C->v1[k] = C->v0[k];
    C->ccn[k] = (n->ccn_l * land + n->ccn_o * (1.0 - land)) * 1.0e6;
    C->w[k] = f->w ? f->w[i] : 0.0;
    C->qaz[k] = f->qaz[i];
  }
  if (p->use_ccn) {
    C->ccn[nz - 1] = C->ccn[nz - 1] * RDGAS * C->tz[nz - 1] / C->p1[nz - 1];
    for (int k = nz - 2; k >= 0; --k) C->ccn[k] = C->ccn[k + 1];
  }
  for (int k = 0; k < nz; ++k) {
    C->c_praut[k] = p->cpaut * pow(C->ccn[k] * RHOR, -1.0 / 3.0);
    double s_leng = sqrt(sqrt(area * 1.0e-10));
    double t_land = n->dw_land * s_leng;
    double t_ocean = n->dw_ocean * s_leng;
    double h_var = t_land * land + t_ocean * (1.0 - land);
    C->h_var[k] = dmin(0.2, dmax(0.01, h_var));
    C->rh_adj[k] = 1.0 - C->h_var[k] - n->rh_inc;
    C->rh_rain[k] = dmax(0.35, C->rh_adj[k] - n->rh_inr);
    if (n->fix_negative) {
      double cvm = p->c_air + C->qv[k] * p->c_vap + (C->qr[k] + C->ql[k]) * C_LIQ +
                   (C->qi[k] + C->qs[k] + C->qg[k]) * C_ICE;
      double lcpk = (p->lv00 + p->d0_vap * C->tz[k]) / cvm;
      double icpk = (LI00 + DC_ICE * C->tz[k]) / cvm;
      if (C->qi[k] < 0.0) {
        C->qs[k] = C->qs[k] + C->qi[k];
        C->qi[k] = 0.0;
      }
      if (C->qs[k] < 0.0) {
        C->qg[k] = C->qg[k] + C->qs[k];
        C->qs[k] = 0.0;
      }
      if (C->qg[k] < 0.0) {
        C->qr[k] = C->qr[k] + C->qg[k];
        C->tz[k] = C->tz[k] - C->qg[k] * icpk;
        C->qg[k] = 0.0;
      }
      if (C->qr[k] < 0.0) {
        C->ql[k] = C->ql[k] + C->qr[k];
        C->qr[k] = 0.0;
      }
      if (C->ql[k] < 0.0) {
        C->qv[k] = C->qv[k] + C->ql[k];
        C->tz[k] = C->tz[k] - C->ql[k] * lcpk;
        C->ql[k] = 0.0;
      }
    }
  }
  if (n->fix_negative) {
    for (int k = 1; k < nz; ++k)
      if (C->qv[k - 1] < 0.0) C->qv[k] = C->qv[k] + C->qv[k - 1] * C->dp1[k - 1] / C->dp1[k];
    for (int k = 0; k < nz - 1; ++k)
      if (C->qv[k] < 0.0) C->qv[k] = 0.0;
    if (nz >= 2) {
      const int kb = nz - 1, ka = nz - 2;
      if ((C->qv[kb] < 0.0) && (C->qv[ka] > 0.0)) {
        double dq_a = dmin(-C->qv[kb] * C->dp1[kb], C->qv[ka] * C->dp1[ka]);
        double dq_b = dmin(-C->qv[kb] * C->dp1[kb], C->qv[ka] * C->dp1[ka]);
        C->qv[ka] = C->qv[ka] - dq_a / C->dp1[ka];
        C->qv[kb] = C->qv[kb] + dq_b / C->dp1[kb];
      }
    }
  }
}

/* fields_update (microphysics.py:1452-1553) */
static void fields_update(const P_t* p, col_t* C, int nz, const oracle_fields* f, size_t base, size_t col) {
  const oracle_nml* n = &p->n;
  const double rdt = p->rdt;
  if (n->sedi_transport) {
    for (int k = 1; k < nz; ++k) {
      C->u1[k] = (C->dp0[k] * C->u1[k] + C->m1[k - 1] * C->u1[k - 1]) / (C->dp0[k] + C->m1[k - 1]);
      C->v1[k] = (C->dp0[k] * C->v1[k] + C->m1[k - 1] * C->v1[k - 1]) / (C->dp0[k] + C->m1[k - 1]);
    }
    for (int k = 1; k < nz; ++k) {
      size_t i = base + k;
      f->udt[i] = f->udt[i] + (C->u1[k] - C->u0[k]) * rdt;
      f->vdt[i] = f->vdt[i] + (C->v1[k] - C->v0[k]) * rdt;
    }
  }
  for (int k = 0; k < nz; ++k) {
    size_t i = base + k;
    double omq = C->dp1[k] / C->dp0[k];
    f->qv_dt[i] = f->qv_dt[i] + rdt * (C->qv[k] - C->qv0[k]) * omq;
    f->ql_dt[i] = f->ql_dt[i] + rdt * (C->ql[k] - C->ql0[k]) * omq;
    f->qr_dt[i] = f->qr_dt[i] + rdt * (C->qr[k] - C->qr0[k]) * omq;
    f->qi_dt[i] = f->qi_dt[i] + rdt * (C->qi[k] - C->qi0[k]) * omq;
    f->qs_dt[i] = f->qs_dt[i] + rdt * (C->qs[k] - C->qs0[k]) * omq;
    f->qg_dt[i] = f->qg_dt[i] + rdt * (C->qg[k] - C->qg0[k]) * omq;
    double cvm = p->c_air + C->qv[k] * p->c_vap + (C->qr[k] + C->ql[k]) * C_LIQ +
                 (C->qi[k] + C->qs[k] + C->qg[k]) * C_ICE;
    f->pt_dt[i] = f->pt_dt[i] + rdt * (C->tz[k] - C->t0[k]) * cvm / CP_AIR;
    if (n->do_qa)
      f->qa_dt[i] = 0.0;
    else
      f->qa_dt[i] = f->qa_dt[i] + rdt * (C->qaz[k] / p->ntimes - C->qa0[k]);
    f->qaz[i] = C->qaz[k];
    if (f->w) f->w[i] = C->w[k];
  }
  const double convt = 86400.0 * rdt * RGRAV;
  /* the accumulators are column-constant; the surface-layer value is exported */
  if (f->rain) f->rain[col] = C->rain[nz - 1] * convt;
  if (f->snow) f->snow[col] = C->snow[nz - 1] * convt;
  if (f->ice) f->ice[col] = C->ice[nz - 1] * convt;
  if (f->graupel) f->graupel[col] = C->graupel[nz - 1] * convt;
}

static void mp_column(const P_t* p, col_t* C, int nz, const oracle_fields* f, size_t col, double timestep) {
  const size_t base = col * (size_t)nz;
  if (p->n.reset_qaz)
    for (int k = 0; k < nz; ++k) f->qaz[base + k] = 0.0;
  fields_init(p, C, nz, f, base, col, timestep);
  for (int it = 0; it < p->ntimes; ++it) {
    warm_rain(p, C, nz, 1);
    sedimentation(p, C, nz);
    warm_rain(p, C, nz, 0);
    icloud(p, C, nz);
  }
  fields_update(p, C, nz, f, base, col);
}

/* ------------------------------------------------------------------------
 * Exported API.
 * ---------------------------------------------------------------------- */
int oracle_mp_derived(const oracle_nml* nml, double timestep, double* out, int nout) {
  P_t p;
  setup(&p, nml, timestep);
  double v[] = {p.c_air, p.c_vap, p.d0_vap, p.lv00, p.cpaut, p.so3, p.zs, p.log_10, p.tice0, p.t_wfr,
                p.fac_rc, p.cracs, p.csacr, p.cgacr, p.cgacs,
                p.acco[0][0], p.acco[0][1], p.acco[0][2], p.acco[0][3],
                p.acco[1][0], p.acco[1][1], p.acco[1][2], p.acco[1][3],
                p.acco[2][0], p.acco[2][1], p.acco[2][2], p.acco[2][3],
                p.csacw, p.csaci, p.cgacw, p.cgaci, p.cracw,
                p.cssub[0], p.cssub[1], p.cssub[2], p.cssub[3], p.cssub[4],
                p.cgsub[0], p.cgsub[1], p.cgsub[2], p.cgsub[3], p.cgsub[4],
                p.crevp[0], p.crevp[1], p.crevp[2], p.crevp[3], p.crevp[4],
                p.cgfr[0], p.cgfr[1],
                p.csmlt[0], p.csmlt[1], p.csmlt[2], p.csmlt[3], p.csmlt[4],
                p.cgmlt[0], p.cgmlt[1], p.cgmlt[2], p.cgmlt[3], p.cgmlt[4],
                p.ces0, timestep, p.mpdt, p.rdt, p.dts, p.dt_rain, p.rdts, p.dt_evap,
                p.fac_i2s, p.fac_g2v, p.fac_v2g, p.fac_imlt, p.fac_l2v, (double)p.ntimes};
  int nv = (int)(sizeof(v) / sizeof(v[0]));
  for (int i = 0; i < nv && i < nout; ++i) out[i] = v[i];
  return nv;
}

/* Point-wise helpers exported for the known-answer tests. */
double oracle_wqs1(double ta, double den) { return wqs1(ta, den); }
double oracle_iqs1(double ta, double den) { return iqs1(ta, den); }
double oracle_wqs2(double ta, double den, double* d) { return wqs2(ta, den, d); }
double oracle_iqs2(double ta, double den, double* d) { return iqs2(ta, den, d); }

int oracle_mp_run(int64_t ncol, int nz, const oracle_nml* nml, double timestep, const oracle_fields* f,
                  int nthreads) {
  if (nz < 2 || nz > NZMAX) return -2;
  if (nml->use_ppm || nml->prog_ccn) return -3;
  if (!(timestep > 0.0)) return -7;
  P_t p;
  setup(&p, nml, timestep);
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
  {
    col_t* C = (col_t*)malloc(sizeof(col_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (int64_t col = 0; col < ncol; ++col) mp_column(&p, C, nz, f, (size_t)col, timestep);
    free(C);
  }
  (void)nthreads;
  return 0;
}

/* One implicit upwind fall of a single column, with the layer-bottom heights
 * built from dz1 as the stencils do; exported for the closed-form test.
 * q (dry mixing ratio) is updated in place, flux gets the mass flux through
 * each layer bottom.  Returns the surface flux. */
double oracle_fall_column(int nz, double dt, const double* dz1, const double* dp1, const double* vt, double* q,
                          double* flux) {
  if (nz < 2 || nz > NZMAX) return NAN;
  col_t* C = (col_t*)calloc(1, sizeof(col_t));
  for (int k = 0; k < nz; ++k) {
    C->dz1[k] = dz1[k];
    C->dp1[k] = dp1[k];
  }
  C->ze[nz - 1] = 0.0 - C->dz1[nz - 1];
  for (int k = nz - 2; k >= 0; --k) C->ze[k] = C->ze[k + 1] - C->dz1[k];
  double s = implicit_fall(C, nz, dt, vt, q, flux, 0, 0.0);
  free(C);
  return s;
}
