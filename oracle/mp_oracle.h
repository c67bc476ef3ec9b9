/*
 * mp_oracle.h -- TEST INFRASTRUCTURE.  Interface of the FP64 CPU restatement
 * of pySHiELD's GFDL cloud microphysics (gfdl_mp_oracle.c).  Independent of
 * the product's headers on purpose.
 */
#ifndef MP_ORACLE_H
#define MP_ORACLE_H
#include <stdint.h>

#define ORACLE_NZMAX 256

/* PhysicsConfig microphysics fields (pySHiELD/_config.py:21-109). */
typedef struct {
  double c_cracw, c_paut, c_pgacs, c_psaci, ccn_l, ccn_o;
  double vs_fac, vg_fac, vi_fac, vr_fac;
  double tau_imlt, tau_i2s, tau_g2v, tau_v2g, tau_l2v;
  double ql_mlt, qs_mlt, t_sub, qi_gen, qi_lim, qi0_max;
  double dw_ocean, dw_land, rh_inc, rh_inr;
  double qc_crt, qi0_crt, qs0_crt, rthresh, mp_time;
  double vg_max, vi_max, vr_max, vs_max;
  double tice, alin, clin;
  int32_t hydrostatic, do_qa, const_vg, const_vi, const_vr, const_vs, de_ice;
  int32_t rad_snow, rad_rain, do_sedi_heat, do_sedi_w, fast_sat_adj;
  int32_t fix_negative, irain_f, prog_ccn, sedi_transport, use_ppm;
  int32_t z_slope_ice, z_slope_liq;
  int32_t reset_qaz;
} oracle_nml;

/* Dense [ncol][nz] FP64 fields; 2D fields are [ncol]. */
typedef struct {
  const double *pt, *qv, *ql, *qr;
  double *qi, *qs; /* written back when de_ice, as the reference does */
  const double *qg, *qa, *ua, *va, *delp, *delz;
  double* w;
  const double *land, *area;
  double *pt_dt, *qv_dt, *ql_dt, *qr_dt, *qi_dt, *qs_dt, *qg_dt, *qa_dt, *udt, *vdt;
  double* qaz;
  double *rain, *snow, *ice, *graupel;
} oracle_fields;

int oracle_mp_run(int64_t ncol, int nz, const oracle_nml* nml, double timestep,
                  const oracle_fields* f, int nthreads);
int oracle_mp_derived(const oracle_nml* nml, double timestep, double* out, int nout);
double oracle_wqs1(double ta, double den);
double oracle_iqs1(double ta, double den);
double oracle_wqs2(double ta, double den, double* dqdt);
double oracle_iqs2(double ta, double den, double* dqdt);
double oracle_fall_column(int nz, double dt, const double* dz1, const double* dp1, const double* vt,
                          double* q, double* flux);

#endif
