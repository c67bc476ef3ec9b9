// Host build of the hand-written reverses (mp_point_adj.h) beside the
// dual-number vector-Jacobian products of the same process code, and of the
// icloud block forward beside the generic process code; driven by
// tests/test_point_reverse.py.
#include "../../pyshield_amd/csrc/mp_point_adj.h"
using namespace shield;

static void generic_fwd(const Coef<double>& c, const shield_mp_params& P, double hv, double rh_adj, double rh_rain,
                        const double* in, Cell& s) {
  double q_liq, q_sol, cvm;
  icloud_melt_freeze<double>(c, P.nml.tice, P.t_wfr, P.nml.ql_mlt, s, in[0], q_liq, q_sol, cvm);
  const double di = mmax(QVMIN, hv * s.qi);
  icloud_main<double>(c, P.nml.tice, P.tice0, s, hv, rh_adj, rh_rain, in[0], in[1], in[2], in[3], in[4], in[5], di,
                      q_liq, q_sol, cvm);
}

extern "C" void icl_check(const shield_mp_params* Pp, int n, const double* cells, const double* ins,
                          const double* hvr, const double* ybs, double* out_dual, double* out_hand,
                          double* fwd_gen, double* fwd_blk) {
  const shield_mp_params& P = *Pp;
  const Coef<double> c = make_coef<double>(P);
  const Coef<Dual> cd0 = make_coef<Dual>(P);
  const double hv = hvr[0], rh_adj = hvr[1], rh_rain = hvr[2];
  for (int j = 0; j < n; ++j) {
    const double* z = cells + 8 * j;
    const double* in = ins + 6 * j;
    const double* yb = ybs + 8 * j;
    // generic forward
    Cell s{z[0], z[1], z[2], z[3], z[4], z[5], z[6], z[7]};
    generic_fwd(c, P, hv, rh_adj, rh_rain, in, s);
    double* fg = fwd_gen + 8 * j;
    fg[0] = s.qv; fg[1] = s.ql; fg[2] = s.qr; fg[3] = s.qi; fg[4] = s.qs; fg[5] = s.qg; fg[6] = s.tz; fg[7] = s.qa;
    // one dual pass per input: 0-6 the cell, 8 den, 9 denfac, 10 vtg, 11 vtr, 12 vts, 13 p1, 14 qi_gen
    double* od = out_dual + 15 * j;
    for (int i = 0; i < 15; ++i) {
      if (i == 7) { od[i] = yb[7]; continue; }
      Coef<Dual> cd = cd0;
      if (i == 14) cd.qi_gen.d = 1.0;
      auto sd = [&](double v, int q) { return Dual(v, i == q ? 1.0 : 0.0); };
      CellT<Dual> cs{sd(z[0], 0), sd(z[1], 1), sd(z[2], 2), sd(z[3], 3), sd(z[4], 4), sd(z[5], 5), sd(z[6], 6),
                     Dual(z[7])};
      const Dual den = sd(in[0], 8);
      Dual q_liq, q_sol, cvm;
      icloud_melt_freeze<Dual>(cd, P.nml.tice, P.t_wfr, P.nml.ql_mlt, cs, den, q_liq, q_sol, cvm);
      const Dual di = mmax(QVMIN, Dual(hv) * cs.qi);
      icloud_main<Dual>(cd, P.nml.tice, P.tice0, cs, Dual(hv), rh_adj, rh_rain, den, sd(in[1], 9), sd(in[2], 10),
                        sd(in[3], 11), sd(in[4], 12), sd(in[5], 13), di, q_liq, q_sol, cvm);
      od[i] = yb[0] * cs.qv.d + yb[1] * cs.ql.d + yb[2] * cs.qr.d + yb[3] * cs.qi.d + yb[4] * cs.qs.d +
              yb[5] * cs.qg.d + yb[6] * cs.tz.d + yb[7] * cs.qa.d;
    }
    // hand-written
    IcK k{c, P.nml.tice, P.t_wfr, P.nml.ql_mlt, P.tice0, hv, rh_adj, rh_rain, {in[0], in[1], in[2], in[3], in[4], in[5]}};
    double y[8];
    for (int q = 0; q < 8; ++q) y[q] = yb[q];
    IcInB ib{0, 0, 0, 0, 0, 0, 0};
    Cell o;
    icloud_vjp(k, Cell{z[0], z[1], z[2], z[3], z[4], z[5], z[6], z[7]}, y, ib, &o);
    double* oh = out_hand + 15 * j;
    for (int q = 0; q < 8; ++q) oh[q] = y[q];
    oh[8] = ib.den; oh[9] = ib.denfac; oh[10] = ib.vtg; oh[11] = ib.vtr; oh[12] = ib.vts; oh[13] = ib.p1;
    oh[14] = ib.qi_gen;
    double* fb = fwd_blk + 8 * j;
    fb[0] = o.qv; fb[1] = o.ql; fb[2] = o.qr; fb[3] = o.qi; fb[4] = o.qs; fb[5] = o.qg; fb[6] = o.tz; fb[7] = o.qa;
  }
}

// revap_racc: dual passes over qv..tz (0-6), den (8), denfac (9) against revap_vjp
extern "C" void revap_check(const shield_mp_params* Pp, int n, const double* cells, const double* ins,
                            double h_var, const double* ybs, double* out_dual, double* out_hand) {
  const shield_mp_params& P = *Pp;
  const Coef<double> c = make_coef<double>(P);
  const Coef<Dual> cd = make_coef<Dual>(P);
  const double dt5 = 0.5 * c.dt_rain;
  for (int j = 0; j < n; ++j) {
    const double* z = cells + 8 * j;
    const double* in = ins + 6 * j;
    const double* yb = ybs + 8 * j;
    double* od = out_dual + 10 * j;
    for (int i = 0; i < 10; ++i) {
      if (i == 7) { od[i] = yb[7]; continue; }
      auto sd = [&](double v, int q) { return Dual(v, i == q ? 1.0 : 0.0); };
      CellT<Dual> cs{sd(z[0], 0), sd(z[1], 1), sd(z[2], 2), sd(z[3], 3), sd(z[4], 4), sd(z[5], 5), sd(z[6], 6),
                     Dual(z[7])};
      revap_racc<Dual>(cd, P.t_wfr, Dual(dt5), Dual(h_var), cs, sd(in[0], 8), sd(in[1], 9));
      od[i] = yb[0] * cs.qv.d + yb[1] * cs.ql.d + yb[2] * cs.qr.d + yb[3] * cs.qi.d + yb[4] * cs.qs.d +
              yb[5] * cs.qg.d + yb[6] * cs.tz.d + yb[7] * cs.qa.d;
    }
    IcX xb{yb[0], yb[1], yb[2], yb[3], yb[4], yb[5], yb[6], yb[7], 0.0, 0.0, 0.0};
    double denb = 0.0, denfacb = 0.0;
    revap_vjp(c, P.t_wfr, dt5, h_var, Cell{z[0], z[1], z[2], z[3], z[4], z[5], z[6], z[7]}, in[0], in[1], xb, denb,
              denfacb);
    double* oh = out_hand + 10 * j;
    oh[0] = xb.qv; oh[1] = xb.ql; oh[2] = xb.qr; oh[3] = xb.qi; oh[4] = xb.qs; oh[5] = xb.qg; oh[6] = xb.tz;
    oh[7] = xb.qa; oh[8] = denb; oh[9] = denfacb;
  }
}

// autoconversion: dual passes over ql, qr, den, ccn, c_praut, dl, fac_rc against autoconv_vjp
extern "C" void autoconv_check(const shield_mp_params* Pp, int subgrid, int n, const double* v, const double* ybs,
                               double* out_dual, double* out_hand) {
  const shield_mp_params& P = *Pp;
  const Coef<double> c = make_coef<double>(P);
  for (int j = 0; j < n; ++j) {
    const double* x = v + 7 * j;  // tz, ql, qr, den, ccn, c_praut, dl
    const double* yb = ybs + 2 * j;
    double* od = out_dual + 7 * j;
    for (int i = 0; i < 7; ++i) {
      Coef<Dual> cd = make_coef<Dual>(P);
      if (i == 6) cd.fac_rc.d = 1.0;
      auto sd = [&](double w, int q) { return Dual(w, i == q ? 1.0 : 0.0); };
      CellT<Dual> cs{Dual(0.0), sd(x[1], 0), sd(x[2], 1), Dual(0.0), Dual(0.0), Dual(0.0), Dual(x[0]), Dual(0.0)};
      if (subgrid)
        autoconv_subgrid_var<Dual>(cd, P.t_wfr, cs, sd(x[3], 2), sd(x[4], 3), sd(x[5], 4), sd(x[6], 5));
      else
        autoconv_no_subgrid_var<Dual>(cd, P.t_wfr, cs, sd(x[3], 2), sd(x[4], 3), sd(x[5], 4));
      od[i] = yb[0] * cs.ql.d + yb[1] * cs.qr.d;
    }
    double qlb = yb[0], qrb = yb[1];
    AcB ab{0.0, 0.0, 0.0, 0.0, 0.0};
    autoconv_vjp(c, P.t_wfr, subgrid != 0, x[0], x[1], x[3], x[4], x[5], x[6], qlb, qrb, ab);
    double* oh = out_hand + 7 * j;
    oh[0] = qlb; oh[1] = qrb; oh[2] = ab.den; oh[3] = ab.ccn; oh[4] = ab.c_praut; oh[5] = ab.dl; oh[6] = ab.fac_rc;
  }
}

// fall_speed: dual passes over qg, qi, qs, tc, den, vs_fac, vg_fac, vi_fac against fall_speed_vjp;
// rain_fspeed: over qr, den, vr_fac against rain_fspeed_vjp (v: qg, qi, qs, tc, den, qr)
extern "C" void speed_check(const shield_mp_params* Pp, int n, const double* v, const double* ybs, double* out_dual,
                            double* out_hand) {
  const shield_mp_params& P = *Pp;
  const Coef<double> c = make_coef<double>(P);
  for (int j = 0; j < n; ++j) {
    const double* x = v + 6 * j;
    const double* yb = ybs + 4 * j;  // cotangents of vtg, vti, vts, vtr
    double* od = out_dual + 11 * j;
    for (int i = 0; i < 8; ++i) {
      Coef<Dual> cd = make_coef<Dual>(P);
      if (i == 5) cd.vs_fac.d = 1.0;
      if (i == 6) cd.vg_fac.d = 1.0;
      if (i == 7) cd.vi_fac.d = 1.0;
      auto sd = [&](double w, int q) { return Dual(w, i == q ? 1.0 : 0.0); };
      Dual vg, vi, vs;
      fall_speed<Dual>(cd, sd(x[0], 0), sd(x[1], 1), sd(x[2], 2), sd(x[3], 3), sd(x[4], 4), vg, vi, vs);
      od[i] = yb[0] * vg.d + yb[1] * vi.d + yb[2] * vs.d;
    }
    for (int i = 0; i < 3; ++i) {
      Coef<Dual> cd = make_coef<Dual>(P);
      if (i == 2) cd.vr_fac.d = 1.0;
      const Dual vr = rain_fspeed<Dual>(cd, false, Dual(x[5], i == 0 ? 1.0 : 0.0), Dual(x[4], i == 1 ? 1.0 : 0.0));
      od[8 + i] = yb[3] * vr.d;
    }
    FsB fb{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    fall_speed_vjp(c, x[0], x[1], x[2], x[3], x[4], yb[0], yb[1], yb[2], fb);
    double* oh = out_hand + 11 * j;
    oh[0] = fb.qg; oh[1] = fb.qi; oh[2] = fb.qs; oh[3] = fb.tc; oh[4] = fb.den; oh[5] = fb.vs_fac;
    oh[6] = fb.vg_fac; oh[7] = fb.vi_fac;
    double qrb = 0.0, denb = 0.0, vrb = 0.0;
    rain_fspeed_vjp(c, false, x[5], x[4], yb[3], qrb, denb, vrb);
    oh[8] = qrb; oh[9] = denb; oh[10] = vrb;
  }
}

// sedi_ice_melt: dual passes over qv..tz against sedi_ice_melt_vjp
extern "C" void melt_check(const shield_mp_params* Pp, int n, const double* cells, const double* ybs,
                           double* out_dual, double* out_hand) {
  const shield_mp_params& P = *Pp;
  const Coef<double> c = make_coef<double>(P);
  const Coef<Dual> cd = make_coef<Dual>(P);
  for (int j = 0; j < n; ++j) {
    const double* z = cells + 8 * j;
    const double* yb = ybs + 8 * j;
    double* od = out_dual + 7 * j;
    for (int i = 0; i < 7; ++i) {
      auto sd = [&](double w, int q) { return Dual(w, i == q ? 1.0 : 0.0); };
      CellT<Dual> cs{sd(z[0], 0), sd(z[1], 1), sd(z[2], 2), sd(z[3], 3), sd(z[4], 4), sd(z[5], 5), sd(z[6], 6),
                     Dual(z[7])};
      sedi_ice_melt<Dual>(cd, P.nml.tice, P.nml.ql_mlt, cs);
      od[i] = yb[0] * cs.qv.d + yb[1] * cs.ql.d + yb[2] * cs.qr.d + yb[3] * cs.qi.d + yb[4] * cs.qs.d +
              yb[5] * cs.qg.d + yb[6] * cs.tz.d;
    }
    IcX xb{yb[0], yb[1], yb[2], yb[3], yb[4], yb[5], yb[6], yb[7], 0.0, 0.0, 0.0};
    sedi_ice_melt_vjp(c, P.nml.tice, P.nml.ql_mlt, Cell{z[0], z[1], z[2], z[3], z[4], z[5], z[6], z[7]}, xb);
    double* oh = out_hand + 7 * j;
    oh[0] = xb.qv; oh[1] = xb.ql; oh[2] = xb.qr; oh[3] = xb.qi; oh[4] = xb.qs; oh[5] = xb.qg; oh[6] = xb.tz;
  }
}
