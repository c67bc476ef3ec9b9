// mp_forward.hip -- the fused GFDL microphysics forward kernel for gfx950.
//
// One launch = one Microphysics.__call__(state, timestep)
// (pySHiELD/stencils/microphysics.py:1981-2263): fields_init, ntimes x
// [warm_rain(first), sedimentation, warm_rain(second), icloud], then
// fields_update.  The reference runs five GT4Py stencils (~100 generated
// kernels per sub-step, 43 persistent 3D scratch fields); here every
// intermediate lives in registers and LDS, so HBM sees only the compulsory
// traffic: 13 FP64 fields read and 11 written per cell (192 B/cell at the
// default namelist) plus 6 FP64 values per column.
//
// Layout: see mp_wave.h.  One wavefront (= one workgroup) owns `ncols`
// consecutive columns spread over S slots of 64 lanes.
#include "mp_process.h"
#include "mp_wave.h"

namespace shield {

enum { CV_HVAR, CV_RHADJ, CV_RHRAIN, CV_CCN, CV_CPRAUT, CV_RAIN, CV_SNOW, CV_ICE, CV_GRAUPEL, CV_N };
constexpr int PMAX = 64;  // max columns per wavefront

template <int S, class R>
struct St {
  double qv[S], ql[S], qr[S], qi[S], qs[S], qg[S], tz[S];
  double t0[S], dz0[S], dp1[S], dp0[S], den[S], denfac[S], dz1[S];
  double m1[S], m1_sol[S], qaz[S], w[S];
  R vtr[S], vtg[S], vts[S];
};

template <int S>
struct Shared {
  double buf[64 * S];
  double buf2[64 * S];
  double colv[CV_N][PMAX];
};

template <int S, class R>
__device__ __forceinline__ Cell load_cell(const St<S, R>& X, int s) {
  Cell c;
  c.qv = X.qv[s];
  c.ql = X.ql[s];
  c.qr = X.qr[s];
  c.qi = X.qi[s];
  c.qs = X.qs[s];
  c.qg = X.qg[s];
  c.tz = X.tz[s];
  c.qa = X.qaz[s];
  return c;
}
template <int S, class R>
__device__ __forceinline__ void store_cell(St<S, R>& X, int s, const Cell& c) {
  X.qv[s] = c.qv;
  X.ql[s] = c.ql;
  X.qr[s] = c.qr;
  X.qi[s] = c.qi;
  X.qs[s] = c.qs;
  X.qg[s] = c.qg;
  X.tz[s] = c.tz;
  X.qaz[s] = c.qa;
}

// den0 = -dp1 / (g dz0) (microphysics.py:149), recomputed where needed
__device__ __forceinline__ double dry_density(double dp1, double dz0) { return -dp1 / (GRAV * dz0); }

// Layer-bottom heights ze (BACKWARD, microphysics.py:373-380 / :756-765) and
// layer thicknesses dz = ze[k] - ze[k+1], dz[nz-1] = ze - zs (:445-452).
template <int S, class R>
__device__ __forceinline__ void layer_dz(const Lay<S>& L, const St<S, R>& X, Shared<S>& sh, double zs,
                                         double (&dz)[S]) {
  double b[S], ze[S], zn[S];
#pragma unroll
  for (int s = 0; s < S; ++s) b[s] = (L.kb[s] == 0) ? zs - X.dz1[s] : -X.dz1[s];
  sum_bwd(L, b, ze);
  col_next(sh.buf, ze, zn);
#pragma unroll
  for (int s = 0; s < S; ++s) dz[s] = (L.kb[s] == 0) ? ze[s] - zs : ze[s] - zn[s];
}

// Implicit upwind fall of one species over dt with speeds vt (the pattern
// shared by microphysics.py:454-515, :825-900, :956-1035, :1089-1168).
// q holds dry mixing ratios.  Only columns with fall[s] change; flux gets the
// mass flux through each layer bottom (0 in non-falling columns) and surf
// the surface flux of the cell's column.  w is transported when do_sedi_w.
template <int S, class R>
__device__ __forceinline__ void implicit_fall(const Lay<S>& L, St<S, R>& X, Shared<S>& sh, const double (&dz)[S],
                                              double (&q)[S], const R (&vt)[S], double dt, const bool (&fall)[S],
                                              bool do_sedi_w, double (&flux)[S], double (&surf)[S]) {
  double dm[S], dd[S], qq[S], a[S], b[S], qm[S], ddp[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (do_sedi_w)
      dm[s] = X.dp1[s] * (1.0 + X.qv[s] + X.ql[s] + X.qr[s] + X.qi[s] + X.qs[s] + X.qg[s]);
    dd[s] = dt * double(vt[s]);
    qq[s] = q[s] * X.dp1[s];
  }
  col_prev(sh.buf, dd, ddp);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const double r = 1.0 / (dz[s] + dd[s]);
    a[s] = ddp[s] * r;
    b[s] = qq[s] * r;
  }
  scan_fwd(L, a, b, qm);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    qm[s] = qm[s] * dz[s];
    b[s] = qq[s] - qm[s];
  }
  sum_fwd(L, b, flux);
  col_bottom(L, sh.buf, flux, surf);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (fall[s]) {
      q[s] = qm[s] / X.dp1[s];
    } else {
      flux[s] = 0.0;
      surf[s] = 0.0;
    }
  }
  if (do_sedi_w) {
    double fp[S], vp[S], vd[S];
#pragma unroll
    for (int s = 0; s < S; ++s) vd[s] = double(vt[s]);
    col_prev(sh.buf, flux, fp);
    col_prev(sh.buf, vd, vp);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (!fall[s]) continue;
      if (L.k[s] == 0)
        X.w[s] = (dm[s] * X.w[s] + flux[s] * vd[s]) / (dm[s] - flux[s]);
      else
        X.w[s] = (dm[s] * X.w[s] - fp[s] * vp[s] + flux[s] * vd[s]) / (dm[s] + fp[s] - flux[s]);
    }
  }
}

// Sedimentation heating, explicit at the top and implicit below
// (microphysics.py:516-557 with cq = C_LIQ, :1176-1217 with cq = C_ICE).
template <int S, class R>
__device__ __forceinline__ void sedi_heat(const Lay<S>& L, St<S, R>& X, Shared<S>& sh, const double (&flux)[S],
                                          double cq, const bool (&apply)[S]) {
  double fp[S], a[S], b[S], tn[S];
  col_prev(sh.buf, flux, fp);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const double dgz = -0.5 * GRAV * X.dz1[s];
    const double cvn = X.dp1[s] * (CV_AIR + X.qv[s] * CV_VAP + (X.qr[s] + X.ql[s]) * C_LIQ +
                                   (X.qi[s] + X.qs[s] + X.qg[s]) * C_ICE);
    if (L.k[s] == 0) {
      a[s] = 0.0;
      b[s] = X.tz[s] + flux[s] * dgz / (cvn + flux[s] * cq);
    } else {
      const double r = 1.0 / (cvn + cq * flux[s]);
      a[s] = fp[s] * cq * r;
      b[s] = ((cvn + cq * (flux[s] - fp[s])) * X.tz[s] + dgz * (fp[s] + flux[s])) * r;
    }
  }
  scan_fwd(L, a, b, tn);
#pragma unroll
  for (int s = 0; s < S; ++s)
    if (apply[s]) X.tz[s] = tn[s];
}

// Vertical subgrid variability limiter (microphysics.py:595-620 for liquid,
// :1342-1367 for ice): returns dl (0 at the top and surface layers).
template <int S>
__device__ __forceinline__ void slope_limiter(const Lay<S>& L, Shared<S>& sh, const double (&q)[S],
                                              double (&dl)[S]) {
  double qp[S], dq[S], dqn[S];
  col_prev(sh.buf, q, qp);
#pragma unroll
  for (int s = 0; s < S; ++s) dq[s] = 0.5 * (q[s] - qp[s]);
  col_next(sh.buf, dq, dqn);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (L.k[s] == 0 || L.kb[s] == 0) {
      dl[s] = 0.0;
    } else {
      double d = 0.5 * mmin(fabs(dq[s] + dqn[s]), 0.5 * q[s]);
      if (dq[s] * dqn[s] <= 0.0) {
        if (dq[s] > 0.0)
          d = mmin(d, mmin(dq[s], -dqn[s]));
        else
          d = 0.0;
      }
      dl[s] = d;
    }
  }
}

// ------------------------------------------------------ warm_rain (:279-653)
template <int S, class R>
__device__ __forceinline__ void warm_rain(const Lay<S>& L, St<S, R>& X, Shared<S>& sh, const shield_mp_params& P,
                                          const Coef<R>& c, bool is_first) {
  const int nz = L.nz;
  (void)nz;
  const double t_wfr = P.t_wfr;
  if (is_first) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      X.dz1[s] = X.dz0[s] * X.tz[s] / X.t0[s];
      X.den[s] = dry_density(X.dp1[s], X.dz0[s]) * X.dz0[s] / X.dz1[s];
      X.denfac[s] = sqrt(SFCRHO / X.den[s]);
    }
  }
  bool pred[S], fall[S];
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.qr[s] > QRMIN);
  col_any(L, pred, fall);
#pragma unroll
  for (int s = 0; s < S; ++s) X.vtr[s] = rain_fspeed<R>(c, !fall[s], R(X.qr[s]), R(X.den[s]));
  double flux[S], surf[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    flux[s] = 0.0;
    surf[s] = 0.0;
  }
  const R dt5 = R(0.5) * c.dt_rain;
  if (wave_any(fall)) {
    double dz[S];
    layer_dz(L, X, sh, P.zs, dz);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (fall[s]) {
        Cell cs = load_cell(X, s);
        revap_racc<R>(c, t_wfr, dt5, R(sh.colv[CV_HVAR][L.cl[s]]), cs, R(X.den[s]), R(X.denfac[s]));
        store_cell(X, s, cs);
      }
    }
    implicit_fall(L, X, sh, dz, X.qr, X.vtr, P.dt_rain, fall, P.do_sedi_w != 0, flux, surf);
    if (P.nml.do_sedi_heat) sedi_heat(L, X, sh, flux, C_LIQ, fall);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (fall[s]) {
        Cell cs = load_cell(X, s);
        revap_racc<R>(c, t_wfr, dt5, R(sh.colv[CV_HVAR][L.cl[s]]), cs, R(X.den[s]), R(X.denfac[s]));
        store_cell(X, s, cs);
      }
    }
  }
  // auto-conversion (:587-645)
  if (P.nml.irain_f != 0) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      Cell cs = load_cell(X, s);
      autoconv_no_subgrid_var<R>(c, t_wfr, cs, R(X.den[s]), R(sh.colv[CV_CCN][L.cl[s]]),
                                 R(sh.colv[CV_CPRAUT][L.cl[s]]));
      store_cell(X, s, cs);
    }
  } else {
    double dl[S];
    if (P.nml.z_slope_liq) slope_limiter(L, sh, X.ql, dl);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const double hv = sh.colv[CV_HVAR][L.cl[s]];
      double d;
      if (P.nml.z_slope_liq)
        d = mmax(dl[s], mmax(QVMIN, hv * X.ql[s]));
      else
        d = mmax(QVMIN, hv * X.ql[s]);
      Cell cs = load_cell(X, s);
      autoconv_subgrid_var<R>(c, t_wfr, cs, R(X.den[s]), R(sh.colv[CV_CCN][L.cl[s]]),
                              R(sh.colv[CV_CPRAUT][L.cl[s]]), R(d));
      store_cell(X, s, cs);
    }
  }
  // rain += r1; m1 += m1_rain (+ m1_sol)  (:647-653)
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (L.v[s] && L.kb[s] == 0) sh.colv[CV_RAIN][L.cl[s]] += surf[s];
    X.m1[s] = is_first ? X.m1[s] + flux[s] : X.m1[s] + flux[s] + X.m1_sol[s];
  }
  __syncthreads();
}

// ---------------------------------------------------- sedimentation (:656-1217)
template <int S, class R>
__device__ __forceinline__ void sedimentation(const Lay<S>& L, St<S, R>& X, Shared<S>& sh,
                                              const shield_mp_params& P, const Coef<R>& c) {
  const double tice_d = P.nml.tice;
  R vti[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    fall_speed<R>(c, R(X.qg[s]), R(X.qi[s]), R(X.qs[s]), R(X.tz[s]), R(X.den[s]), X.vtg[s], vti[s], X.vts[s]);
    X.m1_sol[s] = 0.0;
  }
  // melting of cloud ice at and below the first level warmer than tice, and
  // always in the surface layer (:702-746)
  bool pred[S], stop[S];
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.tz[s] > tice_d);
  col_from_first(L, pred, stop);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const double tcd = X.tz[s] - tice_d;
    if (stop[s] && (X.qi[s] > QCMIN) && (tcd > 0.0)) {
      const R lhi = R(LI00) + R(DC_ICE) * R(X.tz[s]);
      R q_liq = R(X.ql[s]) + R(X.qr[s]);
      R q_sol = R(X.qi[s]) + R(X.qs[s]) + R(X.qg[s]);
      R cvm = c.c_air + R(X.qv[s]) * c.c_vap + q_liq * R(C_LIQ) + q_sol * R(C_ICE);
      const R icpk = lhi / cvm;
      const double sink = mmin(X.qi[s], double(c.fac_imlt * R(tcd) / icpk));
      const double tmp = mmin(sink, dim(double(P.nml.ql_mlt), X.ql[s]));
      X.ql[s] = X.ql[s] + tmp;
      X.qr[s] = X.qr[s] + sink - tmp;
      X.qi[s] = X.qi[s] - sink;
      q_liq = q_liq + R(sink);
      q_sol = q_sol - R(sink);
      cvm = c.c_air + R(X.qv[s]) * c.c_vap + q_liq * R(C_LIQ) + q_sol * R(C_ICE);
      X.tz[s] = X.tz[s] - double(R(sink) * lhi / cvm);
    }
  }
  double dz[S], flux[S], surf[S], rsurf[S];
  bool fall[S];
  layer_dz(L, X, sh, P.zs, dz);
  // cloud ice (:773-900)
#pragma unroll
  for (int s = 0; s < S; ++s) {
    pred[s] = L.v[s] && (X.qi[s] > QRMIN);
    rsurf[s] = 0.0;
  }
  col_any(L, pred, fall);
  if (!(P.nml.vi_fac < 1.0e-5) && wave_any(fall)) {
    implicit_fall(L, X, sh, dz, X.qi, vti, P.dts, fall, P.do_sedi_w != 0, flux, surf);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      X.m1_sol[s] = flux[s];
      if (L.v[s] && L.kb[s] == 0) sh.colv[CV_ICE][L.cl[s]] += surf[s];
    }
  }
  // snow (:902-1035)
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.qs[s] > QRMIN);
  col_any(L, pred, fall);
  if (wave_any(fall)) {
    implicit_fall(L, X, sh, dz, X.qs, X.vts, P.dts, fall, P.do_sedi_w != 0, flux, surf);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      X.m1_sol[s] = X.m1_sol[s] + flux[s];
      if (L.v[s] && L.kb[s] == 0) sh.colv[CV_SNOW][L.cl[s]] += surf[s];
    }
  }
  // graupel (:1037-1168)
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.qg[s] > QRMIN);
  col_any(L, pred, fall);
  if (wave_any(fall)) {
    implicit_fall(L, X, sh, dz, X.qg, X.vtg, P.dts, fall, P.do_sedi_w != 0, flux, surf);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      X.m1_sol[s] = X.m1_sol[s] + flux[s];
      if (L.v[s] && L.kb[s] == 0) sh.colv[CV_GRAUPEL][L.cl[s]] += surf[s];
    }
  }
  __syncthreads();
  if (P.nml.do_sedi_heat) {
    bool all[S];
#pragma unroll
    for (int s = 0; s < S; ++s) all[s] = L.v[s];
    sedi_heat(L, X, sh, X.m1_sol, C_ICE, all);
  }
}

// ------------------------------------------------------------ icloud (:1220-1449)
template <int S, class R>
__device__ __forceinline__ void icloud(const Lay<S>& L, St<S, R>& X, Shared<S>& sh, const shield_mp_params& P,
                                       const Coef<R>& c) {
  const double tice_d = P.nml.tice, t_wfr_d = P.t_wfr;
  R q_liq[S], q_sol[S], cvm[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const R lhi = R(LI00) + R(DC_ICE) * R(X.tz[s]);
    q_liq[s] = R(X.ql[s]) + R(X.qr[s]);
    q_sol[s] = R(X.qi[s]) + R(X.qs[s]) + R(X.qg[s]);
    cvm[s] = c.c_air + R(X.qv[s]) * c.c_vap + q_liq[s] * R(C_LIQ) + q_sol[s] * R(C_ICE);
    const R icpk = lhi / cvm[s];
    if ((X.tz[s] > tice_d) && (X.qi[s] > QCMIN)) {
      // pimlt: instant melting of cloud ice
      const double melt = mmin(X.qi[s], double(c.fac_imlt * R(X.tz[s] - tice_d) / icpk));
      const double tmp = mmin(melt, dim(double(P.nml.ql_mlt), X.ql[s]));
      X.ql[s] = X.ql[s] + tmp;
      X.qr[s] = X.qr[s] + melt - tmp;
      X.qi[s] = X.qi[s] - melt;
      q_liq[s] = q_liq[s] + R(melt);
      q_sol[s] = q_sol[s] - R(melt);
      cvm[s] = c.c_air + R(X.qv[s]) * c.c_vap + q_liq[s] * R(C_LIQ) + q_sol[s] * R(C_ICE);
      X.tz[s] = X.tz[s] - double(R(melt) * lhi / cvm[s]);
    } else if ((X.tz[s] < t_wfr_d) && (X.ql[s] > QCMIN)) {
      // pihom: homogeneous freezing of cloud water
      const R dtmp = R(t_wfr_d - X.tz[s]);
      const R factor = mmin(R(1.0), dtmp / R(DT_FR));
      const double sink = mmin(X.ql[s] * double(factor), double(dtmp / icpk));
      const R qi_crt = c.qi_gen * mmin(c.qi_lim, R(0.1) * R(tice_d - X.tz[s])) / R(X.den[s]);
      const double tmp = mmin(sink, dim(double(qi_crt), X.qi[s]));
      X.ql[s] = X.ql[s] - sink;
      X.qs[s] = X.qs[s] + sink - tmp;
      X.qi[s] = X.qi[s] + tmp;
      q_liq[s] = q_liq[s] - R(sink);
      q_sol[s] = q_sol[s] + R(sink);
      cvm[s] = c.c_air + R(X.qv[s]) * c.c_vap + q_liq[s] * R(C_LIQ) + q_sol[s] * R(C_ICE);
      X.tz[s] = X.tz[s] + double(R(sink) * lhi / cvm[s]);
    }
  }
  double dl[S];
  if (P.nml.z_slope_ice) slope_limiter(L, sh, X.qi, dl);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int cl = L.cl[s];
    const double hv = sh.colv[CV_HVAR][cl];
    double di;
    if (P.nml.z_slope_ice)
      di = mmax(dl[s], mmax(QVMIN, hv * X.qi[s]));
    else
      di = mmax(QVMIN, hv * X.qi[s]);
    const double p1 = dry_density(X.dp1[s], X.dz0[s]) * RDGAS * X.t0[s];
    Cell cs = load_cell(X, s);
    icloud_main<R>(c, tice_d, P.tice0, cs, R(hv), R(sh.colv[CV_RHADJ][cl]), R(sh.colv[CV_RHRAIN][cl]),
                   R(X.den[s]), R(X.denfac[s]), X.vtg[s], X.vtr[s], X.vts[s], p1, R(di), q_liq[s], q_sol[s],
                   cvm[s]);
    store_cell(X, s, cs);
  }
}

// ------------------------------------------------------------------ kernel
template <int S, class R>
__global__ __launch_bounds__(64) void mp_forward_kernel(shield_mp_shape sh3, shield_mp_params P, Coef<R> c,
                                                        shield_mp_fields f, int cols_per_wave) {
  __shared__ Shared<S> sh;
  const int lane = lane_id();
  const int nz = (int)sh3.nk;
  const int64_t ncol = sh3.ni * sh3.nj;
  const int64_t c0 = (int64_t)blockIdx.x * cols_per_wave;
  const int ncols = (int)((ncol - c0) < cols_per_wave ? (ncol - c0) : cols_per_wave);
  const shield_mp_namelist& n = P.nml;

  Lay<S> L;
  L.nz = nz;
  L.ncols = ncols;
  int64_t off[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int idx = 64 * s + lane;
    const int cl = idx / nz;
    const int k = idx - cl * nz;
    const bool v = cl < ncols;
    L.v[s] = v;
    L.k[s] = v ? k : 0;
    L.kb[s] = v ? nz - 1 - k : 0;
    L.cl[s] = v ? cl : 0;
    const int64_t col = c0 + L.cl[s];
    const int64_t i = col / sh3.nj, j = col - i * sh3.nj;
    off[s] = i * sh3.si + j * sh3.sj + L.k[s];
  }

  // ------------------------------------------- per-column scalars (:173-199)
  if (lane < ncols) {
    const int64_t col = c0 + lane;
    const int64_t i = col / sh3.nj, j = col - i * sh3.nj;
    const int64_t o2 = i * sh3.si2 + j * sh3.sj2;
    const double land = f.land[o2], area = f.area[o2];
    const double s_leng = sqrt(sqrt(area * 1.0e-10));
    const double t_land = n.dw_land * s_leng;
    const double t_ocean = n.dw_ocean * s_leng;
    double h_var = t_land * land + t_ocean * (1.0 - land);
    h_var = mmin(0.2, mmax(0.01, h_var));
    const double rh_adj = 1.0 - h_var - n.rh_inc;
    sh.colv[CV_HVAR][lane] = h_var;
    sh.colv[CV_RHADJ][lane] = rh_adj;
    sh.colv[CV_RHRAIN][lane] = mmax(0.35, rh_adj - n.rh_inr);
    sh.colv[CV_CCN][lane] = (n.ccn_l * land + n.ccn_o * (1.0 - land)) * 1.0e6;
    sh.colv[CV_RAIN][lane] = 0.0;
    sh.colv[CV_SNOW][lane] = 0.0;
    sh.colv[CV_ICE][lane] = 0.0;
    sh.colv[CV_GRAUPEL][lane] = 0.0;
  }
  __syncthreads();

  // ---------------------------------------------------- fields_init (:101-276)
  St<S, R> X;
  double qv0[S], ql0[S], qr0[S], qi0[S], qs0[S], qg0[S], qa0[S], u0[S], v0[S];
  double qi_dt_in[S], qs_dt_in[S];  // set only where de_ice rewrote them
  bool de_iced[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    de_iced[s] = false;
    qi_dt_in[s] = qs_dt_in[s] = 0.0;
    if (L.v[s]) {
      const int64_t o = off[s];
      double qi = f.qi[o], qs = f.qs[o];
      if (n.de_ice) {
        const double qi_dt = P.accumulate ? f.qi_dt[o] : 0.0;
        const double qs_dt = P.accumulate ? f.qs_dt[o] : 0.0;
        const double qio = qi - P.timestep * qi_dt;
        const double qin = mmax(qio, n.qi0_max);
        if (qi > qin) {
          qs = qs + qi - qin;
          qi = qin;
          const double dqi = (qin - qio) * P.rdt;
          // the de_ice tendencies are the starting values the tendency
          // update below accumulates onto
          qs_dt_in[s] = qs_dt + qi_dt - dqi;
          qi_dt_in[s] = dqi;
          f.qs[o] = qs;
          f.qi[o] = qi;
          de_iced[s] = true;
        }
      }
      X.t0[s] = f.pt[o];
      X.tz[s] = X.t0[s];
      X.dp0[s] = f.delp[o];
      const double qv = f.qv[o];
      X.dp1[s] = X.dp0[s] * (1.0 - qv);
      const double omq = X.dp0[s] / X.dp1[s];
      X.qv[s] = qv * omq;
      X.ql[s] = f.ql[o] * omq;
      X.qr[s] = f.qr[o] * omq;
      X.qi[s] = qi * omq;
      X.qs[s] = qs * omq;
      X.qg[s] = f.qg[o] * omq;
      qa0[s] = f.qa[o];
      X.dz0[s] = f.delz[o];
      u0[s] = f.ua[o];
      v0[s] = f.va[o];
      X.qaz[s] = n.reset_qaz ? 0.0 : f.qaz[o];
      X.w[s] = P.do_sedi_w ? f.w[o] : 0.0;
    } else {
      X.t0[s] = X.tz[s] = 250.0;
      X.dp0[s] = X.dp1[s] = 100.0;
      X.qv[s] = X.ql[s] = X.qr[s] = X.qi[s] = X.qs[s] = X.qg[s] = 0.0;
      qa0[s] = u0[s] = v0[s] = X.qaz[s] = X.w[s] = 0.0;
      X.dz0[s] = -100.0;
    }
    qv0[s] = X.qv[s];
    ql0[s] = X.ql[s];
    qr0[s] = X.qr[s];
    qi0[s] = X.qi[s];
    qs0[s] = X.qs[s];
    qg0[s] = X.qg[s];
    X.m1[s] = 0.0;
    X.m1_sol[s] = 0.0;
    X.den[s] = X.denfac[s] = X.dz1[s] = 1.0;
  }
  // ccn scaled by the surface-layer density, then c_praut (:174-186)
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (L.v[s] && L.kb[s] == 0) {
      const int cl = L.cl[s];
      double ccn = sh.colv[CV_CCN][cl];
      if (P.use_ccn) {
        const double p1 = dry_density(X.dp1[s], X.dz0[s]) * RDGAS * X.t0[s];
        ccn = ccn * RDGAS * X.tz[s] / p1;
      }
      sh.colv[CV_CCN][cl] = ccn;
      sh.colv[CV_CPRAUT][cl] = P.cpaut * pow(ccn * RHOR, -1.0 / 3.0);
    }
  }
  __syncthreads();
  if (n.fix_negative) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const double cvm = P.c_air + X.qv[s] * P.c_vap + (X.qr[s] + X.ql[s]) * C_LIQ +
                         (X.qi[s] + X.qs[s] + X.qg[s]) * C_ICE;
      const double lcpk = (P.lv00 + P.d0_vap * X.tz[s]) / cvm;
      const double icpk = (LI00 + DC_ICE * X.tz[s]) / cvm;
      if (X.qi[s] < 0.0) {
        X.qs[s] = X.qs[s] + X.qi[s];
        X.qi[s] = 0.0;
      }
      if (X.qs[s] < 0.0) {
        X.qg[s] = X.qg[s] + X.qs[s];
        X.qs[s] = 0.0;
      }
      if (X.qg[s] < 0.0) {
        X.qr[s] = X.qr[s] + X.qg[s];
        X.tz[s] = X.tz[s] - X.qg[s] * icpk;
        X.qg[s] = 0.0;
      }
      if (X.qr[s] < 0.0) {
        X.ql[s] = X.ql[s] + X.qr[s];
        X.qr[s] = 0.0;
      }
      if (X.ql[s] < 0.0) {
        X.qv[s] = X.qv[s] + X.ql[s];
        X.tz[s] = X.tz[s] - X.ql[s] * lcpk;
        X.ql[s] = 0.0;
      }
    }
    // vapour borrowing is a non-linear recurrence: one lane per column walks
    // its column in LDS (:244-276)
#pragma unroll
    for (int s = 0; s < S; ++s) {
      sh.buf[64 * s + lane] = X.qv[s];
      sh.buf2[64 * s + lane] = X.dp1[s];
    }
    __syncthreads();
    if (lane < ncols) {
      double* q = sh.buf + lane * nz;
      const double* dp = sh.buf2 + lane * nz;
      for (int k = 1; k < nz; ++k)
        if (q[k - 1] < 0.0) q[k] = q[k] + q[k - 1] * dp[k - 1] / dp[k];
      for (int k = 0; k < nz - 1; ++k)
        if (q[k] < 0.0) q[k] = 0.0;
      const int kb = nz - 1, ka = nz - 2;
      if ((q[kb] < 0.0) && (q[ka] > 0.0)) {
        const double dq = mmin(-q[kb] * dp[kb], q[ka] * dp[ka]);
        q[ka] = q[ka] - dq / dp[ka];
        q[kb] = q[kb] + dq / dp[kb];
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S; ++s) X.qv[s] = sh.buf[64 * s + lane];
    __syncthreads();
  }

  // --------------------------------------------------------- sub-steps (:2045)
  for (int it = 0; it < P.ntimes; ++it) {
    warm_rain(L, X, sh, P, c, true);
    sedimentation(L, X, sh, P, c);
    warm_rain(L, X, sh, P, c, false);
    icloud(L, X, sh, P, c);
  }

  // ---------------------------------------------------- fields_update (:1452-1553)
  const double rdt = P.rdt;
  if (!n.sedi_transport && !P.accumulate) {
#pragma unroll
    for (int s = 0; s < S; ++s)
      if (L.v[s]) {
        f.udt[off[s]] = 0.0;
        f.vdt[off[s]] = 0.0;
      }
  }
  if (n.sedi_transport) {
    double mp[S], a[S], bu[S], bv[S], u1[S], v1[S];
    col_prev(sh.buf, X.m1, mp);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (L.k[s] == 0) {
        a[s] = 0.0;
        bu[s] = u0[s];
        bv[s] = v0[s];
      } else {
        const double r = 1.0 / (X.dp0[s] + mp[s]);
        a[s] = mp[s] * r;
        bu[s] = X.dp0[s] * u0[s] * r;
        bv[s] = X.dp0[s] * v0[s] * r;
      }
    }
    scan_fwd(L, a, bu, u1);
    scan_fwd(L, a, bv, v1);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (L.v[s] && L.k[s] > 0) {
        const int64_t o = off[s];
        f.udt[o] = (P.accumulate ? f.udt[o] : 0.0) + (u1[s] - u0[s]) * rdt;
        f.vdt[o] = (P.accumulate ? f.vdt[o] : 0.0) + (v1[s] - v0[s]) * rdt;
      } else if (L.v[s] && !P.accumulate) {
        f.udt[off[s]] = 0.0;
        f.vdt[off[s]] = 0.0;
      }
    }
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (!L.v[s]) continue;
    const int64_t o = off[s];
    const double omq = X.dp1[s] / X.dp0[s];
    const bool acc = P.accumulate != 0;
    const double qi_dt0 = de_iced[s] ? qi_dt_in[s] : (acc ? f.qi_dt[o] : 0.0);
    const double qs_dt0 = de_iced[s] ? qs_dt_in[s] : (acc ? f.qs_dt[o] : 0.0);
    f.qv_dt[o] = (acc ? f.qv_dt[o] : 0.0) + rdt * (X.qv[s] - qv0[s]) * omq;
    f.ql_dt[o] = (acc ? f.ql_dt[o] : 0.0) + rdt * (X.ql[s] - ql0[s]) * omq;
    f.qr_dt[o] = (acc ? f.qr_dt[o] : 0.0) + rdt * (X.qr[s] - qr0[s]) * omq;
    f.qi_dt[o] = qi_dt0 + rdt * (X.qi[s] - qi0[s]) * omq;
    f.qs_dt[o] = qs_dt0 + rdt * (X.qs[s] - qs0[s]) * omq;
    f.qg_dt[o] = (acc ? f.qg_dt[o] : 0.0) + rdt * (X.qg[s] - qg0[s]) * omq;
    const double cvm = P.c_air + X.qv[s] * P.c_vap + (X.qr[s] + X.ql[s]) * C_LIQ +
                       (X.qi[s] + X.qs[s] + X.qg[s]) * C_ICE;
    f.pt_dt[o] = (acc ? f.pt_dt[o] : 0.0) + rdt * (X.tz[s] - X.t0[s]) * cvm / CP_AIR;
    if (n.do_qa)
      f.qa_dt[o] = 0.0;
    else
      f.qa_dt[o] = (acc ? f.qa_dt[o] : 0.0) + rdt * (X.qaz[s] / P.ntimes - qa0[s]);
    f.qaz[o] = X.qaz[s];
    if (P.do_sedi_w) f.w[o] = X.w[s];
  }
  if (lane < ncols) {
    const int64_t col = c0 + lane;
    const int64_t i = col / sh3.nj, j = col - i * sh3.nj;
    const int64_t o2 = i * sh3.si2 + j * sh3.sj2;
    const double convt = 86400.0 * rdt * RGRAV;
    if (f.rain) f.rain[o2] = sh.colv[CV_RAIN][lane] * convt;
    if (f.snow) f.snow[o2] = sh.colv[CV_SNOW][lane] * convt;
    if (f.ice) f.ice[o2] = sh.colv[CV_ICE][lane] * convt;
    if (f.graupel) f.graupel[o2] = sh.colv[CV_GRAUPEL][lane] * convt;
  }
}

// ---------------------------------------------------------------- launcher
struct LayoutChoice {
  int S, P;
};

// slot counts with a compiled kernel instance
static const int kSlots[] = {1, 2, 5};

LayoutChoice choose_layout(int64_t nz) {
  LayoutChoice best{0, 0};
  double best_util = -1.0;
  for (int S : kSlots) {
    const int64_t P = (64 * S) / nz;
    if (P < 1) continue;
    const int64_t Pc = P > PMAX ? PMAX : P;
    const double util = double(Pc * nz) / double(64 * S);
    if (util > best_util + 1e-9) {
      best_util = util;
      best = {S, (int)Pc};
    }
  }
  return best;
}

template <int S, class R>
static hipError_t launch_s(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f, int P,
                           hipStream_t stream) {
  const int64_t ncol = sh.ni * sh.nj;
  const int64_t grid = (ncol + P - 1) / P;
  const Coef<R> c = make_coef<R>(p);
  hipLaunchKernelGGL((mp_forward_kernel<S, R>), dim3((unsigned)grid), dim3(64), 0, stream, sh, p, c, f, P);
  return hipGetLastError();
}

template <class R>
static hipError_t launch_r(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f,
                           LayoutChoice lc, hipStream_t stream) {
  switch (lc.S) {
    case 1: return launch_s<1, R>(sh, p, f, lc.P, stream);
    case 2: return launch_s<2, R>(sh, p, f, lc.P, stream);
    case 5: return launch_s<5, R>(sh, p, f, lc.P, stream);
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
