// mp_forward.h -- the fused GFDL microphysics forward kernel for gfx950.
//
// One launch = one Microphysics.__call__(state, timestep)
// (pySHiELD/stencils/microphysics.py:1981-2263): fields_init, ntimes x
// [warm_rain(first), sedimentation, warm_rain(second), icloud], then
// fields_update.  The reference runs five GT4Py stencils (~100 generated
// kernels per sub-step, 43 persistent 3D scratch fields); here every
// intermediate lives in registers and LDS.  HBM sees the compulsory traffic
// (13 FP64 fields read and 11 written per cell at the default namelist, plus
// 6 FP64 values per column) and re-reads of inputs that are cheaper to fetch
// again than to hold: pt and delz in the sub-steps, and the state and winds
// in fields_update.
//
// Layout: see mp_wave.h.  One wavefront (= one workgroup) owns `ncols`
// consecutive columns spread over S slots of 64 lanes.  The kernel is
// instantiated per slot count in mp_forward_s<S>.hip.
#pragma once

#include "mp_process.h"
#include "mp_wave.h"

namespace shield {

enum { CV_HVAR, CV_RHADJ, CV_RHRAIN, CV_CCN, CV_CPRAUT, CV_RAIN, CV_SNOW, CV_ICE, CV_GRAUPEL, CV_N };
// per-cell values kept in LDS (each lane touches only its own cells)
enum { CL_M1, CL_M1SOL, CL_DZ1, CL_TMP, CL_DEN, CL_DENFAC, CL_N };

// Per-cell state that stays in registers for the whole call.
template <int S, class R>
struct St {
  double qv[S], ql[S], qr[S], qi[S], qs[S], qg[S], tz[S];  // prognostic (dry mixing ratios)
  double qaz[S];                                            // cloud-fraction accumulator
  double dp1[S];                                            // dry air mass (Pa)
  R vtr[S], vtg[S], vts[S];                                 // fall speeds for icloud
};

template <int S>
struct Shared {
  double buf[64 * S];          // neighbour exchange
  double cell[CL_N][64 * S];   // per-cell values, CL_*
  double colv[CV_N][pmax_for(S)];
  // this lane's cell of slot s in per-cell array `which`
  __device__ __forceinline__ double& at(int which, int s) { return cell[which][64 * s + lane_id()]; }
};

// launder() hides a value's origin from the optimiser (an empty asm that
// claims to rewrite it in place), so the value is materialised exactly as
// rounded and cannot be fused into a neighbouring multiply-add.  The asm is
// not volatile: it orders nothing, so loads around it are scheduled freely.
__device__ __forceinline__ double launder(double v) {
  asm("" : "+v"(v));
  return v;
}

// ------------------------------------------------------ kernel arguments
// The forward kernel takes one argument block.  Its field pointers are read
// where they are used, each with a scalar load from the kernarg segment.
// Argument values are otherwise all loaded at kernel entry and kept for the
// whole kernel; the ~30 pointers and the coefficients outgrow the scalar
// register file, and the compiler spills the excess to VGPR lanes and reloads
// each use with v_readlane, a VALU instruction.  The empty asm hides the
// offset's value, so no two reads are merged and none is hoisted.
template <class R>
struct FwdArgs {
  shield_mp_shape sh3;
  shield_mp_params P;
  Coef<R> c;
  shield_mp_fields f;
  shield_update_fields u;
  int cols_per_wave;
};

using gdouble = __attribute__((address_space(1))) double;
template <uint32_t OFF>
struct KPtr {
  __device__ __forceinline__ gdouble* get() const {
    uint32_t o = OFF;
    asm volatile("" : "+s"(o));
    const __attribute__((address_space(4))) char* base =
        (const __attribute__((address_space(4))) char*)__builtin_amdgcn_kernarg_segment_ptr();
    return *(gdouble* const __attribute__((address_space(4)))*)(base + o);
  }
  __device__ __forceinline__ gdouble& operator[](int64_t i) const { return get()[i]; }
  __device__ __forceinline__ explicit operator bool() const { return get() != nullptr; }
};
#define SH_FIELDS(X)                                                                                 \
  X(pt) X(qv) X(ql) X(qr) X(qi) X(qs) X(qg) X(qa) X(ua) X(va) X(delp) X(delz) X(w) X(land) X(area) X(pt_dt) \
  X(qv_dt) X(ql_dt) X(qr_dt) X(qi_dt) X(qs_dt) X(qg_dt) X(qa_dt) X(udt) X(vdt) X(qaz) X(rain) X(snow) X(ice)    \
  X(graupel)
// every field pointer, read once: the prologue and fields_update take them
// all at their start, so the empty asm of each read (which orders memory
// operations around it) precedes all their loads and stores
struct GFields {
#define SH_GF(n) gdouble* n;
  SH_FIELDS(SH_GF)
#undef SH_GF
};
// shield_mp_fields at kernarg offset B, member by member
template <uint32_t B>
struct KFields {
#define SH_KF(n) KPtr<B + (uint32_t)offsetof(shield_mp_fields, n)> n;
  SH_FIELDS(SH_KF)
#undef SH_KF
  __device__ __forceinline__ GFields all() const {
    GFields g;
#define SH_KA(n) g.n = n.get();
    SH_FIELDS(SH_KA)
#undef SH_KA
    return g;
  }
};

// The point-wise stages are unrolled over the S slots with static indices, so
// every slot's state stays in registers.
template <int S, class R>
__device__ __forceinline__ Cell cell_of(const St<S, R>& X, int s) {
  return Cell{X.qv[s], X.ql[s], X.qr[s], X.qi[s], X.qs[s], X.qg[s], X.tz[s], X.qaz[s]};
}
template <int S, class R>
__device__ __forceinline__ void set_cell(St<S, R>& X, int s, const Cell& c) {
  X.qv[s] = c.qv;
  X.ql[s] = c.ql;
  X.qr[s] = c.qr;
  X.qi[s] = c.qi;
  X.qs[s] = c.qs;
  X.qg[s] = c.qg;
  X.tz[s] = c.tz;
  X.qaz[s] = c.qa;
}

// Global element offsets of this wavefront's cells (the host guarantees
// every addressed element index fits in 31 bits).
template <int S>
struct Glob {
  int32_t off[S];
};

// den0 = -dp1 / (g dz0) (microphysics.py:149), recomputed where needed
template <class R>
__device__ __forceinline__ double dry_density(double dp1, double dz0) { return fdiv<R>(-dp1, GRAV * dz0); }

// Layer thicknesses dz = ze[k] - ze[k+1], dz[nz-1] = ze[nz-1] - zs
// (microphysics.py:445-452) of the layer-bottom heights ze, which the
// reference sums from the surface up, ze[k] = ze[k+1] - dz1[k] with
// ze[nz-1] = zs - dz1[nz-1] (BACKWARD, :373-380 / :756-765).  The difference
// of that sum is -dz1 itself; the reference's sum-then-difference differs
// from it by round-off only (~1e-14 relative at 100 m layers 20 km up).
template <int S>
__device__ __forceinline__ void layer_dz(const Lay<S>& L, Shared<S>& sh, double (&dz)[S]) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) dz[s] = -sh.cell[CL_DZ1][64 * s + lane];
}

// Implicit upwind fall of one species over dt with speeds vt (the pattern
// shared by microphysics.py:454-515, :825-900, :956-1035, :1089-1168).
// q holds dry mixing ratios.  Only columns with fall[s] change; flux gets the
// mass flux through each layer bottom (0 in non-falling columns) and surf
// the surface flux of the cell's column.  When do_sedi_w the vertical
// velocity, which stays in HBM, is transported with the falling mass.
template <int S, class R, class FS>
__device__ __forceinline__ void implicit_fall(const Lay<S>& L, const St<S, R>& X, Shared<S>& sh, const Glob<S>& G,
                                              const FS& f, const double (&dz)[S], double (&q)[S],
                                              const R (&vt)[S], double dt, const bool (&fall)[S], bool do_sedi_w,
                                              double (&flux)[S], double (&surf)[S]) {
  double dd[S], qq[S], a[S], b[S], qm[S], ddp[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    dd[s] = dt * double(vt[s]);
    qq[s] = q[s] * X.dp1[s];
  }
  col_prev(sh.buf, dd, ddp);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const double r = recip<R>(dz[s] + dd[s]);
    a[s] = ddp[s] * r;
    // FP64: the reference's division, exact where nothing falls in from above
    if constexpr (sizeof(R) == 8)
      b[s] = qq[s] / (dz[s] + dd[s]);
    else
      b[s] = qq[s] * r;
  }
  scan_fwd(L, a, b, qm);
  // The reference sums the flux out of each layer bottom from the top,
  // m1[k] = m1[k-1] + qq[k] - qm[k] dz[k] (:474-481).  By the recurrence,
  // qq[k] - qm[k] dz[k] = dd[k] qm[k] - dd[k-1] qm[k-1], so the sum
  // telescopes to m1[k] = dd[k] qm[k] (nothing enters the top layer); the two
  // forms differ by round-off only.
#pragma unroll
  for (int s = 0; s < S; ++s) {
    flux[s] = dd[s] * qm[s];
    qm[s] = qm[s] * dz[s];
  }
  col_bottom(L, sh.buf, flux, surf);
  if (do_sedi_w) {
    // dm uses the mixing ratios before the fall (:478-486)
    double fp[S], vp[S], vd[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      vd[s] = double(vt[s]);
      if (!fall[s]) flux[s] = 0.0;
    }
    col_prev(sh.buf, flux, fp);
    col_prev(sh.buf, vd, vp);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (!fall[s] || !L.v[s]) continue;
      const double dm = X.dp1[s] * (1.0 + X.qv[s] + X.ql[s] + X.qr[s] + X.qi[s] + X.qs[s] + X.qg[s]);
      const int32_t o = G.off[s];
      const double w = f.w[o];
      if (L.k[s] == 0)
        f.w[o] = fdiv<R>(dm * w + flux[s] * vd[s], dm - flux[s]);
      else
        f.w[o] = fdiv<R>(dm * w - fp[s] * vp[s] + flux[s] * vd[s], dm + fp[s] - flux[s]);
    }
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (fall[s]) {
      q[s] = fdiv<R>(qm[s], X.dp1[s]);
    } else {
      flux[s] = 0.0;
      surf[s] = 0.0;
    }
  }
}

// Sedimentation heating, explicit at the top and implicit below
// (microphysics.py:516-557 with cq = C_LIQ, :1176-1217 with cq = C_ICE).
template <int S, class R>
__device__ __forceinline__ void sedi_heat(const Lay<S>& L, St<S, R>& X, Shared<S>& sh, const double (&flux)[S],
                                          double cq, const bool (&apply)[S]) {
  // separately rounded products and sums, as the reference's: in a cell no
  // flux passes, (cvn tz) / cvn then rounds as it does there
#pragma clang fp contract(off)
  const int lane = lane_id();
  double fp[S], a[S], b[S], tn[S];
  col_prev(sh.buf, flux, fp);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const double dgz = -0.5 * GRAV * sh.cell[CL_DZ1][64 * s + lane];
    const double cvn = X.dp1[s] * (CV_AIR + X.qv[s] * CV_VAP + (X.qr[s] + X.ql[s]) * C_LIQ +
                                   (X.qi[s] + X.qs[s] + X.qg[s]) * C_ICE);
    if (L.k[s] == 0) {
      a[s] = 0.0;
      b[s] = X.tz[s] + fdiv<R>(flux[s] * dgz, cvn + flux[s] * cq);
    } else if constexpr (sizeof(R) == 8) {
      // FP64: the reference's division; with no flux through the cell it
      // returns (cvn tz) / cvn, which a reciprocal can miss by an ulp of tz
      const double d = cvn + cq * flux[s];
      a[s] = fp[s] * cq * recip<R>(d);
      b[s] = ((cvn + cq * (flux[s] - fp[s])) * X.tz[s] + dgz * (fp[s] + flux[s])) / d;
    } else {
      const double r = recip<R>(cvn + cq * flux[s]);
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

// Air density from hydrostatic balance (:331-341) with t0 and dz0 the call's
// input pt and delz: the layer thickness dz1, den and denfac of the sub-step,
// and icloud's p1 = den0 rdgas t0 (:149-151), into their LDS cells (p1 into
// CL_TMP, which icloud reads before it reuses the cell).
template <int S, class R>
__device__ __forceinline__ void air_density(const Lay<S>& L, const St<S, R>& X, Shared<S>& sh,
                                            const double (&t0)[S], const double (&dz0)[S]) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const double dz1 = fdiv<R>(dz0[s] * X.tz[s], t0[s]);
    sh.cell[CL_DZ1][64 * s + lane] = dz1;
    const double den0 = dry_density<R>(X.dp1[s], dz0[s]);
    const double den = fdiv<R>(den0 * dz0[s], dz1);
    sh.at(CL_DEN, s) = den;
    // denfac only enters the rate laws, so mixed mode takes it in FP32
    if constexpr (sizeof(R) == 4)
      sh.at(CL_DENFAC, s) = double(msqrt(R(SFCRHO) / R(den)));
    else
      sh.at(CL_DENFAC, s) = sqrt(SFCRHO / den);
    sh.cell[CL_TMP][64 * s + lane] = den0 * RDGAS * t0[s];
  }
}

// revap_racc on the cells of the columns with falling rain.  It changes only
// cells warmer than t_wfr with rain (its first branch), so those are packed
// into one slot when there are at most 64 of them, and it runs once there;
// qi, qs and qg enter only through their sum, which is packed in qi's place
// (the sum plus two zeros is the sum).  Otherwise it runs per slot.
template <int S, class R>
__device__ __forceinline__ void evaporation(const Lay<S>& L, St<S, R>& X, Shared<S>& sh, const Coef<R>& c,
                                            double t_wfr, R dt5, const bool (&fall)[S]) {
  bool pred[S];
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = fall[s] && (X.tz[s] > t_wfr) && (X.qr[s] > QRMIN);
  const Packed<S> pk = pack_cells(pred);
  if (pk.n <= 64) {
    if (pk.n == 0) return;
    double q_sol[S];
#pragma unroll
    for (int s = 0; s < S; ++s) q_sol[s] = X.qi[s] + X.qs[s] + X.qg[s];
    Cell cs{gather_packed(X.qv, pk.src), gather_packed(X.ql, pk.src), gather_packed(X.qr, pk.src),
            gather_packed(q_sol, pk.src), 0.0, 0.0, gather_packed(X.tz, pk.src), 0.0};
    const int lane = lane_id();
    if (lane < pk.n) {
      const int id = pk.src;
      // column of cell id: floor((id + 1/2) / nz) in FP32 is exact here
      const int cl = int((float(id) + 0.5f) * (1.0f / float(L.nz)));
      revap_racc<R>(c, t_wfr, dt5, R(sh.colv[CV_HVAR][cl]), cs, sh.cell[CL_DEN][id], R(sh.cell[CL_DENFAC][id]));
    }
    scatter_packed(X.qv, cs.qv, pk, pred);
    scatter_packed(X.ql, cs.ql, pk, pred);
    scatter_packed(X.qr, cs.qr, pk, pred);
    scatter_packed(X.tz, cs.tz, pk, pred);
    return;
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (fall[s]) {
      Cell cs = cell_of(X, s);
      revap_racc<R>(c, t_wfr, dt5, R(sh.colv[CV_HVAR][L.cl[s]]), cs, sh.at(CL_DEN, s), R(sh.at(CL_DENFAC, s)));
      set_cell(X, s, cs);
    }
  }
}

// ------------------------------------------------------ warm_rain (:279-653)
// `density`: re-read t0 and dz0 and recompute the air density (sub-steps
// after the first; the kernel computes it for the first from the inputs it
// has loaded).
template <int S, class R, class FS>
__device__ __forceinline__ void warm_rain(const Lay<S>& L, St<S, R>& X, Shared<S>& sh, const Glob<S>& G,
                                          const FS& f, const shield_mp_params& P, const Coef<R>& c,
                                          bool is_first, bool density) {
  const int lane = lane_id();
  const double t_wfr = P.t_wfr;
  if (is_first && density) {
    double t0[S], dz0[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      t0[s] = 250.0;
      dz0[s] = -100.0;
      if (L.v[s]) {
        t0[s] = f.pt[G.off[s]];
        dz0[s] = f.delz[G.off[s]];
      }
    }
    air_density(L, X, sh, t0, dz0);
  }
  bool pred[S], fall[S];
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.qr[s] > QRMIN);
  col_any(L, pred, fall);
#pragma unroll
  for (int s = 0; s < S; ++s) X.vtr[s] = rain_fspeed<R>(c, !fall[s], R(X.qr[s]), R(sh.at(CL_DEN, s)));
  double flux[S], surf[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    flux[s] = 0.0;
    surf[s] = 0.0;
  }
  const R dt5 = R(0.5) * c.dt_rain;
  if (wave_any(fall)) {
    evaporation(L, X, sh, c, t_wfr, dt5, fall);
    double dz[S];
    layer_dz(L, sh, dz);
    implicit_fall(L, X, sh, G, f, dz, X.qr, X.vtr, P.dt_rain, fall, P.do_sedi_w != 0, flux, surf);
    if (P.nml.do_sedi_heat) sedi_heat(L, X, sh, flux, C_LIQ, fall);
    evaporation(L, X, sh, c, t_wfr, dt5, fall);
  }
  // auto-conversion (:587-645)
  if (P.nml.irain_f != 0) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int cl = L.cl[s];
      Cell cs = cell_of(X, s);
      autoconv_no_subgrid_var<R>(c, t_wfr, cs, R(sh.at(CL_DEN, s)), R(sh.colv[CV_CCN][cl]),
                                 R(sh.colv[CV_CPRAUT][cl]));
      set_cell(X, s, cs);
    }
  } else {
    double dl[S];
    if (P.nml.z_slope_liq) slope_limiter(L, sh, X.ql, dl);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int cl = L.cl[s];
      const double hv = sh.colv[CV_HVAR][cl];
      Cell cs = cell_of(X, s);
      double d;
      if (P.nml.z_slope_liq)
        d = mmax(dl[s], mmax(QVMIN, hv * cs.ql));
      else
        d = mmax(QVMIN, hv * cs.ql);
      autoconv_subgrid_var<R>(c, t_wfr, cs, R(sh.at(CL_DEN, s)), R(sh.colv[CV_CCN][cl]),
                              R(sh.colv[CV_CPRAUT][cl]), R(d));
      set_cell(X, s, cs);
    }
  }
  // rain += r1; m1 += m1_rain (+ m1_sol)  (:647-653)
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (L.v[s] && L.kb[s] == 0) sh.colv[CV_RAIN][L.cl[s]] += surf[s];
    double& m1 = sh.cell[CL_M1][64 * s + lane];
    m1 = is_first ? m1 + flux[s] : m1 + flux[s] + sh.cell[CL_M1SOL][64 * s + lane];
  }
  __syncthreads();
}

// ---------------------------------------------------- sedimentation (:656-1217)
template <int S, class R, class FS>
__device__ __forceinline__ void sedimentation(const Lay<S>& L, St<S, R>& X, Shared<S>& sh, const Glob<S>& G,
                                              const FS& f, const shield_mp_params& P,
                                              const Coef<R>& c) {
  const int lane = lane_id();
  const double tice_d = P.nml.tice;
  R vti[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    R vg, vi, vs;
    fall_speed<R>(c, R(X.qg[s]), R(X.qi[s]), R(X.qs[s]), R(X.tz[s] - tice_d),
                  R(sh.at(CL_DEN, s)), vg, vi, vs);
    X.vtg[s] = vg;
    vti[s] = vi;
    X.vts[s] = vs;
  }
  // melting of cloud ice at and below the first level warmer than tice, and
  // always in the surface layer (:702-746)
  bool pred[S], stop[S];
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.tz[s] > tice_d);
  col_from_first(L, pred, stop);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (stop[s]) {
      Cell cs = cell_of(X, s);
      sedi_ice_melt<R>(c, tice_d, P.nml.ql_mlt, cs);
      set_cell(X, s, cs);
    }
  }
  double dz[S], flux[S], surf[S], m1_sol[S];
  bool fall[S];
#pragma unroll
  for (int s = 0; s < S; ++s) m1_sol[s] = 0.0;
  layer_dz(L, sh, dz);
  const bool sedi_w = P.do_sedi_w != 0;
  // cloud ice (:773-900)
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.qi[s] > QRMIN);
  col_any(L, pred, fall);
  if (!(P.nml.vi_fac < 1.0e-5) && wave_any(fall)) {
    implicit_fall(L, X, sh, G, f, dz, X.qi, vti, P.dts, fall, sedi_w, flux, surf);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      m1_sol[s] = flux[s];
      if (L.v[s] && L.kb[s] == 0) sh.colv[CV_ICE][L.cl[s]] += surf[s];
    }
  }
  // snow (:902-1035)
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.qs[s] > QRMIN);
  col_any(L, pred, fall);
  if (wave_any(fall)) {
    implicit_fall(L, X, sh, G, f, dz, X.qs, X.vts, P.dts, fall, sedi_w, flux, surf);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      m1_sol[s] = m1_sol[s] + flux[s];
      if (L.v[s] && L.kb[s] == 0) sh.colv[CV_SNOW][L.cl[s]] += surf[s];
    }
  }
  // graupel (:1037-1168)
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.qg[s] > QRMIN);
  col_any(L, pred, fall);
  if (wave_any(fall)) {
    implicit_fall(L, X, sh, G, f, dz, X.qg, X.vtg, P.dts, fall, sedi_w, flux, surf);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      m1_sol[s] = m1_sol[s] + flux[s];
      if (L.v[s] && L.kb[s] == 0) sh.colv[CV_GRAUPEL][L.cl[s]] += surf[s];
    }
  }
  __syncthreads();
  if (P.nml.do_sedi_heat) {
    bool all[S];
#pragma unroll
    for (int s = 0; s < S; ++s) all[s] = L.v[s];
    sedi_heat(L, X, sh, m1_sol, C_ICE, all);
  }
#pragma unroll
  for (int s = 0; s < S; ++s) sh.cell[CL_M1SOL][64 * s + lane] = m1_sol[s];
}

// ------------------------------------------------------------ icloud (:1220-1449)
// The point-wise processes of icloud run on the wavefront's cells in a sorted
// order: cells holding more than QRMIN of some condensate first, then those
// holding more than QCMIN, then the rest (and the lanes past the last
// column).  A branch costs the wavefront a pass per slot in which any lane
// takes it; most branches need condensate, and a column's cloud layers are a
// fifth of its levels spread over every slot, so sorted they run in about one
// slot instead of two or three.  The state is moved through LDS cells that are
// free during icloud and moved back after; every cell's arithmetic is the
// same as in cell order.  With the ice slope limiter the cells keep their
// order (the limiter's results are per slot).  SORT = false keeps cell order
// everywhere and reads p1's inputs from HBM (the adjoint's replay, whose
// warm_rain does not leave p1 in LDS).
template <int S, class R, class FS, bool SORT = true>
__device__ __forceinline__ void icloud(const Lay<S>& L, St<S, R>& X, Shared<S>& sh, const Glob<S>& G,
                                       const FS& f, const shield_mp_params& P, const Coef<R>& c) {
  const double tice_d = P.nml.tice, t_wfr_d = P.t_wfr, ql_mlt = P.nml.ql_mlt;
  const bool slope = P.nml.z_slope_ice != 0;
  const int lane = lane_id();
  // p1 = den0 rdgas t0 from the call's input pt, delz (:149-151), in cell
  // order: in the forward kernel from the LDS cell air_density left it in
  // (before the slope limiter's or the sort's use of that cell)
  double p1[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if constexpr (SORT) {
      p1[s] = sh.cell[CL_TMP][64 * s + lane];
    } else {
      const int32_t o = G.off[s];
      const bool v = L.v[s];
      const double t0 = v ? f.pt[o] : 250.0;
      const double dz0 = v ? f.delz[o] : -100.0;
      p1[s] = dry_density<R>(X.dp1[s], dz0) * RDGAS * t0;
    }
  }
  // With the ice slope limiter the whole column's melt / freeze must be done
  // before the limiter sees cloud ice; the heat-capacity terms it leaves are
  // parked in LDS cells that are free during icloud (m1_sol is consumed by
  // the second warm_rain, dz1 is rebuilt by the next sub-step).
  double* q_liq_l = sh.cell[CL_M1SOL];
  double* q_sol_l = sh.cell[CL_DZ1];
  double* cvm_l = sh.cell[CL_TMP];
  double dl[S];
  if (slope) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      Cell cs = cell_of(X, s);
      double q_liq, q_sol, cvm;
      icloud_melt_freeze<R>(c, tice_d, t_wfr_d, ql_mlt, cs, sh.at(CL_DEN, s), q_liq, q_sol, cvm);
      set_cell(X, s, cs);
      q_liq_l[64 * s + lane] = q_liq;
      q_sol_l[64 * s + lane] = q_sol;
      cvm_l[64 * s + lane] = cvm;
    }
    slope_limiter(L, sh, X.qi, dl);
  }
  int pos[S], src[S];
  const bool sorted = SORT && !slope;
  if (sorted) {
    int cls[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const double qm = mmax(mmax(mmax(X.ql[s], X.qr[s]), mmax(X.qi[s], X.qs[s])), X.qg[s]);
      cls[s] = !L.v[s] ? 2 : qm > QRMIN ? 0 : qm > QCMIN ? 1 : 2;
    }
    sort_cells3(cls, pos, src, reinterpret_cast<int*>(sh.buf));
    double* b1 = sh.cell[CL_M1SOL];
    double* b2 = sh.cell[CL_DZ1];
    double* b3 = sh.cell[CL_TMP];
    put_cells(X.tz, sh.buf); put_cells(X.qv, b1); put_cells(X.ql, b2); put_cells(X.qr, b3);
    __syncthreads();
    get_cells(X.tz, sh.buf, src); get_cells(X.qv, b1, src); get_cells(X.ql, b2, src); get_cells(X.qr, b3, src);
    __syncthreads();
    put_cells(X.qi, sh.buf); put_cells(X.qs, b1); put_cells(X.qg, b2); put_cells(X.qaz, b3);
    __syncthreads();
    get_cells(X.qi, sh.buf, src); get_cells(X.qs, b1, src); get_cells(X.qg, b2, src); get_cells(X.qaz, b3, src);
    __syncthreads();
    // the fall speeds are not used again before sedimentation recomputes them
    put_cells(p1, sh.buf);
    put_cells(X.vtg, reinterpret_cast<R*>(b1));
    put_cells(X.vtr, reinterpret_cast<R*>(b2));
    put_cells(X.vts, reinterpret_cast<R*>(b3));
    __syncthreads();
    get_cells(p1, sh.buf, src);
    get_cells(X.vtg, reinterpret_cast<const R*>(b1), src);
    get_cells(X.vtr, reinterpret_cast<const R*>(b2), src);
    get_cells(X.vts, reinterpret_cast<const R*>(b3), src);
    __syncthreads();
  } else {
#pragma unroll
    for (int s = 0; s < S; ++s) src[s] = pos[s] = 64 * s + lane;
  }
  const int ncell = L.ncols * L.nz;
  // column of cell id: floor((id + 1/2) / nz) in FP32 is exact here, the
  // quotient stays >= 1/(2 nz) away from an integer
  const float rnz = 1.0f / float(L.nz);
  for_slots<S>([&](auto si) __attribute__((always_inline)) {
    constexpr int s = decltype(si)::value;
    const int id = src[s];
    const int cl = id < ncell ? int((float(id) + 0.5f) * rnz) : 0;
    const double hv = sh.colv[CV_HVAR][cl];
    const double den = sh.cell[CL_DEN][id];
    Cell cs = cell_of(X, s);
    double q_liq, q_sol, cvm;
    double di;
    if (slope) {
      q_liq = q_liq_l[id];
      q_sol = q_sol_l[id];
      cvm = cvm_l[id];
      di = mmax(dl[s], mmax(QVMIN, hv * cs.qi));
    } else {
      icloud_melt_freeze<R>(c, tice_d, t_wfr_d, ql_mlt, cs, den, q_liq, q_sol, cvm);
      di = mmax(QVMIN, hv * cs.qi);
    }
    icloud_main<R>(c, tice_d, P.tice0, cs, R(hv), sh.colv[CV_RHADJ][cl], sh.colv[CV_RHRAIN][cl], den,
                   R(sh.cell[CL_DENFAC][id]), X.vtg[s], X.vtr[s], X.vts[s], p1[s], R(di), q_liq, q_sol,
                   cvm);
    set_cell(X, s, cs);
  });
  if (sorted) {
    double* b1 = sh.cell[CL_M1SOL];
    double* b2 = sh.cell[CL_DZ1];
    double* b3 = sh.cell[CL_TMP];
    put_cells(X.tz, sh.buf); put_cells(X.qv, b1); put_cells(X.ql, b2); put_cells(X.qr, b3);
    __syncthreads();
    get_cells(X.tz, sh.buf, pos); get_cells(X.qv, b1, pos); get_cells(X.ql, b2, pos); get_cells(X.qr, b3, pos);
    __syncthreads();
    put_cells(X.qi, sh.buf); put_cells(X.qs, b1); put_cells(X.qg, b2); put_cells(X.qaz, b3);
    __syncthreads();
    get_cells(X.qi, sh.buf, pos); get_cells(X.qs, b1, pos); get_cells(X.qg, b2, pos); get_cells(X.qaz, b3, pos);
    __syncthreads();
  }
}

// fields_init's fix_negative (microphysics.py:201-276).  At one cell: a
// negative species borrows from the next one (ice from snow, snow from
// graupel, graupel from rain with the fusion heat, rain from cloud water,
// cloud water from vapour with the condensation heat), the heats per unit
// mass formed from the entry state.
template <class R>
SH_FN void fix_negative_cell(const shield_mp_params& P, Cell& x) {
  const double cvm = P.c_air + x.qv * P.c_vap + (x.qr + x.ql) * C_LIQ + (x.qi + x.qs + x.qg) * C_ICE;
  const double tz0 = x.tz;
  if (x.qi < 0.0) {
    x.qs = x.qs + x.qi;
    x.qi = 0.0;
  }
  if (x.qs < 0.0) {
    x.qg = x.qg + x.qs;
    x.qs = 0.0;
  }
  if (x.qg < 0.0) {
    const double icpk = fdiv<R>(LI00 + DC_ICE * tz0, cvm);
    x.qr = x.qr + x.qg;
    x.tz = x.tz - x.qg * icpk;
    x.qg = 0.0;
  }
  if (x.qr < 0.0) {
    x.ql = x.ql + x.qr;
    x.qr = 0.0;
  }
  if (x.ql < 0.0) {
    const double lcpk = fdiv<R>(P.lv00 + P.d0_vap * tz0, cvm);
    x.qv = x.qv + x.ql;
    x.tz = x.tz - x.ql * lcpk;
    x.ql = 0.0;
  }
}
// Then the vapour of one column (q, dp: nz levels, top first): a negative
// layer borrows from the one below (a forward recurrence), negative vapour
// above the surface layer is cleared, and a negative surface layer borrows
// from the layer above.
SH_FN void fix_negative_column(double* q, const double* dp, int nz) {
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
// The wavefront's cells and columns.  The vapour borrowing is a non-linear
// recurrence: one lane per column walks its column in LDS, dp1 staged where
// m1 lives later.  It changes nothing unless some cell holds negative vapour.
template <int S, class R>
__device__ __forceinline__ void fix_negative(const Lay<S>& L, St<S, R>& X, Shared<S>& sh, const shield_mp_params& P) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    Cell cs = cell_of(X, s);
    fix_negative_cell<R>(P, cs);
    set_cell(X, s, cs);
  }
  bool neg = false;
#pragma unroll
  for (int s = 0; s < S; ++s) neg = neg || (L.v[s] && X.qv[s] < 0.0);
  if (__ballot(neg) != 0ull) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      sh.buf[64 * s + lane] = X.qv[s];
      sh.cell[CL_M1][64 * s + lane] = X.dp1[s];
    }
    __syncthreads();
    if (lane < L.ncols) fix_negative_column(sh.buf + lane * L.nz, sh.cell[CL_M1] + lane * L.nz, L.nz);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S; ++s) X.qv[s] = sh.buf[64 * s + lane];
    __syncthreads();
  }
}

// The state update that follows the microphysics for this wavefront's
// columns: update_physics_state_with_tendencies (physics.py:146-194),
// fill_gfs_delp on the updated vapour with the dycore delp
// (update_atmos_state.py:21-37) and prepare_tendencies_and_update_tracers
// (:40-92).  The tendencies were just written by this wavefront and are read
// back from the cache, so every expression is the one the separate kernels
// evaluate.
template <int S, class FS>
__device__ __forceinline__ void state_update(const Lay<S>& L, const Glob<S>& G, Shared<S>& sh, const shield_mp_shape& sh3,
                                             int64_t c0, const FS& f, const shield_update_fields& u) {
  const int lane = lane_id();
  const int nz = L.nz;
  const double* in[10] = {f.qv, f.ql, f.qr, f.qi, f.qs, f.qg, f.qa, f.pt, f.ua, f.va};
  const double* td[10] = {f.qv_dt, f.ql_dt, f.qr_dt, f.qi_dt, f.qs_dt, f.qg_dt, f.qa_dt, f.pt_dt, f.udt, f.vdt};
  // In memory phases: every load of a phase precedes its first store.  A
  // store may alias a later load as far as the compiler knows (the entry
  // point rejects calls where outputs alias), so interleaved per field the
  // loads and stores serialise one HBM round trip per field and slot.
  // forward Euler; the updated vapour and the dycore delp go to LDS, and the
  // updated values the tracer conversion reads stay in registers
  double xin[10][S], xtd[10][S], dp0[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int32_t o = G.off[s];
    dp0[s] = 1.0;
#pragma unroll
    for (int n = 0; n < 10; ++n) xin[n][s] = xtd[n][s] = 0.0;
    if (L.v[s]) {
#pragma unroll
      for (int n = 0; n < 10; ++n) {
        xin[n][s] = in[n][o];
        xtd[n][s] = td[n][o];
      }
      dp0[s] = u.delp[o];
    }
  }
  double x[10][S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
#pragma unroll
    for (int n = 0; n < 10; ++n) x[n][s] = xin[n][s] + xtd[n][s] * u.dt;
    if (L.v[s]) {
      const int32_t o = G.off[s];
      // updated[0] is written after the fill
#pragma unroll
      for (int n = 1; n < 10; ++n) u.updated[n][o] = x[n][s];
    }
    sh.buf[64 * s + lane] = L.v[s] ? x[0][s] : 0.0;
    sh.cell[CL_TMP][64 * s + lane] = dp0[s];
  }
  __syncthreads();
  // fill_gfs_delp: a non-linear vertical recurrence, one lane per column
  if (lane < L.ncols) fill_column(sh.buf + lane * nz, sh.cell[CL_TMP] + lane * nz, nz, u.q_min);
  __syncthreads();
  // prepare_tendencies_and_update_tracers: its loads, then its arithmetic
  // and stores
  double* const* tr = u.tracers;  // qv, ql, qr, qs, qi, qg
  double gu[S], gv[S], gt[S], pdn[S], pup[S], trq[6][S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    gu[s] = gv[s] = gt[s] = pdn[s] = pup[s] = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) trq[q][s] = 0.0;
    if (!L.v[s]) continue;
    const int32_t o = G.off[s];
    const int64_t col = c0 + L.cl[s];
    const int64_t i = col / sh3.nj, j = col - i * sh3.nj;
    const int64_t oi = i * u.si_i + j * u.sj_i + L.k[s];
    gu[s] = u.u_dt[o];
    gv[s] = u.v_dt[o];
    gt[s] = u.pt_dt[o];
    pup[s] = u.prsi[oi];
    pdn[s] = u.prsi[oi + 1];
#pragma unroll
    for (int q = 0; q < 6; ++q) trq[q][s] = tr[q][o];
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (!L.v[s]) continue;
    const int32_t o = G.off[s];
    const double qv1 = sh.buf[64 * s + lane];
    u.updated[0][o] = qv1;
    u.u_dt[o] = gu[s] + (x[8][s] - xin[8][s]) * u.rdt;
    u.v_dt[o] = gv[s] + (x[9][s] - xin[9][s]) * u.rdt;
    u.pt_dt[o] = gt[s] + (x[7][s] - xin[7][s]) * u.rdt;
    const double dp = pdn[s] - pup[s];
    const double qwat_qv = dp * qv1;
    const double qwat_ql = dp * x[1][s];
    const double qwat_qr = dp * x[2][s];
    const double qwat_qs = dp * x[4][s];
    const double qwat_qi = dp * x[3][s];
    const double qwat_qg = dp * x[5][s];
    const double qt = qwat_qv + qwat_ql + qwat_qr + qwat_qs + qwat_qi + qwat_qg;
    const double q_sum = trq[0][s] + trq[1][s] + trq[2][s] + trq[3][s] + trq[4][s] + trq[5][s];
    const double q0 = dp0[s] * (1.0 - q_sum) + qt;
    u.delp[o] = q0;
    tr[0][o] = qwat_qv / q0;
    tr[1][o] = qwat_ql / q0;
    tr[2][o] = qwat_qr / q0;
    tr[3][o] = qwat_qs / q0;
    tr[4][o] = qwat_qi / q0;
    tr[5][o] = qwat_qg / q0;
  }
}

// One sub-step (microphysics.py:2045-2220).
template <int S, class R, class FS>
__device__ __forceinline__ void substep(const Lay<S>& L, St<S, R>& X, Shared<S>& sh, const Glob<S>& G,
                                        const FS& f, const shield_mp_params& P, const Coef<R>& c, bool density) {
  warm_rain(L, X, sh, G, f, P, c, true, density);
  sedimentation(L, X, sh, G, f, P, c);
  warm_rain(L, X, sh, G, f, P, c, false, false);
  icloud(L, X, sh, G, f, P, c);
}

// ------------------------------------------------------------------ kernel
// Waves per SIMD the kernel is compiled for.  The single-sub-step
// mixed-precision kernel with at most 3 slots fits three waves (168 VGPRs)
// with a few bytes of spill, and so do the single-sub-step fused instances,
// whose state-update epilogue would otherwise push the whole kernel to two
// waves; every other instance is left to the register allocator.
template <int S, class R, bool MULTI, bool FUSED>
constexpr int min_waves() {
  return (S <= 3 && !MULTI) ? 3 : 1;
}

// MULTI: more than one sub-step.  With a single sub-step the body is not
// wrapped in a loop: loop-invariant code motion would otherwise keep every
// hoisted constant of the process code in a register across the whole body.
// FUSED: the state update of shield_mp_forward_update follows (a separate
// instance, so the plain kernel keeps its registers).
template <int S, class R, bool MULTI, bool FUSED>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(min_waves<S, R, MULTI, FUSED>(), 3))) void mp_forward_kernel(const FwdArgs<R> a) {
  const shield_mp_shape& sh3 = a.sh3;
  const shield_mp_params& P = a.P;
  const Coef<R>& c = a.c;
  const KFields<(uint32_t)offsetof(FwdArgs<R>, f)> f;
  const shield_update_fields& u = a.u;
  const int cols_per_wave = a.cols_per_wave;
  __shared__ Shared<S> sh;
  const int lane = lane_id();
  const int nz = (int)sh3.nk;
  const int64_t ncol = sh3.ni * sh3.nj;
  const int64_t c0 = (int64_t)blockIdx.x * cols_per_wave;
  const int ncols = (int)((ncol - c0) < cols_per_wave ? (ncol - c0) : cols_per_wave);
  const shield_mp_namelist& n = P.nml;

  Lay<S> L;
  Glob<S> G;
  L.nz = nz;
  L.ncols = ncols;
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
    G.off[s] = (int32_t)(i * sh3.si + j * sh3.sj + L.k[s]);
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
  fill_coef_tab(c, lane);
  __syncthreads();

  // ---------------------------------------------------- fields_init (:101-276)
  // de_ice (:110-120) only acts when the incoming tendencies are kept: with
  // overwritten (zero) tendencies qi never exceeds max(qi, qi0_max).  It
  // rewrites state.qice, state.qsnow and both tendencies in place, as the
  // reference does; fields_update then reads them back.
  const bool de_ice = n.de_ice && P.accumulate;
  St<S, R> X;
  // the call's input pt and delz, for the first sub-step's air density
  double t0_in[S], dz0_in[S];
  {
    // every input of every slot is loaded before any is used, so the loads
    // of the wavefront are in flight together
    const GFields F = f.all();
    double r_pt[S], r_dp[S], r_qv[S], r_ql[S], r_qr[S], r_qi[S], r_qs[S], r_qg[S], r_qa[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int32_t o = G.off[s];
      t0_in[s] = 250.0;
      dz0_in[s] = -100.0;
      if (L.v[s]) {
        t0_in[s] = F.pt[o];
        dz0_in[s] = F.delz[o];
        r_pt[s] = t0_in[s];
        r_dp[s] = F.delp[o];
        r_qv[s] = F.qv[o];
        r_ql[s] = F.ql[o];
        r_qr[s] = F.qr[o];
        r_qi[s] = F.qi[o];
        r_qs[s] = F.qs[o];
        r_qg[s] = F.qg[o];
        r_qa[s] = n.reset_qaz ? 0.0 : F.qaz[o];
      }
    }
    if (de_ice) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        if (!L.v[s]) continue;
        const int32_t o = G.off[s];
        const double qi_dt = F.qi_dt[o];
        const double qio = r_qi[s] - P.timestep * qi_dt;
        const double qin = mmax(qio, n.qi0_max);
        if (r_qi[s] > qin) {
          r_qs[s] = r_qs[s] + r_qi[s] - qin;
          r_qi[s] = qin;
          const double dqi = (qin - qio) * P.rdt;
          F.qs_dt[o] = F.qs_dt[o] + qi_dt - dqi;
          F.qi_dt[o] = dqi;
          F.qs[o] = r_qs[s];
          F.qi[o] = r_qi[s];
        }
      }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (L.v[s]) {
        X.tz[s] = r_pt[s];
        const double dp0 = r_dp[s];
        const double qv = r_qv[s];
        X.dp1[s] = dp0 * (1.0 - qv);
        const double omq = fdiv<R>(dp0, X.dp1[s]);
        // the products are materialised exactly (launder) because
        // fields_update forms the same products again and subtracts
        X.qv[s] = launder(qv * omq);
        X.ql[s] = launder(r_ql[s] * omq);
        X.qr[s] = launder(r_qr[s] * omq);
        X.qi[s] = launder(r_qi[s] * omq);
        X.qs[s] = launder(r_qs[s] * omq);
        X.qg[s] = launder(r_qg[s] * omq);
        X.qaz[s] = r_qa[s];
      } else {
        X.tz[s] = 250.0;
        X.dp1[s] = 100.0;
        X.qv[s] = X.ql[s] = X.qr[s] = X.qi[s] = X.qs[s] = X.qg[s] = X.qaz[s] = 0.0;
      }
      X.vtr[s] = X.vtg[s] = X.vts[s] = R(VF_MIN);
    }
  }
  // ccn scaled by the surface-layer density, then c_praut (:174-186)
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (L.v[s] && L.kb[s] == 0) {
      const int cl = L.cl[s];
      double ccn = sh.colv[CV_CCN][cl];
      if (P.use_ccn) {
        const int32_t o = G.off[s];
        const double p1 = dry_density<R>(X.dp1[s], f.delz[o]) * RDGAS * f.pt[o];
        ccn = fdiv<R>(ccn * RDGAS * X.tz[s], p1);
      }
      sh.colv[CV_CCN][cl] = ccn;
    }
  }
  __syncthreads();
  if (lane < ncols) {
    const double x = sh.colv[CV_CCN][lane] * RHOR;
    if constexpr (sizeof(R) == 4)
      sh.colv[CV_CPRAUT][lane] = P.cpaut * exp_sat(log_sat(x) * (-1.0 / 3.0));
    else
      sh.colv[CV_CPRAUT][lane] = P.cpaut * pow(x, -1.0 / 3.0);
  }
  __syncthreads();
  if (n.fix_negative) fix_negative(L, X, sh, P);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    sh.cell[CL_M1][64 * s + lane] = 0.0;
    sh.cell[CL_M1SOL][64 * s + lane] = 0.0;
  }
  air_density(L, X, sh, t0_in, dz0_in);

  // --------------------------------------------------------- sub-steps (:2045)
  if constexpr (MULTI) {
    for (int it = 0; it < P.ntimes; ++it) substep(L, X, sh, G, f, P, c, it > 0);
  } else {
    substep(L, X, sh, G, f, P, c, false);
  }

  // ---------------------------------------------------- fields_update (:1452-1553)
  // The initial values (q0 = q_in dp0 / dp1, qa0, u0, v0, t0, dp0) are read
  // again from the inputs and recomputed with the operations fields_init used.
  const double rdt = P.rdt;
  const bool acc = P.accumulate != 0;
  // every input of every slot is loaded before the first store, so the loads
  // of the wavefront are in flight together (a store may alias a later load
  // as far as the compiler knows, which would otherwise serialise them)
  const GFields F = f.all();
  double e_dp[S], e_u[S], e_v[S], e_pt[S], e_qa[S], e_q[6][S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int32_t o = G.off[s];
    if (L.v[s]) {
      e_dp[s] = F.delp[o];
      e_u[s] = F.ua[o];
      e_v[s] = F.va[o];
      e_pt[s] = F.pt[o];
      e_qa[s] = F.qa[o];
      e_q[0][s] = F.qv[o];
      e_q[1][s] = F.ql[o];
      e_q[2][s] = F.qr[o];
      e_q[3][s] = F.qi[o];
      e_q[4][s] = F.qs[o];
      e_q[5][s] = F.qg[o];
    } else {
      e_dp[s] = 100.0;
      e_u[s] = e_v[s] = e_pt[s] = e_qa[s] = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q) e_q[q][s] = 0.0;
    }
  }
  if (n.sedi_transport) {
    double m1[S], mp[S], a[S], bu[S], bv[S], u1[S], v1[S];
#pragma unroll
    for (int s = 0; s < S; ++s) m1[s] = sh.cell[CL_M1][64 * s + lane];
    col_prev(sh.buf, m1, mp);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const double u0 = e_u[s];
      const double v0 = e_v[s];
      if (L.k[s] == 0) {
        a[s] = 0.0;
        bu[s] = u0;
        bv[s] = v0;
      } else {
        const double dp0 = e_dp[s];
        const double r = recip<R>(dp0 + mp[s]);
        a[s] = mp[s] * r;
        if constexpr (sizeof(R) == 8) {
          // FP64: the reference's division, exact where no mass falls in
          bu[s] = dp0 * u0 / (dp0 + mp[s]);
          bv[s] = dp0 * v0 / (dp0 + mp[s]);
        } else {
          bu[s] = dp0 * u0 * r;
          bv[s] = dp0 * v0 * r;
        }
      }
    }
    scan_fwd2(L, a, bu, bv, u1, v1);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (!L.v[s]) continue;
      const int32_t o = G.off[s];
      if (L.k[s] > 0) {
        F.udt[o] = (acc ? F.udt[o] : 0.0) + (u1[s] - e_u[s]) * rdt;
        F.vdt[o] = (acc ? F.vdt[o] : 0.0) + (v1[s] - e_v[s]) * rdt;
      } else if (!acc) {
        F.udt[o] = 0.0;
        F.vdt[o] = 0.0;
      }
    }
  } else if (!acc) {
#pragma unroll
    for (int s = 0; s < S; ++s)
      if (L.v[s]) {
        F.udt[G.off[s]] = 0.0;
        F.vdt[G.off[s]] = 0.0;
      }
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (!L.v[s]) continue;
    const int32_t o = G.off[s];
    const double dp0 = e_dp[s];
    const double omq0 = fdiv<R>(dp0, X.dp1[s]);
    const double omq = fdiv<R>(X.dp1[s], dp0);
    // q0 is rounded exactly as fields_init rounded it (launder keeps the
    // product out of a fused multiply-add with the difference), so an
    // untouched species has a zero tendency
    F.qv_dt[o] = (acc ? F.qv_dt[o] : 0.0) + rdt * (X.qv[s] - launder(e_q[0][s] * omq0)) * omq;
    F.ql_dt[o] = (acc ? F.ql_dt[o] : 0.0) + rdt * (X.ql[s] - launder(e_q[1][s] * omq0)) * omq;
    F.qr_dt[o] = (acc ? F.qr_dt[o] : 0.0) + rdt * (X.qr[s] - launder(e_q[2][s] * omq0)) * omq;
    F.qi_dt[o] = (acc ? F.qi_dt[o] : 0.0) + rdt * (X.qi[s] - launder(e_q[3][s] * omq0)) * omq;
    F.qs_dt[o] = (acc ? F.qs_dt[o] : 0.0) + rdt * (X.qs[s] - launder(e_q[4][s] * omq0)) * omq;
    F.qg_dt[o] = (acc ? F.qg_dt[o] : 0.0) + rdt * (X.qg[s] - launder(e_q[5][s] * omq0)) * omq;
    const double cvm = P.c_air + X.qv[s] * P.c_vap + (X.qr[s] + X.ql[s]) * C_LIQ +
                       (X.qi[s] + X.qs[s] + X.qg[s]) * C_ICE;
    F.pt_dt[o] = (acc ? F.pt_dt[o] : 0.0) + fdiv<R>(rdt * (X.tz[s] - e_pt[s]) * cvm, CP_AIR);
    if (n.do_qa)
      F.qa_dt[o] = 0.0;
    else
      F.qa_dt[o] = (acc ? F.qa_dt[o] : 0.0) + rdt * (fdiv<R>(X.qaz[s], P.ntimes) - e_qa[s]);
    F.qaz[o] = X.qaz[s];
  }
  if (lane < ncols) {
    const int64_t col = c0 + lane;
    const int64_t i = col / sh3.nj, j = col - i * sh3.nj;
    const int64_t o2 = i * sh3.si2 + j * sh3.sj2;
    const double convt = 86400.0 * rdt * RGRAV;
    if (F.rain) F.rain[o2] = sh.colv[CV_RAIN][lane] * convt;
    if (F.snow) F.snow[o2] = sh.colv[CV_SNOW][lane] * convt;
    if (F.ice) F.ice[o2] = sh.colv[CV_ICE][lane] * convt;
    if (F.graupel) F.graupel[o2] = sh.colv[CV_GRAUPEL][lane] * convt;
  }
  if constexpr (FUSED) state_update(L, G, sh, sh3, c0, a.f, u);
}

// Launch the kernel for one slot count: a grid of ceil(ncol / P) one-wave
// workgroups.  u.delp != NULL selects the fused state update.  The single-
// sub-step and the looping (MULTI) instances are instantiated in files of
// their own (mp_forward_s<S>.hip, mp_forward_multi_s<S>.hip), so each is
// compiled with the flags that suit it.
template <int S, class R, bool MULTI>
hipError_t launch_kernels(const FwdArgs<R>& a, bool fused, unsigned nblocks, hipStream_t stream) {
  const dim3 grid(nblocks), block(64);
  if (fused)
    hipLaunchKernelGGL((mp_forward_kernel<S, R, MULTI, true>), grid, block, 0, stream, a);
  else
    hipLaunchKernelGGL((mp_forward_kernel<S, R, MULTI, false>), grid, block, 0, stream, a);
  return hipGetLastError();
}
template <int S, class R>
hipError_t launch_slots(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f,
                        const shield_update_fields& u, int P, hipStream_t stream) {
  const int64_t ncol = sh.ni * sh.nj;
  const unsigned nblocks = (unsigned)((ncol + P - 1) / P);
  const FwdArgs<R> a{sh, p, make_coef<R>(p), f, u, P};
  const bool fused = u.delp != nullptr;
  return p.ntimes > 1 ? launch_kernels<S, R, true>(a, fused, nblocks, stream)
                      : launch_kernels<S, R, false>(a, fused, nblocks, stream);
}

// Resident wavefronts per SIMD of one kernel instance on the current device
// (0 when the query fails, e.g. without a device).
template <int S, class R, bool MULTI>
int waves_kernels(bool fused) {
  int blocks = 0;
  const hipError_t e =
      fused ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, mp_forward_kernel<S, R, MULTI, true>, 64, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, mp_forward_kernel<S, R, MULTI, false>, 64, 0);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return blocks / 4;  // one-wave workgroups over the 4 SIMDs of a CU
}
template <int S, class R>
int waves_per_simd(bool multi, bool fused) {
  return multi ? waves_kernels<S, R, true>(fused) : waves_kernels<S, R, false>(fused);
}

}  // namespace shield
