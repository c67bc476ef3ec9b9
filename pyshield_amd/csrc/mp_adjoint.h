// mp_adjoint.h -- the reverse-mode adjoint of one FP64 microphysics call.
//
// Given the call's inputs and cotangents of its outputs (the ten tendencies
// and the four surface precipitation rates), the adjoint returns cotangents
// of the twelve input fields (pt, qv, ql, qr, qi, qs, qg, qa, ua, va, delp,
// delz), of the incoming cloud-fraction accumulator qaz, and the gradient of
// six process parameters: fac_rc (autoconversion threshold qc0 = fac_rc ccn),
// vr_fac, vs_fac, vg_fac, vi_fac (fall-speed coefficients) and qi_gen (ice
// generation / nucleation).
//
// Layout: P columns per wavefront in S slots (the levels of a column on
// consecutive lanes, as in the forward kernel; 2 columns in 3 slots at
// nz = 79), and a persistent grid whose wavefronts walk the columns.  Each
// wavefront
//  1. replays the forward of its columns (the FP64 forward's building blocks,
//     in the forward's order) and records, at 12 points of every sub-step,
//     the per-cell values the reverse reads there (77 per cell and sub-step)
//     in its own region of the workspace;
//  2. walks the steps backwards.  Every point-wise step (density update,
//     rain and solid fall speeds, evaporation / accretion, autoconversion,
//     sedimentation ice melt, and the icloud step: ice melt / freeze, the
//     ice-phase collections and the saturation adjustment of subgrid_z_proc)
//     has a hand-written reverse (mp_point_adj.h) that re-evaluates the
//     step's FP64 forward and follows the branch it took at the point; the
//     liquid slope limiter's three-point stencil uses dual numbers
//     (mp_dual.h).  The vertical steps have hand-written transposes: the
//     implicit fall and the sedimentation heating (backward affine scans),
//     the liquid slope limiter, the momentum transport, and the
//     precipitation and flux sums.
// Supported switches: all of the forward's but de_ice on accumulated
// tendencies (the adjoint differentiates overwritten ones, on which de_ice
// changes nothing), which the host rejects.  fix_negative, z_slope_ice and
// do_sedi_w run in an instance of their own (EXT), which keeps their code out
// of the default one.
#pragma once

#include "mp_forward.h"
#include "mp_point_adj.h"

namespace shield {

enum { AP_FAC_RC, AP_VR_FAC, AP_VS_FAC, AP_VG_FAC, AP_VI_FAC, AP_QI_GEN, AP_N };

// per-cell record of the state: the prognostic cell and the auxiliary fields
enum { Z_QV, Z_QL, Z_QR, Z_QI, Z_QS, Z_QG, Z_TZ, Z_QA, Z_DEN, Z_DENFAC, Z_DZ1, Z_VTR, Z_VTG, Z_VTS, Z_N };
// record points of one sub-step: warm_rain (first) before the evaporation,
// before the fall, before the second evaporation, before autoconversion;
// sedimentation at entry, after the ice melt, after the falls; the same four
// for warm_rain (second); icloud at entry; and the density, its surface ratio
// and the layer thickness, which are fixed after the density update that
// opens the sub-step
enum { RW1_B, RW1_C, RW1_D, RW1_E, RS_F, RS_G, RS_H, RW2_B, RW2_C, RW2_D, RW2_E, RI, RAUX, REC_N };

// The fields a record keeps: those the reverse step reads from it.  The cell
// values (qv .. tz) at most points; the rain fall speed beside them before
// the rain fall; the solid species and their fall speeds after the ice melt;
// temperature and cloud water before autoconversion.  The reverse reads den,
// denfac and dz1 from RAUX, icloud's fall speeds from RW2_C and RS_G, and
// never the cloud-fraction accumulator.
constexpr uint32_t zbit(int z) { return 1u << z; }
constexpr uint32_t ZM_CELL = zbit(Z_QV) | zbit(Z_QL) | zbit(Z_QR) | zbit(Z_QI) | zbit(Z_QS) | zbit(Z_QG) | zbit(Z_TZ);
constexpr uint32_t ZM_AUX = zbit(Z_DEN) | zbit(Z_DENFAC) | zbit(Z_DZ1);
__host__ __device__ constexpr uint32_t rec_fields(int rec) {
  return (rec == RW1_C || rec == RW2_C)   ? ZM_CELL | zbit(Z_VTR)
         : (rec == RW1_E || rec == RW2_E) ? zbit(Z_QL) | zbit(Z_TZ)
         : (rec == RS_G) ? zbit(Z_QI) | zbit(Z_QS) | zbit(Z_QG) | zbit(Z_VTG) | zbit(Z_VTS)
         : (rec == RAUX) ? ZM_AUX
                         : ZM_CELL;
}
__host__ __device__ constexpr int nbits(uint32_t m) { return __builtin_popcount(m); }
// offset of record rec in a sub-step's block, in fields (a loop, not
// recursion, so it folds to a constant wherever rec is one)
__host__ __device__ constexpr int rec_offset(int rec) {
  int o = 0;
  for (int r = 0; r < rec; ++r) o += nbits(rec_fields(r));
  return o;
}
constexpr int REC_FIELDS = rec_offset(REC_N);
static_assert(REC_FIELDS == 77, "fields per sub-step");
// With do_sedi_w, per sub-step after all the records: w before each
// transporting stage (the first rain fall, the solid falls, the second rain
// fall) and qv, ql, qr before the solid falls (the ice melt changes ql, qr).
enum { SW_W1, SW_WS, SW_W2, SW_QV, SW_QL, SW_QR, SW_FIELDS };

// The adjoint kernel's argument block.  As in the forward kernel, the field
// pointers of f and a are read from the kernarg segment where they are used
// (KPtr), not held in scalar registers for the whole kernel.
struct AdjArgs {
  shield_mp_shape sh3;
  shield_mp_params P;
  Coef<double> c;
  shield_mp_fields f;
  shield_mp_adjoint_fields a;
  double* tape;
  int64_t tape_stride;
  double* part;
  int cols_per_wave;
};
#define SH_AFIELDS(X)                                                                                            \
  X(pt_dt) X(qv_dt) X(ql_dt) X(qr_dt) X(qi_dt) X(qs_dt) X(qg_dt) X(qa_dt) X(udt) X(vdt) X(rain) X(snow) X(ice) \
  X(graupel) X(pt) X(qv) X(ql) X(qr) X(qi) X(qs) X(qg) X(qa) X(ua) X(va) X(delp) X(delz) X(qaz) X(w_after) X(w)
// shield_mp_adjoint_fields at kernarg offset B, member by member
template <uint32_t B>
struct KAdjFields {
#define SH_KAF(n) KPtr<B + (uint32_t)offsetof(shield_mp_adjoint_fields, n)> n;
  SH_AFIELDS(SH_KAF)
#undef SH_KAF
};

// Member MEMBER of the argument block, read from the kernarg segment through
// an offset the compiler cannot see (as KPtr does for the field pointers).
// The coefficient block and the parameters are passed to each stage of the
// persistent loop this way: read directly, their loads are loop-invariant and
// the compiler hoists them all to the kernel's entry, where the scalar
// registers cannot hold them and they spill to VGPR lanes, every use then a
// v_readlane.  Through karg each stage loads what it uses with scalar loads.
template <class T, uint32_t OFF>
__device__ __forceinline__ const T& karg() {
  uint32_t o = OFF;
  asm volatile("" : "+s"(o));
  const __attribute__((address_space(4))) char* base =
      (const __attribute__((address_space(4))) char*)__builtin_amdgcn_kernarg_segment_ptr();
  return *(const T*)(const __attribute__((address_space(4))) T*)(base + o);
}

// per-cell cotangents of the auxiliary fields, kept in LDS
enum { AB_DEN, AB_DENFAC, AB_DZ1, AB_VTR, AB_VTG, AB_VTS, AB_M1, AB_M1SOL, AB_DP1, AB_T0, AB_DZ0, AB_N };

template <int S>
struct AdjShared {
  Shared<S> sh;
  double bar[AB_N][64 * S];
  // per column: rain, snow, ice, graupel (per unit of the accumulated flux), ccn, c_praut
  double colbar[8][pmax_for(S)];
};
enum { CB_RAIN, CB_SNOW, CB_ICE, CB_GRAUPEL, CB_CCN, CB_CPRAUT };

template <int S>
struct TapeView {
  double* base;  // this wavefront's region: ntimes * REC_FIELDS * 64 S doubles
  double* sw;    // then, with do_sedi_w, ntimes * SW_FIELDS * 64 S doubles
  __device__ __forceinline__ double* at_sw(int it, int z) const { return sw + ((size_t)it * SW_FIELDS + z) * (64 * S); }
  // field z of record rec (z must be one of rec_fields(rec))
  __device__ __forceinline__ double* at(int it, int rec, int z) const {
    const int k = rec_offset(rec) + nbits(rec_fields(rec) & (zbit(z) - 1u));
    return base + ((size_t)it * REC_FIELDS + k) * (64 * S);
  }
};

// value of field z of the replay's state
template <int S>
__device__ __forceinline__ double rec_value(const St<S, double>& X, Shared<S>& sh, int z, int s) {
  switch (z) {
    case Z_QV: return X.qv[s];
    case Z_QL: return X.ql[s];
    case Z_QR: return X.qr[s];
    case Z_QI: return X.qi[s];
    case Z_QS: return X.qs[s];
    case Z_QG: return X.qg[s];
    case Z_TZ: return X.tz[s];
    case Z_QA: return X.qaz[s];
    case Z_DEN: return sh.at(CL_DEN, s);
    case Z_DENFAC: return sh.at(CL_DENFAC, s);
    case Z_DZ1: return sh.at(CL_DZ1, s);
    case Z_VTR: return X.vtr[s];
    case Z_VTG: return X.vtg[s];
    default: return X.vts[s];
  }
}

// record the fields of `mask` (by default all that record rec keeps)
template <int S>
__device__ __forceinline__ void tape_save(const TapeView<S>& T, int it, int rec, const St<S, double>& X,
                                          Shared<S>& sh, uint32_t mask = 0xFFFFFFFFu) {
  const int lane = lane_id();
  mask &= rec_fields(rec);
#pragma unroll
  for (int z = 0; z < Z_N; ++z) {
    if (!(mask & zbit(z))) continue;
#pragma unroll
    for (int s = 0; s < S; ++s) T.at(it, rec, z)[64 * s + lane] = rec_value(X, sh, z, s);
  }
}

// a recorded state: z[field][slot]
template <int S>
struct Rec {
  double z[Z_N][S];
};
// Load record rec, with den, denfac and dz1 from LDS (where the reverse walk
// put the sub-step's RAUX record); fields the record does not keep read as
// NaN.
template <int S>
__device__ __forceinline__ void tape_load(const TapeView<S>& T, int it, int rec, Rec<S>& r, Shared<S>& sh) {
  const int lane = lane_id();
  const uint32_t m = rec_fields(rec);
#pragma unroll
  for (int q = 0; q < Z_N; ++q)
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (q == Z_DEN || q == Z_DENFAC || q == Z_DZ1)
        r.z[q][s] = sh.cell[q == Z_DEN ? CL_DEN : q == Z_DENFAC ? CL_DENFAC : CL_DZ1][64 * s + lane];
      else
        r.z[q][s] = (m & zbit(q)) ? T.at(it, rec, q)[64 * s + lane] : __builtin_nan("");
    }
}

// put the sub-step's den, denfac and dz1 into LDS for tape_load
template <int S>
__device__ __forceinline__ void aux_load(const TapeView<S>& T, int it, Shared<S>& sh) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    sh.cell[CL_DEN][64 * s + lane] = T.at(it, RAUX, Z_DEN)[64 * s + lane];
    sh.cell[CL_DENFAC][64 * s + lane] = T.at(it, RAUX, Z_DENFAC)[64 * s + lane];
    sh.cell[CL_DZ1][64 * s + lane] = T.at(it, RAUX, Z_DZ1)[64 * s + lane];
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

// x[k] = a[k] x[k+1] + b[k] for k < nz-1, x[nz-1] = b[nz-1]  (surface -> top),
// the transpose direction of scan_fwd
template <int S>
__device__ __forceinline__ void scan_bwd(const Lay<S>& L, const double (&a)[S], const double (&b)[S],
                                         double (&x)[S]) {
  const int lane = lane_id();
  double carry = 0.0;
#pragma unroll
  for (int s = S - 1; s >= 0; --s) {
    double A = a[s], B = b[s];
    const int kb = L.kb[s];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const double Ap = __shfl_down(A, d), Bp = __shfl_down(B, d);
      if (lane + d < 64 && kb >= d) {
        B = A * Bp + B;
        A = A * Ap;
      }
    }
    x[s] = (kb <= 63 - lane) ? B : A * carry + B;
    carry = readlane(x[s], 0);
  }
}
// Two of them with the same multipliers a (each output gets exactly the
// operations of scan_bwd; the multipliers' composition is shared)
template <int S>
__device__ __forceinline__ void scan_bwd2(const Lay<S>& L, const double (&a)[S], const double (&b)[S],
                                          const double (&c)[S], double (&x)[S], double (&y)[S]) {
  const int lane = lane_id();
  double carry_x = 0.0, carry_y = 0.0;
#pragma unroll
  for (int s = S - 1; s >= 0; --s) {
    double A = a[s], B = b[s], C = c[s];
    const int kb = L.kb[s];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const double Ap = __shfl_down(A, d), Bp = __shfl_down(B, d), Cp = __shfl_down(C, d);
      if (lane + d < 64 && kb >= d) {
        B = A * Bp + B;
        C = A * Cp + C;
        A = A * Ap;
      }
    }
    x[s] = (kb <= 63 - lane) ? B : A * carry_x + B;
    y[s] = (kb <= 63 - lane) ? C : A * carry_y + C;
    carry_x = readlane(x[s], 0);
    carry_y = readlane(y[s], 0);
  }
}

// cotangents of the eight cell values, in registers
template <int S>
struct CellBar {
  double v[8][S];
};

// the liquid slope limiter at one interior level (the forward's
// slope_limiter, microphysics.py:595-620), on any scalar type
template <class W>
__device__ __forceinline__ W slope_point(W qm, W q0, W qp) {
  const W dq = 0.5 * (q0 - qm);
  const W dqn = 0.5 * (qp - q0);
  W d = 0.5 * mmin(mabs(dq + dqn), 0.5 * q0);
  if (dq * dqn <= 0.0) {
    if (dq > 0.0)
      d = mmin(d, mmin(dq, -dqn));
    else
      d = W(0.0);
  }
  return d;
}

// ---------------------------------------------------- vertical transposes
// Transpose of implicit_fall for one species.  q, vt: the fall's inputs;
// qbar: in, the cotangent of q after the fall; out, before it.  fluxbar:
// cotangent of the flux out of each layer bottom (the surface flux included
// in the surface layer's).  Accumulates vt, dz1 and dp1 cotangents.
template <int S>
__device__ __forceinline__ void fall_adj(const Lay<S>& L, Shared<S>& sh, const double (&q)[S], const double (&vt)[S],
                                         const double (&dz1)[S], const double (&dp1)[S], double dt,
                                         const bool (&fall)[S], double (&qbar)[S], const double (&fluxbar)[S],
                                         double (&vtbar)[S], double (&dz1bar)[S], double (&dp1bar)[S]) {
  double dz[S], dd[S], qq[S], ddp[S], r[S], a[S], b[S], qm[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    dz[s] = -dz1[s];
    dd[s] = dt * vt[s];
    qq[s] = q[s] * dp1[s];
  }
  col_prev(sh.buf, dd, ddp);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    r[s] = 1.0 / (dz[s] + dd[s]);
    a[s] = ddp[s] * r[s];
    b[s] = qq[s] * r[s];
  }
  scan_fwd(L, a, b, qm);  // mass per unit thickness
  double g[S], ddbar[S], dzbar[S], an[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (fall[s] && L.v[s]) {
      const double qn = qm[s] * dz[s] / dp1[s];
      g[s] = qbar[s] * dz[s] / dp1[s] + fluxbar[s] * dd[s];
      dzbar[s] = qbar[s] * qm[s] * rcp64(dp1[s]);
      dp1bar[s] -= qbar[s] * qn * rcp64(dp1[s]);
      ddbar[s] = fluxbar[s] * qm[s];
    } else {
      g[s] = dzbar[s] = ddbar[s] = 0.0;
    }
  }
  col_next(sh.buf, a, an);
#pragma unroll
  for (int s = 0; s < S; ++s)
    if (L.kb[s] == 0) an[s] = 0.0;
  double lam[S], qmp[S], t[S], tn[S];
  scan_bwd(L, an, g, lam);
  col_prev(sh.buf, qm, qmp);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    t[s] = (fall[s] && L.v[s] && L.k[s] > 0) ? lam[s] * r[s] * qmp[s] : 0.0;
  }
  col_next(sh.buf, t, tn);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (!(fall[s] && L.v[s])) continue;
    const double qqbar = lam[s] * r[s];
    const double sbar = -lam[s] * qm[s] * r[s];  // d qm / d (dz + dd)
    ddbar[s] += sbar + (L.kb[s] > 0 ? tn[s] : 0.0);
    dzbar[s] += sbar;
    qbar[s] = qqbar * dp1[s];
    dp1bar[s] += qqbar * q[s];
    vtbar[s] += ddbar[s] * dt;
    dz1bar[s] -= dzbar[s];
  }
}

// Transpose of sedi_heat.  st: the state before the heating (tz and the
// mixing ratios of cvn); flux: the mass flux it uses; tbar: in, cotangent of
// tz after; out, before.  Accumulates the flux, mixing-ratio, dp1 and dz1
// cotangents.
template <int S>
__device__ __forceinline__ void heat_adj(const Lay<S>& L, Shared<S>& sh, const Rec<S>& st, const double (&dp1)[S],
                                         const double (&flux)[S], double cq, const bool (&apply)[S],
                                         CellBar<S>& xb, double (&fluxbar)[S], double (&dp1bar)[S],
                                         double (&dz1bar)[S]) {
  double fp[S], cvn[S], dgz[S], a[S], b[S], tn[S], D[S];
  col_prev(sh.buf, flux, fp);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    dgz[s] = -0.5 * GRAV * st.z[Z_DZ1][s];
    cvn[s] = dp1[s] * (CV_AIR + st.z[Z_QV][s] * CV_VAP + (st.z[Z_QR][s] + st.z[Z_QL][s]) * C_LIQ +
                       (st.z[Z_QI][s] + st.z[Z_QS][s] + st.z[Z_QG][s]) * C_ICE);
    D[s] = cvn[s] + cq * flux[s];
    const double tz = st.z[Z_TZ][s];
    if (L.k[s] == 0) {
      a[s] = 0.0;
      b[s] = tz + flux[s] * dgz[s] / (cvn[s] + flux[s] * cq);
    } else {
      const double r = 1.0 / (cvn[s] + cq * flux[s]);
      a[s] = fp[s] * cq * r;
      b[s] = ((cvn[s] + cq * (flux[s] - fp[s])) * tz + dgz[s] * (fp[s] + flux[s])) * r;
    }
  }
  scan_fwd(L, a, b, tn);
  double g[S], an[S], mu[S], tnp[S], fpb[S], fpbn[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const bool ap = apply[s] && L.v[s];
    g[s] = ap ? xb.v[6][s] : 0.0;
    if (ap) xb.v[6][s] = 0.0;
  }
  col_next(sh.buf, a, an);
#pragma unroll
  for (int s = 0; s < S; ++s)
    if (L.kb[s] == 0) an[s] = 0.0;
  scan_bwd(L, an, g, mu);
  col_prev(sh.buf, tn, tnp);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    fpb[s] = 0.0;
    if (!(apply[s] && L.v[s])) continue;
    const double m = mu[s], tz = st.z[Z_TZ][s];
    double cvnbar, dgzbar;
    if (L.k[s] == 0) {
      xb.v[6][s] += m;
      fluxbar[s] += m * dgz[s] * cvn[s] * rcp64(D[s] * D[s]);
      dgzbar = m * flux[s] * rcp64(D[s]);
      cvnbar = -m * flux[s] * dgz[s] * rcp64(D[s] * D[s]);
    } else {
      xb.v[6][s] += m * (cvn[s] + cq * (flux[s] - fp[s])) * rcp64(D[s]);
      fluxbar[s] += m * (cq * tz + dgz[s] - cq * tn[s]) * rcp64(D[s]);
      fpb[s] = m * (-cq * tz + dgz[s] + cq * tnp[s]) * rcp64(D[s]);
      cvnbar = m * (tz - tn[s]) * rcp64(D[s]);
      dgzbar = m * (fp[s] + flux[s]) * rcp64(D[s]);
    }
    dp1bar[s] += cvnbar * cvn[s] * rcp64(dp1[s]);
    xb.v[0][s] += cvnbar * dp1[s] * CV_VAP;
    xb.v[1][s] += cvnbar * dp1[s] * C_LIQ;
    xb.v[2][s] += cvnbar * dp1[s] * C_LIQ;
    xb.v[3][s] += cvnbar * dp1[s] * C_ICE;
    xb.v[4][s] += cvnbar * dp1[s] * C_ICE;
    xb.v[5][s] += cvnbar * dp1[s] * C_ICE;
    dz1bar[s] += dgzbar * (-0.5 * GRAV);
  }
  col_next(sh.buf, fpb, fpbn);  // flux[k-1] cotangents land on level k-1
#pragma unroll
  for (int s = 0; s < S; ++s)
    if (L.v[s] && L.kb[s] > 0) fluxbar[s] += fpbn[s];
}

// forward implicit fall of a recorded species: returns the flux and the
// species after the fall
template <int S>
__device__ __forceinline__ void fall_fwd(const Lay<S>& L, Shared<S>& sh, const double (&q)[S], const double (&vt)[S],
                                         const double (&dz1)[S], const double (&dp1)[S], double dt,
                                         const bool (&fall)[S], double (&qn)[S], double (&flux)[S]) {
  double dz[S], dd[S], qq[S], ddp[S], a[S], b[S], qm[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    dz[s] = -dz1[s];
    dd[s] = dt * vt[s];
    qq[s] = q[s] * dp1[s];
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
    if (fall[s]) {
      flux[s] = dd[s] * qm[s];
      qn[s] = qm[s] * dz[s] / dp1[s];
    } else {
      flux[s] = 0.0;
      qn[s] = q[s];
    }
  }
}


// ------------------------------------------------------- w transport (do_sedi_w)
// The forward's transport of the vertical velocity with a species' falling
// mass (implicit_fall with do_sedi_w, microphysics.py:499-515): in falling
// columns w[k] becomes (dm w - F[k-1] v[k-1] + F[k] v[k]) / (dm + F[k-1] - F[k])
// (no k-1 terms at the top), dm the layer's total mass before the fall, F
// the fall's mass flux through each layer bottom, v its fall speed.
template <int S>
__device__ __forceinline__ void w_transport(const Lay<S>& L, Shared<S>& sh, const double (&dm)[S],
                                            const double (&flux)[S], const double (&vt)[S], const bool (&fall)[S],
                                            double (&w)[S]) {
  double fl[S], fp[S], vp[S];
#pragma unroll
  for (int s = 0; s < S; ++s) fl[s] = fall[s] ? flux[s] : 0.0;
  col_prev(sh.buf, fl, fp);
  col_prev(sh.buf, vt, vp);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (!fall[s] || !L.v[s]) continue;
    if (L.k[s] == 0)
      w[s] = (dm[s] * w[s] + fl[s] * vt[s]) / (dm[s] - fl[s]);
    else
      w[s] = (dm[s] * w[s] - fp[s] * vp[s] + fl[s] * vt[s]) / (dm[s] + fp[s] - fl[s]);
  }
}
// Its transpose: w0, the velocity before; wb in: the cotangent after, out:
// before.  Accumulates the cotangents of dm, the flux and the fall speed (the
// k-1 terms land on level k-1).
template <int S>
__device__ __forceinline__ void w_transport_adj(const Lay<S>& L, Shared<S>& sh, const double (&dm)[S],
                                                const double (&flux)[S], const double (&vt)[S], const bool (&fall)[S],
                                                const double (&w0)[S], double (&wb)[S], double (&dmb)[S],
                                                double (&fluxb)[S], double (&vtb)[S]) {
  double fl[S], fp[S], vp[S], fpb[S], vpb[S], fpbn[S], vpbn[S];
#pragma unroll
  for (int s = 0; s < S; ++s) fl[s] = fall[s] ? flux[s] : 0.0;
  col_prev(sh.buf, fl, fp);
  col_prev(sh.buf, vt, vp);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    fpb[s] = vpb[s] = 0.0;
    if (!fall[s] || !L.v[s]) continue;
    const bool top = L.k[s] == 0;
    const double N = top ? dm[s] * w0[s] + fl[s] * vt[s] : dm[s] * w0[s] - fp[s] * vp[s] + fl[s] * vt[s];
    const double D = top ? dm[s] - fl[s] : dm[s] + fp[s] - fl[s];
    const double wn = N / D;
    const double Nb = wb[s] * rcp64(D), Db = -wb[s] * wn * rcp64(D);
    wb[s] = Nb * dm[s];
    dmb[s] += Nb * w0[s] + Db;
    fluxb[s] += Nb * vt[s] - Db;
    vtb[s] += Nb * fl[s];
    if (!top) {
      fpb[s] = Db - Nb * vp[s];
      vpb[s] = -Nb * fp[s];
    }
  }
  col_next(sh.buf, fpb, fpbn);
  col_next(sh.buf, vpb, vpbn);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (!L.v[s] || L.kb[s] == 0) continue;
    fluxb[s] += fpbn[s];
    vtb[s] += vpbn[s];
  }
}
// dm = dp1 (1 + qv + ql + qr + qi + qs + qg) of a cell
__device__ __forceinline__ double total_mass(double dp1, double qv, double ql, double qr, double qi, double qs,
                                             double qg) {
  return dp1 * (1.0 + qv + ql + qr + qi + qs + qg);
}

// ------------------------------------------------------- taped forward
template <int S, class FS>
__device__ __forceinline__ void warm_rain_rec(const Lay<S>& L, St<S, double>& X, Shared<S>& sh, const Glob<S>& G,
                                              const FS& f, const shield_mp_params& P,
                                              const Coef<double>& c, bool is_first, const TapeView<S>& T, int it,
                                              int rb, double (&wv)[S], bool sw) {
  const int lane = lane_id();
  const double t_wfr = P.t_wfr;
  if (is_first) {
    // the forward kernel's density update (dz1, den, denfac, and icloud's p1)
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
    tape_save(T, it, RAUX, X, sh);
  }
  tape_save(T, it, rb, X, sh);
  bool pred[S], fall[S];
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.qr[s] > QRMIN);
  col_any(L, pred, fall);
#pragma unroll
  for (int s = 0; s < S; ++s) X.vtr[s] = rain_fspeed<double>(c, !fall[s], X.qr[s], sh.at(CL_DEN, s));
  double flux[S], surf[S];
#pragma unroll
  for (int s = 0; s < S; ++s) flux[s] = surf[s] = 0.0;
  const double dt5 = 0.5 * c.dt_rain;
  if (wave_any(fall)) {
    evaporation(L, X, sh, c, t_wfr, dt5, fall);
    tape_save(T, it, rb + 1, X, sh);
    double dz[S], dm[S];
    layer_dz(L, sh, dz);
    if (sw) {
      // w and the total mass before the fall (the adjoint leaves f.w alone)
#pragma unroll
      for (int s = 0; s < S; ++s) {
        T.at_sw(it, is_first ? SW_W1 : SW_W2)[64 * s + lane] = wv[s];
        dm[s] = total_mass(X.dp1[s], X.qv[s], X.ql[s], X.qr[s], X.qi[s], X.qs[s], X.qg[s]);
      }
    }
    implicit_fall(L, X, sh, G, f, dz, X.qr, X.vtr, P.dt_rain, fall, false, flux, surf);
    if (sw) w_transport(L, sh, dm, flux, X.vtr, fall, wv);
    if (P.nml.do_sedi_heat) sedi_heat(L, X, sh, flux, C_LIQ, fall);
    tape_save(T, it, rb + 2, X, sh);
    evaporation(L, X, sh, c, t_wfr, dt5, fall);
  } else {
    // no rain falls: the reverse reads only the fall speed (icloud's)
    tape_save(T, it, rb + 1, X, sh, zbit(Z_VTR));
  }
  tape_save(T, it, rb + 3, X, sh);
  if (P.nml.irain_f != 0) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      Cell cs = cell_of(X, s);
      autoconv_no_subgrid_var<double>(c, t_wfr, cs, sh.at(CL_DEN, s), sh.colv[CV_CCN][L.cl[s]],
                                      sh.colv[CV_CPRAUT][L.cl[s]]);
      set_cell(X, s, cs);
    }
  } else {
    double dl[S];
    if (P.nml.z_slope_liq) slope_limiter(L, sh, X.ql, dl);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      Cell cs = cell_of(X, s);
      const double hv = sh.colv[CV_HVAR][L.cl[s]];
      const double d = P.nml.z_slope_liq ? mmax(dl[s], mmax(QVMIN, hv * cs.ql)) : mmax(QVMIN, hv * cs.ql);
      autoconv_subgrid_var<double>(c, t_wfr, cs, sh.at(CL_DEN, s), sh.colv[CV_CCN][L.cl[s]],
                                   sh.colv[CV_CPRAUT][L.cl[s]], d);
      set_cell(X, s, cs);
    }
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (L.v[s] && L.kb[s] == 0) sh.colv[CV_RAIN][L.cl[s]] += surf[s];
    double& m1 = sh.cell[CL_M1][64 * s + lane];
    m1 = is_first ? m1 + flux[s] : m1 + flux[s] + sh.cell[CL_M1SOL][64 * s + lane];
  }
  __syncthreads();
}

template <int S, class FS>
__device__ __forceinline__ void sed_rec(const Lay<S>& L, St<S, double>& X, Shared<S>& sh, const Glob<S>& G,
                                        const FS& f, const shield_mp_params& P, const Coef<double>& c,
                                        const TapeView<S>& T, int it, double (&wv)[S], bool sw) {
  const int lane = lane_id();
  const double tice_d = P.nml.tice;
  tape_save(T, it, RS_F, X, sh);
  double vti[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    double vg, vi, vs;
    fall_speed<double>(c, X.qg[s], X.qi[s], X.qs[s], X.tz[s] - tice_d, sh.at(CL_DEN, s), vg, vi, vs);
    X.vtg[s] = vg;
    vti[s] = vi;
    X.vts[s] = vs;
  }
  bool pred[S], stop[S], fall[S];
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.tz[s] > tice_d);
  col_from_first(L, pred, stop);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (stop[s]) {
      Cell cs = cell_of(X, s);
      sedi_ice_melt<double>(c, tice_d, P.nml.ql_mlt, cs);
      set_cell(X, s, cs);
    }
  }
  tape_save(T, it, RS_G, X, sh);
  double dz[S], flux[S], surf[S], m1_sol[S], dm[S];
#pragma unroll
  for (int s = 0; s < S; ++s) m1_sol[s] = 0.0;
  layer_dz(L, sh, dz);
  if (sw) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int idx = 64 * s + lane;
      T.at_sw(it, SW_WS)[idx] = wv[s];
      T.at_sw(it, SW_QV)[idx] = X.qv[s];
      T.at_sw(it, SW_QL)[idx] = X.ql[s];
      T.at_sw(it, SW_QR)[idx] = X.qr[s];
    }
  }
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.qi[s] > QRMIN);
  col_any(L, pred, fall);
  if (!(P.nml.vi_fac < 1.0e-5) && wave_any(fall)) {
    if (sw)
#pragma unroll
      for (int s = 0; s < S; ++s) dm[s] = total_mass(X.dp1[s], X.qv[s], X.ql[s], X.qr[s], X.qi[s], X.qs[s], X.qg[s]);
    implicit_fall(L, X, sh, G, f, dz, X.qi, vti, P.dts, fall, false, flux, surf);
    if (sw) w_transport(L, sh, dm, flux, vti, fall, wv);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      m1_sol[s] = flux[s];
      if (L.v[s] && L.kb[s] == 0) sh.colv[CV_ICE][L.cl[s]] += surf[s];
    }
  }
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.qs[s] > QRMIN);
  col_any(L, pred, fall);
  if (wave_any(fall)) {
    if (sw)
#pragma unroll
      for (int s = 0; s < S; ++s) dm[s] = total_mass(X.dp1[s], X.qv[s], X.ql[s], X.qr[s], X.qi[s], X.qs[s], X.qg[s]);
    implicit_fall(L, X, sh, G, f, dz, X.qs, X.vts, P.dts, fall, false, flux, surf);
    if (sw) w_transport(L, sh, dm, flux, X.vts, fall, wv);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      m1_sol[s] = m1_sol[s] + flux[s];
      if (L.v[s] && L.kb[s] == 0) sh.colv[CV_SNOW][L.cl[s]] += surf[s];
    }
  }
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (X.qg[s] > QRMIN);
  col_any(L, pred, fall);
  if (wave_any(fall)) {
    if (sw)
#pragma unroll
      for (int s = 0; s < S; ++s) dm[s] = total_mass(X.dp1[s], X.qv[s], X.ql[s], X.qr[s], X.qi[s], X.qs[s], X.qg[s]);
    implicit_fall(L, X, sh, G, f, dz, X.qg, X.vtg, P.dts, fall, false, flux, surf);
    if (sw) w_transport(L, sh, dm, flux, X.vtg, fall, wv);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      m1_sol[s] = m1_sol[s] + flux[s];
      if (L.v[s] && L.kb[s] == 0) sh.colv[CV_GRAUPEL][L.cl[s]] += surf[s];
    }
  }
  __syncthreads();
  tape_save(T, it, RS_H, X, sh);
  if (P.nml.do_sedi_heat) {
    bool all[S];
#pragma unroll
    for (int s = 0; s < S; ++s) all[s] = L.v[s];
    sedi_heat(L, X, sh, m1_sol, C_ICE, all);
  }
#pragma unroll
  for (int s = 0; s < S; ++s) sh.cell[CL_M1SOL][64 * s + lane] = m1_sol[s];
}


// ----------------------------------------------------------- reverse steps
// Register-side state of the reverse walk of one column.
template <int S>
struct AdjRegs {
  CellBar<S> xb;     // cotangent of the cell values
  double gp[AP_N];   // parameter gradient (this lane's share)
  double wb[S];      // cotangent of the vertical velocity (do_sedi_w)
};



template <int S>
__device__ __forceinline__ double& abar(AdjShared<S>& A, int which, int s) {
  return A.bar[which][64 * s + lane_id()];
}

// Evaporation / accretion of rain (revap_racc) at record r, in falling
// columns (hand-written reverse, mp_point_adj.h).  It changes only cells
// warmer than t_wfr with rain; as in the forward, those are packed into one
// slot when there are at most 64 of them.  qi, qs and qg enter through their
// sum, packed in qi's place, and each gains the cotangent the sum gets (the
// reverse adds it to each once), so the result is that of cell order.
template <int S>
__device__ __forceinline__ void revap_adj(const Lay<S>& L, AdjShared<S>& A, AdjRegs<S>& G, const Rec<S>& r,
                                          const bool (&fall)[S], const Coef<double>& c, double t_wfr, double dt5) {
  bool pred[S];
#pragma unroll
  for (int s = 0; s < S; ++s)
    pred[s] = fall[s] && L.v[s] && (r.z[Z_TZ][s] > t_wfr) && (r.z[Z_QR][s] > QRMIN);
  const Packed<S> pk = pack_cells(pred);
  if (pk.n <= 64) {
    if (pk.n == 0) return;
    double q_sol[S];
#pragma unroll
    for (int s = 0; s < S; ++s) q_sol[s] = r.z[Z_QI][s] + r.z[Z_QS][s] + r.z[Z_QG][s];
    const Cell cs{gather_packed(r.z[Z_QV], pk.src), gather_packed(r.z[Z_QL], pk.src),
                  gather_packed(r.z[Z_QR], pk.src), gather_packed(q_sol, pk.src), 0.0, 0.0,
                  gather_packed(r.z[Z_TZ], pk.src), 0.0};
    IcX xb{gather_packed(G.xb.v[0], pk.src), gather_packed(G.xb.v[1], pk.src), gather_packed(G.xb.v[2], pk.src),
           0.0, 0.0, 0.0, gather_packed(G.xb.v[6], pk.src), 0.0, 0.0, 0.0, 0.0};
    if (lane_id() < pk.n) {
      const int id = pk.src;
      // column of cell id: floor((id + 1/2) / nz) in FP32 is exact here
      const int cl = int((float(id) + 0.5f) * (1.0f / float(L.nz)));
      double denb = 0.0, denfacb = 0.0;
      revap_vjp(c, t_wfr, dt5, A.sh.colv[CV_HVAR][cl], cs, A.sh.cell[CL_DEN][id], A.sh.cell[CL_DENFAC][id], xb,
                denb, denfacb);
      A.bar[AB_DEN][id] += denb;
      A.bar[AB_DENFAC][id] += denfacb;
    }
    scatter_packed(G.xb.v[0], xb.qv, pk, pred);
    scatter_packed(G.xb.v[1], xb.ql, pk, pred);
    scatter_packed(G.xb.v[2], xb.qr, pk, pred);
    scatter_packed(G.xb.v[6], xb.tz, pk, pred);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const double t = pull(xb.qi, pk.pos[s]);
      if (pred[s]) {
        G.xb.v[3][s] += t;
        G.xb.v[4][s] += t;
        G.xb.v[5][s] += t;
      }
    }
    __syncthreads();
    return;
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (!(fall[s] && L.v[s])) continue;
    IcX xb{G.xb.v[0][s], G.xb.v[1][s], G.xb.v[2][s], G.xb.v[3][s], G.xb.v[4][s], G.xb.v[5][s], G.xb.v[6][s],
           G.xb.v[7][s], 0.0, 0.0, 0.0};
    double denb = 0.0, denfacb = 0.0;
    revap_vjp(c, t_wfr, dt5, A.sh.colv[CV_HVAR][L.cl[s]],
              Cell{r.z[Z_QV][s], r.z[Z_QL][s], r.z[Z_QR][s], r.z[Z_QI][s], r.z[Z_QS][s], r.z[Z_QG][s], r.z[Z_TZ][s],
                   0.0},
              r.z[Z_DEN][s], r.z[Z_DENFAC][s], xb, denb, denfacb);
    G.xb.v[0][s] = xb.qv;
    G.xb.v[1][s] = xb.ql;
    G.xb.v[2][s] = xb.qr;
    G.xb.v[3][s] = xb.qi;
    G.xb.v[4][s] = xb.qs;
    G.xb.v[5][s] = xb.qg;
    G.xb.v[6][s] = xb.tz;
    abar(A, AB_DEN, s) += denb;
    abar(A, AB_DENFAC, s) += denfacb;
  }
}

// Autoconversion at record r, with the liquid slope limiter's dl (hand-written
// reverse, mp_point_adj.h): cotangents of ql, qr, den, ccn, c_praut, dl and
// fac_rc.
template <int S>
__device__ __forceinline__ void autoconv_adj(const Lay<S>& L, AdjShared<S>& A, AdjRegs<S>& G, const Rec<S>& r,
                                             const shield_mp_params& P, const Coef<double>& c) {
  Shared<S>& sh = A.sh;
  const bool slope = P.nml.z_slope_liq != 0 && P.nml.irain_f == 0;
  double ql[S], dl[S], dlbar[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    ql[s] = r.z[Z_QL][s];
    dlbar[s] = 0.0;
  }
  if (slope) slope_limiter(L, sh, ql, dl);
  double ccnbar[S], cprautbar[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    ccnbar[s] = cprautbar[s] = 0.0;
    if (!L.v[s]) continue;
    const int cl = L.cl[s];
    const double hv = sh.colv[CV_HVAR][cl], ccn = sh.colv[CV_CCN][cl], cpraut = sh.colv[CV_CPRAUT][cl];
    double qlb = G.xb.v[1][s], qrb = G.xb.v[2][s];
    AcB ab{0.0, 0.0, 0.0, 0.0, 0.0};
    const double tz = r.z[Z_TZ][s], qlv = r.z[Z_QL][s], den = r.z[Z_DEN][s];
    if (P.nml.irain_f != 0) {
      autoconv_vjp(c, P.t_wfr, false, tz, qlv, den, ccn, cpraut, 0.0, qlb, qrb, ab);
    } else {
      const double base = mmax(QVMIN, hv * qlv);
      const bool use_dl = slope && (dl[s] > base);
      autoconv_vjp(c, P.t_wfr, true, tz, qlv, den, ccn, cpraut, use_dl ? dl[s] : base, qlb, qrb, ab);
      if (use_dl)
        dlbar[s] = ab.dl;
      else if (!(QVMIN > hv * qlv))
        qlb += ab.dl * hv;
    }
    G.xb.v[1][s] = qlb;
    G.xb.v[2][s] = qrb;
    abar(A, AB_DEN, s) += ab.den;
    ccnbar[s] = ab.ccn;
    cprautbar[s] = ab.c_praut;
    G.gp[AP_FAC_RC] += ab.fac_rc;
  }
  // per-column sums
  for (int cc = 0; cc < L.ncols; ++cc) {
    double a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (L.v[s] && L.cl[s] == cc) {
        a1 += ccnbar[s];
        a2 += cprautbar[s];
      }
    }
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    if (lane_id() == 0) {
      A.colbar[CB_CCN][cc] += a1;
      A.colbar[CB_CPRAUT][cc] += a2;
    }
  }
  __syncthreads();
  if (slope) {
    // dl[k] = slope_point(ql[k-1], ql[k], ql[k+1]) at interior levels
    double qp[S], qn[S], gm[S], gp[S], gmn[S], gpp[S];
    col_prev(sh.buf, ql, qp);
    col_next(sh.buf, ql, qn);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      gm[s] = gp[s] = 0.0;
      if (!L.v[s] || L.k[s] == 0 || L.kb[s] == 0 || dlbar[s] == 0.0) continue;
      const Dual dm = slope_point<Dual>(Dual(qp[s], 1.0), Dual(ql[s]), Dual(qn[s]));
      const Dual d0 = slope_point<Dual>(Dual(qp[s]), Dual(ql[s], 1.0), Dual(qn[s]));
      const Dual dp = slope_point<Dual>(Dual(qp[s]), Dual(ql[s]), Dual(qn[s], 1.0));
      G.xb.v[1][s] += dlbar[s] * d0.d;
      gm[s] = dlbar[s] * dm.d;  // lands on level k-1
      gp[s] = dlbar[s] * dp.d;  // lands on level k+1
    }
    col_next(sh.buf, gm, gmn);
    col_prev(sh.buf, gp, gpp);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (!L.v[s]) continue;
      if (L.kb[s] > 0) G.xb.v[1][s] += gmn[s];
      if (L.k[s] > 0) G.xb.v[1][s] += gpp[s];
    }
  }
}

// The reverse of one warm_rain call (records rb .. rb+3).
template <int S, class FS>
__device__ __forceinline__ void warm_rain_adj(const Lay<S>& L, AdjShared<S>& A, AdjRegs<S>& G, const Glob<S>& Gl,
                                              const FS& f, const shield_mp_params& P,
                                              const Coef<double>& c, bool is_first,
                                              const TapeView<S>& T, int it, int rb, const double (&dp1)[S],
                                              bool sw) {
  Shared<S>& sh = A.sh;
  const double t_wfr = P.t_wfr;
  const double dt5 = 0.5 * c.dt_rain;
  Rec<S> rB;
  tape_load(T, it, rb, rB, sh);
  bool pred[S], fall[S];
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (rB.z[Z_QR][s] > QRMIN);
  col_any(L, pred, fall);
  // m1 += flux (+ m1_sol); rain += surface flux
  double fluxbar[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const double m1b = abar(A, AB_M1, s);
    fluxbar[s] = m1b;
    if (!is_first) abar(A, AB_M1SOL, s) += m1b;
    if (L.v[s] && L.kb[s] == 0) fluxbar[s] += A.colbar[CB_RAIN][L.cl[s]];
  }
  {
    Rec<S> rE;
    tape_load(T, it, rb + 3, rE, sh);
    autoconv_adj(L, A, G, rE, P, c);
  }
  if (wave_any(fall)) {
    {
      Rec<S> rD;
      tape_load(T, it, rb + 2, rD, sh);
      revap_adj(L, A, G, rD, fall, c, t_wfr, dt5);
    }
    Rec<S> rC;
    tape_load(T, it, rb + 1, rC, sh);
    double qr[S], vt[S], dz1[S], qn[S], flux[S], vtbar[S], dz1bar[S], dp1bar[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      qr[s] = rC.z[Z_QR][s];
      vt[s] = rC.z[Z_VTR][s];
      dz1[s] = rC.z[Z_DZ1][s];
      vtbar[s] = dz1bar[s] = dp1bar[s] = 0.0;
    }
    fall_fwd(L, sh, qr, vt, dz1, dp1, P.dt_rain, fall, qn, flux);
    if (P.nml.do_sedi_heat) {
      Rec<S> st = rC;
#pragma unroll
      for (int s = 0; s < S; ++s) st.z[Z_QR][s] = qn[s];
      heat_adj(L, sh, st, dp1, flux, C_LIQ, fall, G.xb, fluxbar, dp1bar, dz1bar);
    }
    // the w transport (do_sedi_w), inside the fall: its total mass is that of
    // the cell before the fall, whose rain cotangent fall_adj forms
    double dmq[S];
#pragma unroll
    for (int s = 0; s < S; ++s) dmq[s] = 0.0;
    if (sw) {
      const int lane = lane_id();
      double dm[S], w0[S], dmb[S];
#pragma unroll
      for (int s = 0; s < S; ++s) {
        dm[s] = total_mass(dp1[s], rC.z[Z_QV][s], rC.z[Z_QL][s], rC.z[Z_QR][s], rC.z[Z_QI][s], rC.z[Z_QS][s],
                           rC.z[Z_QG][s]);
        w0[s] = T.at_sw(it, is_first ? SW_W1 : SW_W2)[64 * s + lane];
        dmb[s] = 0.0;
      }
      w_transport_adj(L, sh, dm, flux, vt, fall, w0, G.wb, dmb, fluxbar, vtbar);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        dp1bar[s] += dmb[s] * (dm[s] * rcp64(dp1[s]));
        const double g = dmb[s] * dp1[s];
        G.xb.v[0][s] += g;
        G.xb.v[1][s] += g;
        G.xb.v[3][s] += g;
        G.xb.v[4][s] += g;
        G.xb.v[5][s] += g;
        dmq[s] = g;
      }
    }
    fall_adj(L, sh, qr, vt, dz1, dp1, P.dt_rain, fall, G.xb.v[2], fluxbar, vtbar, dz1bar, dp1bar);
#pragma unroll
    for (int s = 0; s < S; ++s) G.xb.v[2][s] += dmq[s];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      abar(A, AB_VTR, s) += vtbar[s];
      abar(A, AB_DZ1, s) += dz1bar[s];
      abar(A, AB_DP1, s) += dp1bar[s];
    }
    revap_adj(L, A, G, rB, fall, c, t_wfr, dt5);
  }
  // rain fall speed: inputs qr, den, vr_fac
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const double vb = abar(A, AB_VTR, s);
    abar(A, AB_VTR, s) = 0.0;
    if (!L.v[s] || vb == 0.0) continue;
    double qrb = 0.0, denb = 0.0, vrb = 0.0;
    rain_fspeed_vjp(c, !fall[s], rB.z[Z_QR][s], rB.z[Z_DEN][s], vb, qrb, denb, vrb);
    G.xb.v[2][s] += qrb;
    abar(A, AB_DEN, s) += denb;
    G.gp[AP_VR_FAC] += vrb;
  }
  if (is_first) {
    // dz1 = dz0 tz / t0, den = den0 dz0 / dz1 = -dp1 / (g dz1), denfac = sqrt(sfcrho / den)
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const double dzb = abar(A, AB_DZ1, s), db = abar(A, AB_DEN, s), dfb = abar(A, AB_DENFAC, s);
      abar(A, AB_DZ1, s) = abar(A, AB_DEN, s) = abar(A, AB_DENFAC, s) = 0.0;
      if (!L.v[s]) continue;
      const int32_t o = Gl.off[s];
      const double t0 = f.pt[o], dz0 = f.delz[o], tz = rB.z[Z_TZ][s];
      const double dz1 = rB.z[Z_DZ1][s], den = rB.z[Z_DEN][s], denfac = rB.z[Z_DENFAC][s];
      const double denb = db - dfb * 0.5 * denfac * rcp64(den);
      const double dz1b = dzb - denb * den * rcp64(dz1);
      G.xb.v[6][s] += dz1b * dz0 * rcp64(t0);
      abar(A, AB_DZ0, s) += dz1b * tz * rcp64(t0);
      abar(A, AB_T0, s) -= dz1b * dz1 * rcp64(t0);
      abar(A, AB_DP1, s) += denb * den * rcp64(dp1[s]);
    }
  }
}

// The reverse of sedimentation (records RS_F, RS_G, RS_H).
template <int S>
__device__ __forceinline__ void sed_adj(const Lay<S>& L, AdjShared<S>& A, AdjRegs<S>& G, const shield_mp_params& P,
                                        const Coef<double>& c, const TapeView<S>& T, int it,
                                        const double (&dp1)[S], bool sw) {
  Shared<S>& sh = A.sh;
  const double tice_d = P.nml.tice;
  Rec<S> rF, rG;
  tape_load(T, it, RS_F, rF, sh);
  tape_load(T, it, RS_G, rG, sh);
  double vti[S], dz1[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    double vg, vi, vs;
    fall_speed<double>(c, rF.z[Z_QG][s], rF.z[Z_QI][s], rF.z[Z_QS][s], rF.z[Z_TZ][s] - tice_d, rF.z[Z_DEN][s], vg,
                       vi, vs);
    vti[s] = vi;
    dz1[s] = rG.z[Z_DZ1][s];
  }
  // the three falls from the post-melt state
  bool pred[S], fi[S], fs[S], fg[S];
  bool any_i, any_s, any_g;
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (rG.z[Z_QI][s] > QRMIN);
  col_any(L, pred, fi);
  any_i = !(P.nml.vi_fac < 1.0e-5) && wave_any(fi);
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (rG.z[Z_QS][s] > QRMIN);
  col_any(L, pred, fs);
  any_s = wave_any(fs);
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (rG.z[Z_QG][s] > QRMIN);
  col_any(L, pred, fg);
  any_g = wave_any(fg);
  double qi[S], qs[S], qg[S], vts[S], vtg[S], qn[S], fl_i[S], fl_s[S], fl_g[S], m1s[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    qi[s] = rG.z[Z_QI][s];
    qs[s] = rG.z[Z_QS][s];
    qg[s] = rG.z[Z_QG][s];
    vts[s] = rG.z[Z_VTS][s];
    vtg[s] = rG.z[Z_VTG][s];
    fl_i[s] = fl_s[s] = fl_g[s] = 0.0;
  }
  // the species after their falls (the total mass of a later fall's w
  // transport reads them)
  double qni[S], qns[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    qni[s] = qi[s];
    qns[s] = qs[s];
  }
  if (any_i) fall_fwd(L, sh, qi, vti, dz1, dp1, P.dts, fi, qni, fl_i);
  if (any_s) fall_fwd(L, sh, qs, vts, dz1, dp1, P.dts, fs, qns, fl_s);
  if (any_g) fall_fwd(L, sh, qg, vtg, dz1, dp1, P.dts, fg, qn, fl_g);
  // do_sedi_w: w before each solid fall and each fall's total mass
  double w0[S], w1[S], w2[S], dmi[S], dms[S], dmg[S];
  if (sw) {
    const int lane = lane_id();
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int idx = 64 * s + lane;
      const double qv = T.at_sw(it, SW_QV)[idx], ql = T.at_sw(it, SW_QL)[idx], qr = T.at_sw(it, SW_QR)[idx];
      w0[s] = w1[s] = T.at_sw(it, SW_WS)[idx];
      dmi[s] = total_mass(dp1[s], qv, ql, qr, qi[s], qs[s], qg[s]);
      dms[s] = total_mass(dp1[s], qv, ql, qr, qni[s], qs[s], qg[s]);
      dmg[s] = total_mass(dp1[s], qv, ql, qr, qni[s], qns[s], qg[s]);
    }
    if (any_i) w_transport(L, sh, dmi, fl_i, vti, fi, w1);
#pragma unroll
    for (int s = 0; s < S; ++s) w2[s] = w1[s];
    if (any_s) w_transport(L, sh, dms, fl_s, vts, fs, w2);
  }
  double fluxbar[S], dz1bar[S], dp1bar[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    m1s[s] = fl_i[s] + fl_s[s] + fl_g[s];
    fluxbar[s] = abar(A, AB_M1SOL, s);
    abar(A, AB_M1SOL, s) = 0.0;
    dz1bar[s] = dp1bar[s] = 0.0;
  }
  if (P.nml.do_sedi_heat) {
    Rec<S> rH;
    tape_load(T, it, RS_H, rH, sh);
    bool all[S];
#pragma unroll
    for (int s = 0; s < S; ++s) all[s] = L.v[s];
    heat_adj(L, sh, rH, dp1, m1s, C_ICE, all, G.xb, fluxbar, dp1bar, dz1bar);
  }
  // graupel, snow, ice falls in reverse order; each flux also carries its
  // surface value into the column's precipitation
  // With do_sedi_w each fall's w transport is reversed first (it adds to the
  // fall's flux and speed cotangents); the cotangent of its total mass goes
  // to the species as that fall saw them: the falling species' share after
  // the fall's transpose (before the fall), the others' at once (the later
  // falls' transposes have already run, the earlier ones' not yet).
  double vtbar[S], dmb[S];
  auto w_rev = [&](const double (&dm)[S], const double (&fl)[S], const double (&vt)[S], const bool (&fall)[S],
                   const double (&wb0)[S], double (&fb)[S], double (&vb)[S], int own) __attribute__((always_inline)) {
#pragma unroll
    for (int s = 0; s < S; ++s) dmb[s] = 0.0;
    w_transport_adj(L, sh, dm, fl, vt, fall, wb0, G.wb, dmb, fb, vb);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      dp1bar[s] += dmb[s] * (dm[s] * rcp64(dp1[s]));
      const double g = dmb[s] * dp1[s];
      dmb[s] = g;
#pragma unroll
      for (int q = 0; q < 6; ++q)
        if (q != own) G.xb.v[q][s] += g;
    }
  };
  if (any_g) {
    double fb[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      fb[s] = fluxbar[s] + ((L.v[s] && L.kb[s] == 0) ? A.colbar[CB_GRAUPEL][L.cl[s]] : 0.0);
      vtbar[s] = 0.0;
    }
    if (sw) w_rev(dmg, fl_g, vtg, fg, w2, fb, vtbar, 5);
    fall_adj(L, sh, qg, vtg, dz1, dp1, P.dts, fg, G.xb.v[5], fb, vtbar, dz1bar, dp1bar);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      abar(A, AB_VTG, s) += vtbar[s];
      if (sw) G.xb.v[5][s] += dmb[s];
    }
  }
  if (any_s) {
    double fb[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      fb[s] = fluxbar[s] + ((L.v[s] && L.kb[s] == 0) ? A.colbar[CB_SNOW][L.cl[s]] : 0.0);
      vtbar[s] = 0.0;
    }
    if (sw) w_rev(dms, fl_s, vts, fs, w1, fb, vtbar, 4);
    fall_adj(L, sh, qs, vts, dz1, dp1, P.dts, fs, G.xb.v[4], fb, vtbar, dz1bar, dp1bar);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      abar(A, AB_VTS, s) += vtbar[s];
      if (sw) G.xb.v[4][s] += dmb[s];
    }
  }
  double vtibar[S];
#pragma unroll
  for (int s = 0; s < S; ++s) vtibar[s] = 0.0;
  if (any_i) {
    double fb[S];
#pragma unroll
    for (int s = 0; s < S; ++s) fb[s] = fluxbar[s] + ((L.v[s] && L.kb[s] == 0) ? A.colbar[CB_ICE][L.cl[s]] : 0.0);
    if (sw) w_rev(dmi, fl_i, vti, fi, w0, fb, vtibar, 3);
    fall_adj(L, sh, qi, vti, dz1, dp1, P.dts, fi, G.xb.v[3], fb, vtibar, dz1bar, dp1bar);
    if (sw)
#pragma unroll
      for (int s = 0; s < S; ++s) G.xb.v[3][s] += dmb[s];
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    abar(A, AB_DZ1, s) += dz1bar[s];
    abar(A, AB_DP1, s) += dp1bar[s];
  }
  // ice melt at and below the first level warmer than tice (record F)
  bool stop[S];
#pragma unroll
  for (int s = 0; s < S; ++s) pred[s] = L.v[s] && (rF.z[Z_TZ][s] > tice_d);
  col_from_first(L, pred, stop);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (!(stop[s] && L.v[s])) continue;
    IcX xb{G.xb.v[0][s], G.xb.v[1][s], G.xb.v[2][s], G.xb.v[3][s], G.xb.v[4][s], G.xb.v[5][s], G.xb.v[6][s],
           G.xb.v[7][s], 0.0, 0.0, 0.0};
    sedi_ice_melt_vjp(c, tice_d, P.nml.ql_mlt,
                      Cell{rF.z[Z_QV][s], rF.z[Z_QL][s], rF.z[Z_QR][s], rF.z[Z_QI][s], rF.z[Z_QS][s], rF.z[Z_QG][s],
                           rF.z[Z_TZ][s], 0.0},
                      xb);
    G.xb.v[0][s] = xb.qv;
    G.xb.v[1][s] = xb.ql;
    G.xb.v[2][s] = xb.qr;
    G.xb.v[3][s] = xb.qi;
    G.xb.v[4][s] = xb.qs;
    G.xb.v[5][s] = xb.qg;
    G.xb.v[6][s] = xb.tz;
  }
  // fall speeds (record F): inputs qg, qi, qs, tz, den and vs_fac, vg_fac, vi_fac
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const double bg = abar(A, AB_VTG, s), bs = abar(A, AB_VTS, s), bi = vtibar[s];
    abar(A, AB_VTG, s) = abar(A, AB_VTS, s) = 0.0;
    if (!L.v[s]) continue;
    FsB fb{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    fall_speed_vjp(c, rF.z[Z_QG][s], rF.z[Z_QI][s], rF.z[Z_QS][s], rF.z[Z_TZ][s] - tice_d, rF.z[Z_DEN][s], bg, bi, bs,
                   fb);
    G.xb.v[5][s] += fb.qg;
    G.xb.v[3][s] += fb.qi;
    G.xb.v[4][s] += fb.qs;
    G.xb.v[6][s] += fb.tc;
    abar(A, AB_DEN, s) += fb.den;
    G.gp[AP_VS_FAC] += fb.vs_fac;
    G.gp[AP_VG_FAC] += fb.vg_fac;
    G.gp[AP_VI_FAC] += fb.vi_fac;
  }
}

// Cotangents of the call's inputs that are held in registers, cell order,
// until fields_init's reverse stores them.
template <int S>
struct InBar {
  double pt[S], q[6][S], dp0[S], ua[S], va[S], qa[S];
};

// fields_update's outputs at one cell, read from the call's inputs and
// output cotangents
struct FuIn {
  double dp0, dp1, pt0, gt, ga, qin[6], gq[6];
};
// cotangents fields_update's reverse gives the call's inputs at one cell
struct FuB {
  double pt, dp0, dp1, qa, q[6];
};
// Transpose of fields_update (microphysics.py:1452-1553) at one cell: the
// tendencies of the final cell x, (x q - q_in dp0 / dp1) dp1 / dp0 / dt for
// the species, (x tz - pt) cvm(x) / cp_air / dt for temperature, and the
// cloud-fraction accumulator's tendency.  Adds to yb (the cotangents of x)
// and to b.
__device__ __forceinline__ void fields_update_adj(const shield_mp_params& P, double rdt, const Cell& x, const FuIn& in,
                                                  double (&yb)[8], FuB& b) {
  const double omq0 = in.dp0 / in.dp1, omq = in.dp1 / in.dp0;
  const double xq[6] = {x.qv, x.ql, x.qr, x.qi, x.qs, x.qg};
  double omqb = 0.0, omq0b = 0.0;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    yb[q] += in.gq[q] * rdt * omq;
    b.q[q] -= in.gq[q] * rdt * omq0 * omq;
    omqb += in.gq[q] * rdt * (xq[q] - in.qin[q] * omq0);
    omq0b -= in.gq[q] * rdt * in.qin[q] * omq;
  }
  b.dp1 += omqb * rcp64(in.dp0) - omq0b * omq0 * rcp64(in.dp1);
  b.dp0 += omq0b * rcp64(in.dp1) - omqb * omq * rcp64(in.dp0);
  const double cvm = P.c_air + x.qv * P.c_vap + (x.qr + x.ql) * C_LIQ + (x.qi + x.qs + x.qg) * C_ICE;
  yb[6] += in.gt * rdt * cvm * (1.0 / CP_AIR);
  b.pt -= in.gt * rdt * cvm * (1.0 / CP_AIR);
  const double cvmb = in.gt * rdt * (x.tz - in.pt0) * (1.0 / CP_AIR);
  yb[0] += cvmb * P.c_vap;
  yb[1] += cvmb * C_LIQ;
  yb[2] += cvmb * C_LIQ;
  yb[3] += cvmb * C_ICE;
  yb[4] += cvmb * C_ICE;
  yb[5] += cvmb * C_ICE;
  if (!P.nml.do_qa) {
    yb[7] += in.ga * rdt * rcp64(P.ntimes);
    b.qa -= in.ga * rdt;
  }
}

// The reverse of icloud with the ice slope limiter (z_slope_ice), in cell
// order as the forward runs it.  The forward melts / freezes every cell of
// the column first, then takes the limiter's dl of the post-melt cloud ice
// for each cell's ice autoconversion width.  So the reverse runs each cell's
// icloud reverse down to the state after melt / freeze (its dl gaining a
// cotangent), transposes the limiter across the column (three dual passes
// per interior level, scattered to k-1 and k+1 as for the liquid limiter),
// and then runs each cell's melt / freeze reverse.  In the last sub-step
// fields_update's reverse runs inside, as in icloud_adj.
template <int S, class FS, class FA>
__device__ __forceinline__ void icloud_adj_slope(const Lay<S>& L, AdjShared<S>& A, AdjRegs<S>& G, const Glob<S>& Gl,
                                                 const FS& f, const FA& a, const shield_mp_params& P,
                                                 const Coef<double>& c, const TapeView<S>& T, int it, bool last,
                                                 InBar<S>& nb) {
  Shared<S>& sh = A.sh;
  const int lane = lane_id();
  const double rdt = P.rdt;
  Cell x0[S];
  double qmf[S], dl[S], p1[S], dp1[S];
  // the post-melt cloud ice of every cell, then the limiter over the column
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int idx = 64 * s + lane;
    x0[s] = Cell{T.at(it, RI, Z_QV)[idx], T.at(it, RI, Z_QL)[idx], T.at(it, RI, Z_QR)[idx], T.at(it, RI, Z_QI)[idx],
                 T.at(it, RI, Z_QS)[idx], T.at(it, RI, Z_QG)[idx], T.at(it, RI, Z_TZ)[idx], 0.0};
    qmf[s] = 0.0;
    p1[s] = dp1[s] = 1.0;
    if (!L.v[s]) continue;
    const int32_t o = Gl.off[s];
    dp1[s] = f.delp[o] * (1.0 - f.qv[o]);  // fields_init's dp1, the same expression
    p1[s] = dry_density<double>(dp1[s], f.delz[o]) * RDGAS * f.pt[o];
    const IcK k{c, P.nml.tice, P.t_wfr, P.nml.ql_mlt, P.tice0, 0.0, 0.0, 0.0,
                {sh.cell[CL_DEN][idx], sh.cell[CL_DENFAC][idx], 0.0, 0.0, 0.0, p1[s]}};
    IcX y{x0[s].qv, x0[s].ql, x0[s].qr, x0[s].qi, x0[s].qs, x0[s].qg, x0[s].tz, 0.0, 0.0, 0.0, 0.0};
    blk_melt_freeze(k, y);
    qmf[s] = y.qi;
  }
  slope_limiter(L, sh, qmf, dl);
  IcX xbm[S];
  double dlb[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    dlb[s] = 0.0;
    xbm[s] = IcX{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  }
  // unrolled by construction: a rolled loop would keep the per-slot arrays in scratch
  for_slots<S>([&](auto si) __attribute__((always_inline)) {
    constexpr int s = decltype(si)::value;
    if (!L.v[s]) return;
    const int idx = 64 * s + lane;
    const int cl = L.cl[s];
    const int32_t o = Gl.off[s];
    IcK k{c, P.nml.tice, P.t_wfr, P.nml.ql_mlt, P.tice0, sh.colv[CV_HVAR][cl], sh.colv[CV_RHADJ][cl],
          sh.colv[CV_RHRAIN][cl],
          {sh.cell[CL_DEN][idx], sh.cell[CL_DENFAC][idx], T.at(it, RS_G, Z_VTG)[idx], T.at(it, RW2_C, Z_VTR)[idx],
           T.at(it, RS_G, Z_VTS)[idx], p1[s]}};
    k.slope = true;
    k.dl = dl[s];
    double yb[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) yb[q] = last ? 0.0 : G.xb.v[q][s];
    IcInB ib{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    FuB fb{0.0, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}};
    icloud_vjp<true>(
        k, x0[s], yb, ib, nullptr,
        [&](const Cell& x1, double (&y)[8]) {
          if (!last) return;
          const FuIn in{f.delp[o], dp1[s], f.pt[o], a.pt_dt[o], a.qa_dt[o],
                        {f.qv[o], f.ql[o], f.qr[o], f.qi[o], f.qs[o], f.qg[o]},
                        {a.qv_dt[o], a.ql_dt[o], a.qr_dt[o], a.qi_dt[o], a.qs_dt[o], a.qg_dt[o]}};
          fields_update_adj(P, rdt, x1, in, y, fb);
        },
        &xbm[s]);
    if (last) {
      nb.pt[s] = fb.pt;
      nb.dp0[s] += fb.dp0;
      nb.qa[s] = fb.qa;
#pragma unroll
      for (int q = 0; q < 6; ++q) nb.q[q][s] = fb.q[q];
      abar(A, AB_DP1, s) += fb.dp1;
      G.xb.v[7][s] = yb[7];
    }
    dlb[s] = ib.dl;
    abar(A, AB_DEN, s) += ib.den;
    abar(A, AB_DENFAC, s) += ib.denfac;
    abar(A, AB_VTG, s) += ib.vtg;
    abar(A, AB_VTR, s) += ib.vtr;
    abar(A, AB_VTS, s) += ib.vts;
    G.gp[AP_QI_GEN] += ib.qi_gen;
    // p1 = -dp1 / (g dz0) rdgas t0
    abar(A, AB_DP1, s) += ib.p1 * p1[s] * rcp64(dp1[s]);
    abar(A, AB_DZ0, s) -= ib.p1 * p1[s] * rcp64(f.delz[o]);
    abar(A, AB_T0, s) += ib.p1 * p1[s] * rcp64(f.pt[o]);
  });
  // the limiter's transpose: dl[k] = slope_point(q[k-1], q[k], q[k+1]) at
  // interior levels
  {
    double qp[S], qn[S], gm[S], gp[S], gmn[S], gpp[S];
    col_prev(sh.buf, qmf, qp);
    col_next(sh.buf, qmf, qn);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      gm[s] = gp[s] = 0.0;
      if (!L.v[s] || L.k[s] == 0 || L.kb[s] == 0 || dlb[s] == 0.0) continue;
      const Dual dm = slope_point<Dual>(Dual(qp[s], 1.0), Dual(qmf[s]), Dual(qn[s]));
      const Dual d0 = slope_point<Dual>(Dual(qp[s]), Dual(qmf[s], 1.0), Dual(qn[s]));
      const Dual dp = slope_point<Dual>(Dual(qp[s]), Dual(qmf[s]), Dual(qn[s], 1.0));
      xbm[s].qi += dlb[s] * d0.d;
      gm[s] = dlb[s] * dm.d;  // lands on level k-1
      gp[s] = dlb[s] * dp.d;  // lands on level k+1
    }
    col_next(sh.buf, gm, gmn);
    col_prev(sh.buf, gp, gpp);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (!L.v[s]) continue;
      if (L.kb[s] > 0) xbm[s].qi += gmn[s];
      if (L.k[s] > 0) xbm[s].qi += gpp[s];
    }
  }
  // melt / freeze, reversed
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (!L.v[s]) continue;
    const int idx = 64 * s + lane;
    const IcK k{c, P.nml.tice, P.t_wfr, P.nml.ql_mlt, P.tice0, 0.0, 0.0, 0.0,
                {sh.cell[CL_DEN][idx], sh.cell[CL_DENFAC][idx], 0.0, 0.0, 0.0, p1[s]}};
    IcInB ib{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const IcX xe{x0[s].qv, x0[s].ql, x0[s].qr, x0[s].qi, x0[s].qs, x0[s].qg, x0[s].tz, 0.0, 0.0, 0.0, 0.0};
    blk_melt_freeze_rev(k, xe, xbm[s], ib);
    G.xb.v[0][s] = xbm[s].qv;
    G.xb.v[1][s] = xbm[s].ql;
    G.xb.v[2][s] = xbm[s].qr;
    G.xb.v[3][s] = xbm[s].qi;
    G.xb.v[4][s] = xbm[s].qs;
    G.xb.v[5][s] = xbm[s].qg;
    G.xb.v[6][s] = xbm[s].tz;
    abar(A, AB_DEN, s) += ib.den;
    G.gp[AP_QI_GEN] += ib.qi_gen;
  }
  __syncthreads();
}

// The reverse of icloud (record RI): melt / freeze, then icloud_main with
// subgrid_z_proc, on the cells sorted by condensate as the forward's icloud
// sorts them (some condensate above QRMIN, above QCMIN, none), so a slot of
// cells without condensate skips the processes that need it.  Each cell's
// arithmetic is that of cell order; only the order of the qi_gen partial sum
// differs.  The cell cotangents move to sorted order and back through the LDS
// cells the reverse walk leaves free (buf, m1, m1_sol, tmp); the inputs of a
// sorted cell are read at its cell index: the record and fall speeds from the
// tape, den and denfac from LDS, t0, dz0 and dp1 = delp (1 - qv) from the
// call's inputs.
//
// In the last sub-step (`last`) the replay has not run icloud: its forward
// sweep here forms the call's final cell, and fields_update's reverse runs on
// it at once (fields_update_adj), giving the cotangents of icloud's outputs
// and adding the input cotangents of fields_update to nb.
template <int S, class FS, class FA>
__device__ __forceinline__ void icloud_adj(const Lay<S>& L, AdjShared<S>& A, AdjRegs<S>& G, const Glob<S>& Gl,
                                           const FS& f, const FA& a, const shield_mp_params& P,
                                           const Coef<double>& c, const TapeView<S>& T, int it, bool last,
                                           InBar<S>& nb, bool slope) {
  if (slope) {
    icloud_adj_slope(L, A, G, Gl, f, a, P, c, T, it, last, nb);
    return;
  }
  Shared<S>& sh = A.sh;
  const int lane = lane_id();
  int pos[S], src[S], osrc[S];
  {
    int cls[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int idx = 64 * s + lane;
      const double qm = mmax(mmax(mmax(T.at(it, RI, Z_QL)[idx], T.at(it, RI, Z_QR)[idx]),
                                  mmax(T.at(it, RI, Z_QI)[idx], T.at(it, RI, Z_QS)[idx])),
                             T.at(it, RI, Z_QG)[idx]);
      cls[s] = !L.v[s] ? 2 : qm > QRMIN ? 0 : qm > QCMIN ? 1 : 2;
    }
    // the cells' HBM offsets ride in the upper half of buf, which the sort
    // (64 S ints) leaves free
    int* const ob = reinterpret_cast<int*>(sh.buf) + 64 * S;
#pragma unroll
    for (int s = 0; s < S; ++s) ob[64 * s + lane] = Gl.off[s];
    sort_cells3(cls, pos, src, reinterpret_cast<int*>(sh.buf));
#pragma unroll
    for (int s = 0; s < S; ++s) osrc[s] = ob[src[s]];
    __syncthreads();
  }
  double* const b1 = sh.cell[CL_M1];
  double* const b2 = sh.cell[CL_M1SOL];
  double* const b3 = sh.cell[CL_TMP];
  // cotangents of the cell after icloud in sorted order (qa's is read by the
  // cloud fraction and passes through icloud unchanged)
  // (zero in the last sub-step, where fields_update's reverse forms them)
  double xs[8][S];
  // fields_update's input cotangents (last sub-step), sorted order
  double fpt[S], fdp0[S], fqa[S], fq[6][S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    fpt[s] = fdp0[s] = fqa[s] = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) fq[q][s] = 0.0;
  }
  if (last) {
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int q = 0; q < 8; ++q) xs[q][s] = 0.0;
  } else {
    put_cells(G.xb.v[0], sh.buf); put_cells(G.xb.v[1], b1); put_cells(G.xb.v[2], b2); put_cells(G.xb.v[3], b3);
    __syncthreads();
    get_cells(xs[0], sh.buf, src); get_cells(xs[1], b1, src); get_cells(xs[2], b2, src); get_cells(xs[3], b3, src);
    __syncthreads();
    put_cells(G.xb.v[4], sh.buf); put_cells(G.xb.v[5], b1); put_cells(G.xb.v[6], b2); put_cells(G.xb.v[7], b3);
    __syncthreads();
    get_cells(xs[4], sh.buf, src); get_cells(xs[5], b1, src); get_cells(xs[6], b2, src); get_cells(xs[7], b3, src);
    __syncthreads();
  }
  const double rdt = P.rdt;
  const int ncell = L.ncols * L.nz;
  // column of cell id: floor((id + 1/2) / nz) in FP32 is exact here, the
  // quotient stays >= 1/(2 nz) away from an integer
  const float rnz = 1.0f / float(L.nz);
  const double* const rq[7] = {T.at(it, RI, Z_QV), T.at(it, RI, Z_QL), T.at(it, RI, Z_QR), T.at(it, RI, Z_QI),
                               T.at(it, RI, Z_QS), T.at(it, RI, Z_QG), T.at(it, RI, Z_TZ)};
  const double* const rvtg = T.at(it, RS_G, Z_VTG);
  const double* const rvts = T.at(it, RS_G, Z_VTS);
  const double* const rvtr = T.at(it, RW2_C, Z_VTR);
  for_slots<S>([&](auto si) __attribute__((always_inline)) {
    constexpr int s = decltype(si)::value;
    const int id = src[s];
    if (id >= ncell) return;
    const int cl = int((float(id) + 0.5f) * rnz);
    const double hv = sh.colv[CV_HVAR][cl], rh_adj = sh.colv[CV_RHADJ][cl], rh_rain = sh.colv[CV_RHRAIN][cl];
    const int32_t o = osrc[s];
    const double t0 = f.pt[o], dz0 = f.delz[o];
    const double dp1 = f.delp[o] * (1.0 - f.qv[o]);  // fields_init's dp1, the same expression
    const double p1 = dry_density<double>(dp1, dz0) * RDGAS * t0;
    const IcK k{c, P.nml.tice, P.t_wfr, P.nml.ql_mlt, P.tice0, hv, rh_adj, rh_rain,
                {sh.cell[CL_DEN][id], sh.cell[CL_DENFAC][id], rvtg[id], rvtr[id], rvts[id], p1}};
    double yb[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) yb[q] = xs[q][s];
    IcInB ib{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const Cell x0{rq[0][id], rq[1][id], rq[2][id], rq[3][id], rq[4][id], rq[5][id], rq[6][id], 0.0};
    // one instance of icloud_vjp for both cases (the sweep forms the final
    // cell in every sub-step; only the last one uses it)
    FuB fb{0.0, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}};
    icloud_vjp<true>(k, x0, yb, ib, nullptr, [&](const Cell& x1, double (&y)[8]) {
      if (!last) return;
      const FuIn in{f.delp[o], dp1, t0, a.pt_dt[o], a.qa_dt[o],
                    {f.qv[o], f.ql[o], f.qr[o], f.qi[o], f.qs[o], f.qg[o]},
                    {a.qv_dt[o], a.ql_dt[o], a.qr_dt[o], a.qi_dt[o], a.qs_dt[o], a.qg_dt[o]}};
      fields_update_adj(P, rdt, x1, in, y, fb);
    });
    if (last) {
      fpt[s] = fb.pt;
      fdp0[s] = fb.dp0;
      fqa[s] = fb.qa;
#pragma unroll
      for (int q = 0; q < 6; ++q) fq[q][s] = fb.q[q];
      A.bar[AB_DP1][id] += fb.dp1;
      xs[7][s] = yb[7];
    }
#pragma unroll
    for (int q = 0; q < 7; ++q) xs[q][s] = yb[q];
    A.bar[AB_DEN][id] += ib.den;
    A.bar[AB_DENFAC][id] += ib.denfac;
    A.bar[AB_VTG][id] += ib.vtg;
    A.bar[AB_VTR][id] += ib.vtr;
    A.bar[AB_VTS][id] += ib.vts;
    G.gp[AP_QI_GEN] += ib.qi_gen;
    // p1 = -dp1 / (g dz0) rdgas t0
    A.bar[AB_DP1][id] += ib.p1 * p1 * rcp64(dp1);
    A.bar[AB_DZ0][id] -= ib.p1 * p1 * rcp64(dz0);
    A.bar[AB_T0][id] += ib.p1 * p1 * rcp64(t0);
  });
  __syncthreads();
  put_cells(xs[0], sh.buf); put_cells(xs[1], b1); put_cells(xs[2], b2); put_cells(xs[3], b3);
  __syncthreads();
  get_cells(G.xb.v[0], sh.buf, pos); get_cells(G.xb.v[1], b1, pos); get_cells(G.xb.v[2], b2, pos);
  get_cells(G.xb.v[3], b3, pos);
  __syncthreads();
  put_cells(xs[4], sh.buf); put_cells(xs[5], b1); put_cells(xs[6], b2);
  __syncthreads();
  get_cells(G.xb.v[4], sh.buf, pos); get_cells(G.xb.v[5], b1, pos); get_cells(G.xb.v[6], b2, pos);
  __syncthreads();
  if (last) {
    // back to cell order: qa's cotangent and fields_update's input cotangents
    double t[S];
    put_cells(xs[7], sh.buf); put_cells(fpt, b1); put_cells(fdp0, b2); put_cells(fqa, b3);
    __syncthreads();
    get_cells(G.xb.v[7], sh.buf, pos); get_cells(nb.pt, b1, pos); get_cells(t, b2, pos); get_cells(nb.qa, b3, pos);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S; ++s) nb.dp0[s] += t[s];
    put_cells(fq[0], sh.buf); put_cells(fq[1], b1); put_cells(fq[2], b2); put_cells(fq[3], b3);
    __syncthreads();
    get_cells(nb.q[0], sh.buf, pos); get_cells(nb.q[1], b1, pos); get_cells(nb.q[2], b2, pos);
    get_cells(nb.q[3], b3, pos);
    __syncthreads();
    put_cells(fq[4], sh.buf); put_cells(fq[5], b1);
    __syncthreads();
    get_cells(nb.q[4], sh.buf, pos); get_cells(nb.q[5], b1, pos);
    __syncthreads();
  }
}


// ------------------------------------------------------------ fix_negative
// Transpose of fix_negative_cell at one cell (the borrowing chain of
// fields_init, microphysics.py:201-240).  x0: the cell at its entry; b: in,
// the cotangents of qv, ql, qr, qi, qs, qg, tz after it; out, before it.  A
// species cleared to zero passes no cotangent back; a borrowing adds the
// borrower's cotangent to the lender's; the two heatings pass theirs to the
// borrowed amount and, through lcpk and icpk, to the entry temperature and
// heat capacity.
SH_FN void fix_negative_cell_vjp(const shield_mp_params& P, const Cell& x0, double (&b)[7]) {
  const double cvm = P.c_air + x0.qv * P.c_vap + (x0.qr + x0.ql) * C_LIQ + (x0.qi + x0.qs + x0.qg) * C_ICE;
  const double tz0 = x0.tz;
  // the forward's branches and borrowed amounts
  const bool c1 = x0.qi < 0.0;
  const double qs1 = c1 ? x0.qs + x0.qi : x0.qs;
  const bool c2 = qs1 < 0.0;
  const double qg2 = c2 ? x0.qg + qs1 : x0.qg;
  const bool c3 = qg2 < 0.0;
  const double qr3 = c3 ? x0.qr + qg2 : x0.qr;
  const bool c4 = qr3 < 0.0;
  const double ql4 = c4 ? x0.ql + qr3 : x0.ql;
  const bool c5 = ql4 < 0.0;
  double cvmb = 0.0, tz0b = 0.0;
  // cloud water from vapour, with the condensation heat
  double qv0b = b[0], ql4b = b[1];
  const double tz3b = b[6];
  if (c5) {
    const double lcpk = (P.lv00 + P.d0_vap * tz0) / cvm;
    ql4b = b[0] - b[6] * lcpk;
    const double lcpkb = -b[6] * ql4;
    tz0b += lcpkb * P.d0_vap * rcp64(cvm);
    cvmb -= lcpkb * lcpk * rcp64(cvm);
  }
  // rain from cloud water
  const double ql0b = ql4b, qr3b = c4 ? ql4b : b[2];
  // graupel from rain, with the fusion heat
  const double qr0b = qr3b;
  double qg2b = b[5];
  if (c3) {
    const double icpk = (LI00 + DC_ICE * tz0) / cvm;
    qg2b = qr3b - tz3b * icpk;
    const double icpkb = -tz3b * qg2;
    tz0b += icpkb * DC_ICE * rcp64(cvm);
    cvmb -= icpkb * icpk * rcp64(cvm);
  }
  tz0b += tz3b;
  // snow from graupel, ice from snow
  const double qg0b = qg2b, qs1b = c2 ? qg2b : b[4];
  const double qs0b = qs1b, qi0b = c1 ? qs1b : b[3];
  b[0] = qv0b + cvmb * P.c_vap;
  b[1] = ql0b + cvmb * C_LIQ;
  b[2] = qr0b + cvmb * C_LIQ;
  b[3] = qi0b + cvmb * C_ICE;
  b[4] = qs0b + cvmb * C_ICE;
  b[5] = qg0b + cvmb * C_ICE;
  b[6] = tz0b;
}

// Transpose of fix_negative_column for one column (q: in, the vapour before
// the column pass; overwritten).  qb: in, the cotangents of the vapour after
// it; out, before it.  dpb accumulates the cotangents of dp.
SH_FN void fix_negative_column_vjp(double* q, const double* dp, int nz, double* qb, double* dpb) {
  // the forward recurrence, kept: q[k] becomes the value after it
  for (int k = 1; k < nz; ++k)
    if (q[k - 1] < 0.0) q[k] = q[k] + q[k - 1] * dp[k - 1] / dp[k];
  const int kb = nz - 1, ka = nz - 2;
  const double qkb = q[kb], qka = q[ka] < 0.0 ? 0.0 : q[ka];
  // the surface layer's borrowing from the layer above
  if ((qkb < 0.0) && (qka > 0.0)) {
    const double a = -qkb * dp[kb], c = qka * dp[ka];
    const double dq = mmin(a, c);
    const double dqb = -qb[ka] * rcp64(dp[ka]) + qb[kb] * rcp64(dp[kb]);
    dpb[ka] += qb[ka] * dq * rcp64(dp[ka] * dp[ka]);
    dpb[kb] -= qb[kb] * dq * rcp64(dp[kb] * dp[kb]);
    if (a < c) {
      qb[kb] -= dqb * dp[kb];
      dpb[kb] -= dqb * qkb;
    } else {
      qb[ka] += dqb * dp[ka];
      dpb[ka] += dqb * qka;
    }
  }
  // negative vapour above the surface layer cleared
  for (int k = 0; k < nz - 1; ++k)
    if (q[k] < 0.0) qb[k] = 0.0;
  // the recurrence, backwards
  for (int k = nz - 1; k >= 1; --k) {
    if (q[k - 1] < 0.0) {
      const double r = dp[k - 1] / dp[k];
      qb[k - 1] += qb[k] * r;
      dpb[k - 1] += qb[k] * q[k - 1] * rcp64(dp[k]);
      dpb[k] -= qb[k] * q[k - 1] * r * rcp64(dp[k]);
    }
  }
}

// The reverse of fix_negative for the wavefront: the entry cells are formed
// again from the inputs as the replay formed them, the cell pass is redone to
// give the column pass its input, and both passes are transposed (the column
// pass one lane per column in LDS, which the reverse walk has left free).
template <int S, class FS>
__device__ __forceinline__ void fix_negative_adj(const Lay<S>& L, AdjShared<S>& A, AdjRegs<S>& G,
                                                 const St<S, double>& X, const Glob<S>& Gl, const FS& f,
                                                 const shield_mp_params& P) {
  Shared<S>& sh = A.sh;
  const int lane = lane_id();
  Cell x0[S];
  double qv5[S];
  bool neg = false;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    x0[s] = Cell{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 250.0, 0.0};
    qv5[s] = 0.0;
    if (!L.v[s]) continue;
    const int32_t o = Gl.off[s];
    const double omq = f.delp[o] / X.dp1[s];
    x0[s] = Cell{launder(f.qv[o] * omq), launder(f.ql[o] * omq), launder(f.qr[o] * omq), launder(f.qi[o] * omq),
                 launder(f.qs[o] * omq), launder(f.qg[o] * omq), f.pt[o], 0.0};
    Cell x5 = x0[s];
    fix_negative_cell<double>(P, x5);
    qv5[s] = x5.qv;
    neg = neg || x5.qv < 0.0;
  }
  if (__ballot(neg) != 0ull) {
    double* const q = sh.buf;
    double* const dp = sh.cell[CL_M1];
    double* const qb = sh.cell[CL_M1SOL];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      q[64 * s + lane] = qv5[s];
      dp[64 * s + lane] = X.dp1[s];
      qb[64 * s + lane] = G.xb.v[0][s];
    }
    __syncthreads();
    if (lane < L.ncols) {
      const int b0 = lane * L.nz;
      fix_negative_column_vjp(q + b0, dp + b0, L.nz, qb + b0, A.bar[AB_DP1] + b0);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S; ++s) G.xb.v[0][s] = qb[64 * s + lane];
    __syncthreads();
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (!L.v[s]) continue;
    double b[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) b[q] = G.xb.v[q][s];
    fix_negative_cell_vjp(P, x0[s], b);
#pragma unroll
    for (int q = 0; q < 7; ++q) G.xb.v[q][s] = b[q];
  }
}


// ------------------------------------------------------------------ kernel
// EXT: the instance with fix_negative, the ice slope limiter and the w
// transport (fix_negative, z_slope_ice, do_sedi_w); without it their code is
// compiled out, since it raises the register pressure of the whole kernel.
template <int S, bool EXT>
__global__ __launch_bounds__(64) void mp_adjoint_kernel(const AdjArgs args) {
  const shield_mp_shape& sh3 = args.sh3;
  const shield_mp_params& P = args.P;
  const Coef<double>& c = args.c;
  // the coefficients and parameters of one stage (karg)
  const auto kc = []() -> const Coef<double>& { return karg<Coef<double>, (uint32_t)offsetof(AdjArgs, c)>(); };
  const auto kP = []() -> const shield_mp_params& {
    return karg<shield_mp_params, (uint32_t)offsetof(AdjArgs, P)>();
  };
  const KFields<(uint32_t)offsetof(AdjArgs, f)> f;
  const KAdjFields<(uint32_t)offsetof(AdjArgs, a)> a;
  double* const tape = args.tape;
  const int64_t tape_stride = args.tape_stride;
  double* const part = args.part;
  const int cols_per_wave = args.cols_per_wave;
  const bool sedi_w = EXT && P.do_sedi_w;
  const bool fixneg = EXT && P.nml.fix_negative;
  __shared__ AdjShared<S> A;
  Shared<S>& sh = A.sh;
  const int lane = lane_id();
  const int nz = (int)sh3.nk;
  const int64_t ncol = sh3.ni * sh3.nj;
  const shield_mp_namelist& n = P.nml;
  double* const tw = tape + (int64_t)blockIdx.x * tape_stride;
  const TapeView<S> T{tw, tw + (int64_t)P.ntimes * REC_FIELDS * 64 * S};
  const double rdt = P.rdt;
  double gp_acc[AP_N];
#pragma unroll
  for (int p = 0; p < AP_N; ++p) gp_acc[p] = 0.0;
  fill_coef_tab(c, lane);
  __syncthreads();

  const int64_t stride = (int64_t)gridDim.x * cols_per_wave;
  for (int64_t c0 = (int64_t)blockIdx.x * cols_per_wave; c0 < ncol; c0 += stride) {
    const int ncols = (int)((ncol - c0) < cols_per_wave ? (ncol - c0) : cols_per_wave);
    Lay<S> L;
    Glob<S> Gl;
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
      const int64_t ci = col / sh3.nj, cj = col - ci * sh3.nj;
      Gl.off[s] = (int32_t)(ci * sh3.si + cj * sh3.sj + L.k[s]);
    }
    // the 2D offset of column `lane` (lanes below ncols)
    int64_t o2 = 0;
    if (lane < ncols) {
      const int64_t col = c0 + lane;
      const int64_t ci = col / sh3.nj, cj = col - ci * sh3.nj;
      o2 = ci * sh3.si2 + cj * sh3.sj2;
    }
    // ------------------------------------------------------------ fields_init
    if (lane < ncols) {
      const double land = f.land[o2], area = f.area[o2];
      const double s_leng = sqrt(sqrt(area * 1.0e-10));
      double h_var = n.dw_land * s_leng * land + n.dw_ocean * s_leng * (1.0 - land);
      h_var = mmin(0.2, mmax(0.01, h_var));
      const double rh_adj = 1.0 - h_var - n.rh_inc;
      sh.colv[CV_HVAR][lane] = h_var;
      sh.colv[CV_RHADJ][lane] = rh_adj;
      sh.colv[CV_RHRAIN][lane] = mmax(0.35, rh_adj - n.rh_inr);
      sh.colv[CV_CCN][lane] = (n.ccn_l * land + n.ccn_o * (1.0 - land)) * 1.0e6;
      sh.colv[CV_RAIN][lane] = sh.colv[CV_SNOW][lane] = sh.colv[CV_ICE][lane] = sh.colv[CV_GRAUPEL][lane] = 0.0;
    }
    __syncthreads();
    St<S, double> X;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (L.v[s]) {
        const int32_t o = Gl.off[s];
        X.tz[s] = f.pt[o];
        const double dp0 = f.delp[o];
        X.dp1[s] = dp0 * (1.0 - f.qv[o]);
        const double omq = dp0 / X.dp1[s];
        // the products are materialised exactly (launder), so fix_negative's
        // reverse, which forms them again, branches as the replay did
        X.qv[s] = launder(f.qv[o] * omq);
        X.ql[s] = launder(f.ql[o] * omq);
        X.qr[s] = launder(f.qr[o] * omq);
        X.qi[s] = launder(f.qi[o] * omq);
        X.qs[s] = launder(f.qs[o] * omq);
        X.qg[s] = launder(f.qg[o] * omq);
        X.qaz[s] = n.reset_qaz ? 0.0 : f.qaz[o];
      } else {
        X.tz[s] = 250.0;
        X.dp1[s] = 100.0;
        X.qv[s] = X.ql[s] = X.qr[s] = X.qi[s] = X.qs[s] = X.qg[s] = X.qaz[s] = 0.0;
      }
      X.vtr[s] = X.vtg[s] = X.vts[s] = VF_MIN;
    }
    double p1_b[S];  // p1 of the surface layer (in the slot and lane of that cell)
#pragma unroll
    for (int s = 0; s < S; ++s) {
      p1_b[s] = 0.0;
      if (L.v[s] && L.kb[s] == 0) {
        double ccn = sh.colv[CV_CCN][L.cl[s]];
        if (P.use_ccn) {
          const int32_t o = Gl.off[s];
          p1_b[s] = dry_density<double>(X.dp1[s], f.delz[o]) * RDGAS * f.pt[o];
          ccn = ccn * RDGAS * X.tz[s] / p1_b[s];
        }
        sh.colv[CV_CCN][L.cl[s]] = ccn;
      }
    }
    __syncthreads();
    if (lane < ncols) sh.colv[CV_CPRAUT][lane] = P.cpaut * pow(sh.colv[CV_CCN][lane] * RHOR, -1.0 / 3.0);
    if (fixneg) fix_negative(L, X, sh, kP());
#pragma unroll
    for (int s = 0; s < S; ++s) {
      sh.cell[CL_M1][64 * s + lane] = 0.0;
      sh.cell[CL_M1SOL][64 * s + lane] = 0.0;
      sh.cell[CL_DZ1][64 * s + lane] = 1.0;
    }
    // the vertical velocity, carried in registers (do_sedi_w)
    double wv[S];
#pragma unroll
    for (int s = 0; s < S; ++s) wv[s] = (sedi_w && L.v[s]) ? f.w[Gl.off[s]] : 0.0;
    __syncthreads();
    // ------------------------------------------------ sub-steps, recorded
    for (int it = 0; it < P.ntimes; ++it) {
      warm_rain_rec(L, X, sh, Gl, f, kP(), kc(), true, T, it, RW1_B, wv, sedi_w);
      sed_rec(L, X, sh, Gl, f, kP(), kc(), T, it, wv, sedi_w);
      warm_rain_rec(L, X, sh, Gl, f, kP(), kc(), false, T, it, RW2_B, wv, sedi_w);
      tape_save(T, it, RI, X, sh);
      // the last sub-step's icloud runs in its reverse (icloud_adj)
      if (it + 1 < P.ntimes) icloud(L, X, sh, Gl, f, kP(), kc());
    }
    __syncthreads();
    // ----------------------------------------------------- reverse walk
    AdjRegs<S> G;
#pragma unroll
    for (int p = 0; p < AP_N; ++p) G.gp[p] = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      G.wb[s] = (sedi_w && a.w_after && L.v[s]) ? a.w_after[Gl.off[s]] : 0.0;
#pragma unroll
      for (int q = 0; q < 8; ++q) G.xb.v[q][s] = 0.0;
#pragma unroll
      for (int b = 0; b < AB_N; ++b) abar(A, b, s) = 0.0;
    }
    const double convt = 86400.0 * rdt * RGRAV;
    if (lane < ncols) {
      A.colbar[CB_RAIN][lane] = a.rain ? a.rain[o2] * convt : 0.0;
      A.colbar[CB_SNOW][lane] = a.snow ? a.snow[o2] * convt : 0.0;
      A.colbar[CB_ICE][lane] = a.ice ? a.ice[o2] * convt : 0.0;
      A.colbar[CB_GRAUPEL][lane] = a.graupel ? a.graupel[o2] * convt : 0.0;
      A.colbar[CB_CCN][lane] = A.colbar[CB_CPRAUT][lane] = 0.0;
    }
    InBar<S> nb;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      nb.pt[s] = nb.dp0[s] = nb.ua[s] = nb.va[s] = nb.qa[s] = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q) nb.q[q][s] = 0.0;
    }
    // fields_update's reverse runs in the last sub-step's icloud reverse,
    // except the momentum transport's
    if (n.sedi_transport) {
      double m1[S], mp[S], al[S], bu[S], bv[S], u1[S], v1[S], an[S], gu[S], gv[S], nu[S], nv[S];
#pragma unroll
      for (int s = 0; s < S; ++s) m1[s] = sh.cell[CL_M1][64 * s + lane];
      col_prev(sh.buf, m1, mp);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int32_t o = Gl.off[s];
        const double u0 = L.v[s] ? f.ua[o] : 0.0, v0 = L.v[s] ? f.va[o] : 0.0;
        if (L.k[s] == 0) {
          al[s] = 0.0;
          bu[s] = u0;
          bv[s] = v0;
        } else {
          const double dp0 = L.v[s] ? f.delp[o] : 100.0;
          const double r = 1.0 / (dp0 + mp[s]);
          al[s] = mp[s] * r;
          bu[s] = dp0 * u0 * r;
          bv[s] = dp0 * v0 * r;
        }
        gu[s] = (L.v[s] && L.k[s] > 0) ? a.udt[o] * rdt : 0.0;
        gv[s] = (L.v[s] && L.k[s] > 0) ? a.vdt[o] * rdt : 0.0;
        nb.ua[s] -= gu[s];
        nb.va[s] -= gv[s];
      }
      scan_fwd2(L, al, bu, bv, u1, v1);
      col_next(sh.buf, al, an);
#pragma unroll
      for (int s = 0; s < S; ++s)
        if (L.kb[s] == 0) an[s] = 0.0;
      scan_bwd2(L, an, gu, gv, nu, nv);
      double u1p[S], v1p[S], mb[S], mbn[S];
      col_prev(sh.buf, u1, u1p);
      col_prev(sh.buf, v1, v1p);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        mb[s] = 0.0;
        if (!L.v[s]) continue;
        const int32_t o = Gl.off[s];
        if (L.k[s] == 0) {
          nb.ua[s] += nu[s];
          nb.va[s] += nv[s];
        } else {
          const double dp0 = f.delp[o], u0 = f.ua[o], v0 = f.va[o];
          const double r = 1.0 / (dp0 + mp[s]);
          nb.ua[s] += nu[s] * dp0 * r;
          nb.va[s] += nv[s] * dp0 * r;
          nb.dp0[s] += (nu[s] * (u0 - u1[s]) + nv[s] * (v0 - v1[s])) * r;
          mb[s] = (nu[s] * (u1p[s] - u1[s]) + nv[s] * (v1p[s] - v1[s])) * r;  // lands on level k-1
        }
      }
      col_next(sh.buf, mb, mbn);
#pragma unroll
      for (int s = 0; s < S; ++s)
        if (L.v[s] && L.kb[s] > 0) abar(A, AB_M1, s) += mbn[s];
    }
    __syncthreads();
    // ------------------------------------------------ sub-steps, reversed
    for (int it = P.ntimes - 1; it >= 0; --it) {
      aux_load(T, it, sh);
      icloud_adj(L, A, G, Gl, f, a, kP(), kc(), T, it, it == P.ntimes - 1, nb, EXT && P.nml.z_slope_ice);
      __syncthreads();
      warm_rain_adj(L, A, G, Gl, f, kP(), kc(), false, T, it, RW2_B, X.dp1, sedi_w);
      __syncthreads();
      sed_adj(L, A, G, kP(), kc(), T, it, X.dp1, sedi_w);
      __syncthreads();
      warm_rain_adj(L, A, G, Gl, f, kP(), kc(), true, T, it, RW1_B, X.dp1, sedi_w);
      __syncthreads();
    }
    if (fixneg) fix_negative_adj(L, A, G, X, Gl, f, kP());
    // -------------------------------------------------- fields_init, reversed
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (!L.v[s]) continue;
      const int cl = L.cl[s];
      const double ccn = sh.colv[CV_CCN][cl], cpraut = sh.colv[CV_CPRAUT][cl];
      const double ccnbar = A.colbar[CB_CCN][cl] - A.colbar[CB_CPRAUT][cl] * cpraut * rcp64(3.0 * ccn);
      const int32_t o = Gl.off[s];
      const double dp0 = f.delp[o], qv = f.qv[o], dp1 = X.dp1[s];
      const double omq = dp0 / dp1;
      const double qin[6] = {qv, f.ql[o], f.qr[o], f.qi[o], f.qs[o], f.qg[o]};
      double omqb = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        nb.q[q][s] += G.xb.v[q][s] * omq;
        omqb += G.xb.v[q][s] * qin[q];
      }
      nb.pt[s] += G.xb.v[6][s] + abar(A, AB_T0, s);
      double dp1b = abar(A, AB_DP1, s) - omqb * omq * rcp64(dp1);
      nb.dp0[s] += omqb * rcp64(dp1);
      double dz0b = abar(A, AB_DZ0, s);
      if (L.kb[s] == 0 && P.use_ccn) {
        // ccn = ccn0 rdgas tz / p1, p1 = -dp1 / (g dz0) rdgas t0, with tz = t0 = pt
        nb.pt[s] += ccnbar * ccn * rcp64(f.pt[o]);
        const double p1b = -ccnbar * ccn * rcp64(p1_b[s]);
        dp1b += p1b * p1_b[s] * rcp64(dp1);
        dz0b -= p1b * p1_b[s] * rcp64(f.delz[o]);
        nb.pt[s] += p1b * p1_b[s] * rcp64(f.pt[o]);
      }
      nb.dp0[s] += dp1b * (1.0 - qv);
      nb.q[0][s] -= dp1b * dp0;
      a.pt[o] = nb.pt[s];
      a.qv[o] = nb.q[0][s];
      a.ql[o] = nb.q[1][s];
      a.qr[o] = nb.q[2][s];
      a.qi[o] = nb.q[3][s];
      a.qs[o] = nb.q[4][s];
      a.qg[o] = nb.q[5][s];
      a.qa[o] = nb.qa[s];
      a.ua[o] = nb.ua[s];
      a.va[o] = nb.va[s];
      a.delp[o] = nb.dp0[s];
      a.delz[o] = dz0b;
      if (a.qaz) a.qaz[o] = n.reset_qaz ? 0.0 : G.xb.v[7][s];
      if (sedi_w && a.w) a.w[o] = G.wb[s];
    }
#pragma unroll
    for (int p = 0; p < AP_N; ++p) gp_acc[p] += G.gp[p];
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < AP_N; ++p) {
    const double v = wave_sum(gp_acc[p]);
    if (lane == 0) part[(int64_t)blockIdx.x * AP_N + p] = v;
  }
}

// One launch of the adjoint kernel (instantiated with EXT = false in
// mp_adjoint.hip and EXT = true in mp_adjoint_ext.hip, which compile in
// parallel).
template <int S, bool EXT>
hipError_t launch_adjoint_kernel(unsigned nw, const AdjArgs& args, hipStream_t stream) {
  hipLaunchKernelGGL((mp_adjoint_kernel<S, EXT>), dim3(nw), dim3(64), 0, stream, args);
  return hipGetLastError();
}

}  // namespace shield
