/*
 * shield_mp.h -- C ABI of the MI355X GFDL cloud-microphysics column physics.
 *
 * Every entry point takes plain pointers and sizes.  Device pointers are
 * caller-owned FP64 buffers; launches are ordered on the HIP stream passed as
 * `stream` (a hipStream_t, NULL = default stream).  No entry point allocates
 * device memory.  Errors are returned as negative codes; shield_strerror()
 * gives the text.
 *
 * Field layout: a 3D field is addressed as base[i*si + j*sj + k] for
 * 0 <= i < ni, 0 <= j < nj, 0 <= k < nk (k stride 1, k = 0 is the model top,
 * k = nk-1 the surface layer), which is the C-ordered [X, Y, Z] storage of the
 * reference's Quantity (halo offsets folded into the base pointer).  A 2D
 * field is base[i*si2 + j*sj2].
 */
#ifndef SHIELD_MP_H
#define SHIELD_MP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
enum shield_status {
  SHIELD_OK = 0,
  SHIELD_ERR_NULL = -1,        /* a required pointer is NULL                      */
  SHIELD_ERR_SHAPE = -2,       /* ni, nj, nk or strides out of range              */
  SHIELD_ERR_UNSUPPORTED = -3, /* use_ppm or prog_ccn requested (broken upstream)  */
  SHIELD_ERR_MODE = -4,        /* unknown precision mode                          */
  SHIELD_ERR_ALIAS = -5,       /* an output aliases another field                 */
  SHIELD_ERR_LAUNCH = -6,      /* HIP launch or runtime failure                   */
  SHIELD_ERR_TIMESTEP = -7,    /* timestep <= 0 or not finite                     */
  SHIELD_ERR_NOGPU = -8,       /* no HIP device visible                           */
  SHIELD_ERR_ADJ_SWITCH = -9,  /* a switch the adjoint does not differentiate     */
  SHIELD_ERR_WORKSPACE = -10   /* workspace smaller than one wavefront's tape     */
};

/* Precision modes of shield_mp_forward. */
enum shield_mp_mode {
  SHIELD_MP_FP64 = 0,  /* every operation in FP64 (the reference's arithmetic)   */
  SHIELD_MP_MIXED = 1  /* FP32 process rates; FP64 state, transfers, column
                          sedimentation/flux sums and precipitation             */
};

/* ---------------------------------------------------------------- config */
/* Microphysics fields of PhysicsConfig (pySHiELD/_config.py:21-109).        */
typedef struct shield_mp_namelist {
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
  /* Extension, not in the reference: 1 = zero the cloud-fraction accumulator
   * qaz at the start of every call (GFDL Fortran behaviour).  0 = keep it
   * across calls exactly as pySHiELD/stencils/microphysics.py:1674,2156 does. */
  int32_t reset_qaz;
} shield_mp_namelist;

/* Derived scalars: Microphysics.__init__ (microphysics.py:1643-1659,1712-1713),
 * gfdl_cloud_microphys_init/setupm (:1806-1954) and _set_timestep
 * (:1964-1979).  Filled by shield_mp_setup(); passed by value to kernels.   */
typedef struct shield_mp_params {
  shield_mp_namelist nml;
  double c_air, c_vap, d0_vap, lv00, cpaut, so3, zs, log_10, tice0, t_wfr;
  double fac_rc, cracs, csacr, cgacr, cgacs;
  double acco[3][4];
  double csacw, csaci, cgacw, cgaci, cracw;
  double cssub[5], cgsub[5], crevp[5], cgfr[2], csmlt[5], cgmlt[5];
  double ces0;
  double timestep, mpdt, rdt, dts, dt_rain, rdts, dt_evap;
  double fac_i2s, fac_g2v, fac_v2g, fac_imlt, fac_l2v;
  int32_t ntimes;
  int32_t use_ccn;   /* = !prog_ccn                       (microphysics.py:1658) */
  int32_t do_sedi_w; /* = do_sedi_w && !hydrostatic       (microphysics.py:1655) */
  /* 1 (set by shield_mp_setup): tendencies are accumulated with += as
   * Microphysics.__call__ does.  0: they are overwritten, which equals
   * zeroing them first as Physics.__call__ does with prepare_microphysics
   * (physics.py:133-143) and saves reading them.                           */
  int32_t accumulate;
} shield_mp_params;

/* ----------------------------------------------------------------- shape */
typedef struct shield_mp_shape {
  int64_t ni, nj, nk; /* compute domain: ni*nj columns of nk levels          */
  int64_t si, sj;     /* element strides of i and j for 3D fields (k stride 1) */
  int64_t si2, sj2;   /* element strides of i and j for 2D fields            */
} shield_mp_shape;

/* ---------------------------------------------------------------- fields */
/* MicrophysicsState (microphysics.py:1556-1620) plus the driver's persistent
 * buffers.  Names follow the reference; the reference passes delprsi as
 * `delp`, delz as `dz` and wmp as `w` (microphysics.py:2003-2004,2058).     */
typedef struct shield_mp_fields {
  /* state in (3D); qi and qs are rewritten in place when de_ice, as the
   * reference's fields_init does (microphysics.py:110-120)                  */
  const double *pt, *qv, *ql, *qr;
  double *qi, *qs;
  const double *qg, *qa, *ua, *va, *delp, *delz;
  double *w;                  /* vertical velocity; in/out when do_sedi_w     */
  const double *land, *area;  /* 2D land mask and cell area                   */
  /* tendencies, accumulated with += (microphysics.py:1510-1538);
   * qi_dt and qs_dt are also read first when de_ice (:110-120)              */
  double *pt_dt, *qv_dt, *ql_dt, *qr_dt, *qi_dt, *qs_dt, *qg_dt, *qa_dt;
  double *udt, *vdt;
  double *qaz;                /* 3D cloud-fraction accumulator, in/out        */
  /* 2D surface precipitation of this call in mm/day (the reference keeps it
   * in private 3D accumulators, microphysics.py:1547-1553,1664-1667); any of
   * the four may be NULL                                                    */
  double *rain, *snow, *ice, *graupel;
} shield_mp_fields;

/* --------------------------------------------------------- entry points */

/* Fill *out from the namelist for the given timestep: the host-side part of
 * Microphysics.__init__ (microphysics.py:1624-1811).  Pure host code.       */
int shield_mp_setup(const shield_mp_namelist* nml, double timestep,
                    shield_mp_params* out);

/* Recompute the timestep-dependent scalars in place when `timestep` differs
 * from p->timestep: Microphysics._update_timestep_if_needed
 * (microphysics.py:1959-1962).  Pure host code.                             */
int shield_mp_update_timestep(shield_mp_params* p, double timestep);

/* One Microphysics.__call__(state, timestep) (microphysics.py:1981-2263) on
 * the GPU: fields_init, ntimes x [warm_rain, sedimentation, warm_rain,
 * icloud], fields_update, all in one fused kernel.  `mode` is a
 * shield_mp_mode.                                                           */
int shield_mp_forward(const shield_mp_shape* shape, const shield_mp_params* p,
                      const shield_mp_fields* f, int mode, void* stream);

/* The state-update part of shield_mp_forward_update: the sequence
 * Physics.__call__ runs after the microphysics (physics.py:331-366) and
 * UpdateAtmosphereState.__call__ runs next (update_atmos_state.py:301-349,
 * without ApplyPhysicsToDycore).  3D fields use the microphysics shape's
 * si/sj; prsi has nk+1 levels with strides si_i/sj_i and k stride 1.        */
typedef struct shield_update_fields {
  /* physics_updated_* = x + x_dt * dt (physics.py:146-194), written, in the
   * order qvapor, qliquid, qrain, qice, qsnow, qgraupel, qcld, pt, ua, va  */
  double* updated[10];
  /* dycore tendencies, accumulated: += (x_updated - x) * rdt for ua, va, pt
   * (update_atmos_state.py:76-78)                                          */
  double *u_dt, *v_dt, *pt_dt;
  /* dycore (FV3) tracers qvapor, qliquid, qrain, qsnow, qice, qgraupel and
   * delp, rewritten in place by the GFS -> FV3 conversion (:79-92)         */
  double* tracers[6];
  double* delp;
  const double* prsi;
  int64_t si_i, sj_i;
  double dt;     /* forward Euler step (the physics timestep)               */
  double rdt;    /* 1 / dt_atmos of the dycore tendencies                   */
  double q_min;  /* fill_gfs_delp threshold on the updated vapour (1e-9)    */
} shield_update_fields;

/* shield_mp_forward, then in the same kernel
 * update_physics_state_with_tendencies (physics.py:146-194), fill_gfs_delp
 * on physics_updated_specific_humidity with the dycore delp
 * (update_atmos_state.py:21-37, :318-320) and
 * prepare_tendencies_and_update_tracers (:40-92, :321-345): the state the
 * dycore receives, without writing and re-reading the tendencies.  The
 * reference also fills the halo columns of the updated vapour; only the
 * columns in `shape` are filled here, the only ones the conversion reads.  No
 * single reference function: it replaces that sequence of four calls.     */
int shield_mp_forward_update(const shield_mp_shape* shape, const shield_mp_params* p,
                             const shield_mp_fields* f, const shield_update_fields* u, int mode,
                             void* stream);

/* ------------------------------------------------- state-update path */
/* For these entry points shape->si2 and shape->sj2 are the i/j strides of
 * the nk+1-level interface field (phii or prsi), which has k stride 1 like
 * every other field; the nk-level fields use shape->si and shape->sj.       */

/* prepare_microphysics (pySHiELD/stencils/physics.py:105-143):
 * dz = (phii[k+1]-phii[k])/g, wmp = -omga(1+zvir qv) pt/delp Rd/g, and the
 * ten tendencies zeroed.  phii has nk+1 levels (interfaces).  tend[10] = {udt, vdt, pt_dt, qv_dt, ql_dt,
 * qr_dt, qi_dt, qs_dt, qg_dt, qa_dt}.                                       */
int shield_prepare_microphysics(const shield_mp_shape* shape, double* dz,
                                const double* phii, double* wmp,
                                const double* omga, const double* qvapor,
                                const double* pt, const double* delp,
                                double* const tend[10], void* stream);

/* update_physics_state_with_tendencies (physics.py:146-194): for each of the
 * ten fields out[n] = in[n] + tend[n]*dt, in the order qvapor, qliquid,
 * qrain, qice, qsnow, qgraupel, qcld, pt, ua, va.                           */
int shield_update_physics_state_with_tendencies(const shield_mp_shape* shape,
                                                const double* const in[10],
                                                const double* const tend[10],
                                                double* const out[10], double dt,
                                                void* stream);

/* fill_gfs_delp (pySHiELD/update/update_atmos_state.py:21-37): vertical
 * filling of q below q_min, in place, over the nk layers of the reference's
 * nk+1-level domain; 1 <= nk <= SHIELD_FILL_MAX_NK.                         */
#define SHIELD_FILL_MAX_NK 2048
int shield_fill_gfs_delp(const shield_mp_shape* shape, const double* delp,
                         double* q, double q_min, void* stream);

/* prepare_tendencies_and_update_tracers (update_atmos_state.py:40-92) over
 * all nk layers (the reference's interval(0, -1) on an nk+1-level domain):
 * u_dt, v_dt, pt_dt += (x_t1 - x_t0) * rdt; the GFS mixing ratios *_t1 are
 * converted to FV3 ones in *_t0 and delp is replaced by the new total mass.
 * prsi has nk+1 interface levels.                                          */
int shield_prepare_tendencies_and_update_tracers(
    const shield_mp_shape* shape, double* u_dt, double* v_dt, double* pt_dt,
    const double* u_t1, const double* v_t1, const double* pt_t1,
    const double* qv_t1, const double* ql_t1, const double* qr_t1,
    const double* qs_t1, const double* qi_t1, const double* qg_t1,
    const double* u_t0, const double* v_t0, const double* pt_t0,
    double* qv_t0, double* ql_t0, double* qr_t0, double* qs_t0, double* qi_t0,
    double* qg_t0, const double* prsi, double* delp, double rdt, void* stream);

/* ------------------------------------------- Physics driver prologue */
/* Fields of shield_physics_prologue.  Layer fields have nk levels and use
 * shape->si/sj; phii, prsi and prsik have nk+1 interfaces with strides
 * si_i/sj_i.  The reference stores every field with nk+1 levels; the extra
 * level of its layer fields is padding that no result reads.               */
typedef struct shield_prologue_fields {
  const double* delz;
  /* in: total layer mass (Pa); out: the layer-mean pressure, which the
   * reference writes into the same field (physics.py:89-93)                */
  double* delp;
  /* qvapor, qliquid, qrain, qice, qsnow, qgraupel, qo3mr: in per total mass,
   * out per dry mass (physics.py:54-82)                                     */
  double* q[7];
  double* qsgs_tke;  /* divided by the dry layer mass (physics.py:82)       */
  const double* pt;
  double *phii, *prsi, *prsik;     /* out, nk+1 interfaces                  */
  double *dm, *delprsi, *del_gz;   /* out: dry layer mass, prsi[k+1]-prsi[k],
                                    * and del_gz as get_phi_fv3 leaves it   */
  double* phil;                    /* out: geopotential at layer middles     */
  /* prepare_microphysics (physics.py:105-143) when omga != NULL: dz, wmp and
   * the ten tendencies {udt, vdt, pt_dt, qv_dt, ql_dt, qr_dt, qi_dt, qs_dt,
   * qg_dt, qa_dt} zeroed                                                    */
  const double* omga;
  double *dz, *wmp;
  double* tend[10];
  int64_t si_i, sj_i;
  double ptop;  /* model-top pressure (grid_data.ptop)                      */
} shield_prologue_fields;

/* ---------------------------------------------------------- adjoint */
/* Cotangents for shield_mp_adjoint.  Output cotangents (read) belong to the
 * outputs of shield_mp_forward: the ten tendencies (3D, strides of the
 * shape) and the four surface precipitation rates (2D, any may be NULL for a
 * zero cotangent).  Input cotangents (written) belong to the twelve input
 * fields and the incoming cloud-fraction accumulator (qaz may be NULL).
 * params receives the gradient (device array of SHIELD_MP_NPARAMS doubles,
 * summed over the domain) of the six process parameters, in the order
 * fac_rc, vr_fac, vs_fac, vg_fac, vi_fac, qi_gen of shield_mp_params /
 * shield_mp_namelist.                                                        */
#define SHIELD_MP_NPARAMS 6
typedef struct shield_mp_adjoint_fields {
  const double *pt_dt, *qv_dt, *ql_dt, *qr_dt, *qi_dt, *qs_dt, *qg_dt, *qa_dt, *udt, *vdt;
  const double *rain, *snow, *ice, *graupel;
  double *pt, *qv, *ql, *qr, *qi, *qs, *qg, *qa, *ua, *va, *delp, *delz;
  double* qaz;
  double* params;
  /* do_sedi_w: the cotangent of w after the call (read; NULL reads as zero)
   * and of w before it (written; NULL leaves it unwritten).  Neither is used
   * without do_sedi_w.                                                       */
  const double* w_after;
  double* w;
} shield_mp_adjoint_fields;

/* Bytes of workspace for `waves` concurrent wavefronts (each walks groups of
 * columns and records their per-sub-step state there); 0 on a bad shape.    */
size_t shield_mp_adjoint_workspace_bytes(const shield_mp_shape* shape, const shield_mp_params* p, int64_t waves);

/* Reverse-mode adjoint of one FP64 shield_mp_forward call with the inputs
 * f (its output pointers are not used): input cotangents and the parameter
 * gradient from output cotangents, as for overwritten tendencies
 * (accumulate = 0).  The forward of each column is recomputed, so f's inputs
 * must be those of the forward call.  The workspace (ws, ws_bytes) bounds
 * the number of concurrent wavefronts; nk <= 128.  Returns
 * SHIELD_ERR_ADJ_SWITCH for de_ice with accumulate (de_ice changes nothing
 * on overwritten tendencies),
 * and SHIELD_ERR_ALIAS when a written cotangent, the parameter gradient or
 * the workspace shares memory with another of them or with a field read.
 * No reference counterpart: the reference has no derivative code; this is
 * the torch.autograd backward of the microphysics call.                     */
int shield_mp_adjoint(const shield_mp_shape* shape, const shield_mp_params* p, const shield_mp_fields* f,
                      const shield_mp_adjoint_fields* a, void* ws, size_t ws_bytes, void* stream);

/* shield_mp_adjoint without a workspace argument and without any device
 * allocation: the tapes and parameter partial sums go to a device array the
 * library holds (SHIELD_MP_ADJOINT_SCRATCH_BYTES per GPU, for the life of the
 * process), which bounds the number of concurrent wavefronts exactly as a
 * workspace of that size does.  Same inputs, outputs, checks and results as
 * shield_mp_adjoint.  Returns SHIELD_ERR_WORKSPACE when one wavefront's tape
 * (ntimes x 77 x 64 x slots doubles) does not fit the array.  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (two
 * streams, two threads) race and must be serialised by the caller.
 * No counterpart in the reference.  It is for callers that cannot allocate
 * device memory: code under HIP graph capture, serving loops and harnesses
 * that forbid hipMalloc.                                                      */
#define SHIELD_MP_ADJOINT_SCRATCH_BYTES ((size_t)128 << 20)
int shield_mp_adjoint_noalloc(const shield_mp_shape* shape, const shield_mp_params* p, const shield_mp_fields* f,
                              const shield_mp_adjoint_fields* a, void* stream);

/* The column-local part of Physics.__call__ before the microphysics
 * (physics.py:278-330): atmos_phys_driver_statein (physics.py:26-102),
 * get_prs_fv3 (get_prs_fv3.py:7-20), get_phi_fv3 (get_phi_fv3.py:7-25) and,
 * when f->omga is set, prepare_microphysics, in one kernel.  nwat = 6;
 * 1 <= nk <= SHIELD_PROLOGUE_MAX_NK.                                        */
#define SHIELD_PROLOGUE_MAX_NK 1024
int shield_physics_prologue(const shield_mp_shape* shape, const shield_prologue_fields* f, void* stream);

/* Fields of shield_apply_physics_to_dycore for one tile on one rank (all
 * four tile edges local), C-contiguous, k = 0 the top, i in [0, ni), j in
 * [0, nj):
 *   layer fields [ni][nj][nk]; interface fields pe, peln, pk [ni][nj][nk+1];
 *   2D fields [ni][nj];
 *   u [ni][nj+1][nk] and v [ni+1][nj][nk], the D-grid winds;
 *   u_dt, v_dt [ni+2][nj+2][nk], the A-grid wind tendencies with a one-point
 *   halo (compute point (i, j) at [i+1][j+1]), filled by the caller;
 *   vlon_*, vlat_* [ni+2][nj+2] with the same halo; es1_* [ni][nj+1];
 *   ew2_* [ni+1][nj]; edge_vect_w, _e [nj]; edge_vect_s, _n [ni].          */
typedef struct shield_dycore_fields {
  const double *qvapor, *qliquid, *qrain, *qsnow, *qice, *qgraupel;
  double *pt, *t_dt;
  const double* delp;
  double *pe, *peln, *pk, *ps;
  const double *ua, *va;
  double *u_srf, *v_srf;
  double *u, *v, *u_dt, *v_dt;
  const double *vlon1, *vlon2, *vlon3, *vlat1, *vlat2, *vlat3;
  const double *es1_1, *es1_2, *es1_3, *ew2_1, *ew2_2, *ew2_3;
  const double *edge_vect_w, *edge_vect_e, *edge_vect_s, *edge_vect_n;
  double dt;      /* the physics time step (dt_atmos)                       */
  double con_cp;  /* heat capacity of moist_cv (CP_AIR in the reference)     */
  double kappa;   /* RDGAS / CP_AIR                                         */
  int32_t grid_type, npx, npy, pad_;
} shield_dycore_fields;

/* The steps of ApplyPhysicsToDycore.__call__ (fv_update_phys.py:148-195)
 * that pySHiELD holds, for one tile on one rank: moist_cv (:32-50),
 * update_pressure_and_surface_winds (:53-74) and AGrid2DGridPhysics
 * (update_dwind_phys.py:166-701; the doubly periodic form when
 * grid_type > 3).  The u_dt / v_dt halo exchange before the wind update and
 * CubedToLatLon after it (NDSL) are the caller's.  Only shape->ni, nj, nk
 * are read.  The grid fields may be NULL when grid_type > 3.  Errors:
 * SHIELD_ERR_NULL; SHIELD_ERR_SHAPE (ni or nj below 2, nk outside
 * [1, SHIELD_DYCORE_MAX_NK], or
 * npx - 1 != ni or npy - 1 != nj when grid_type <= 3); SHIELD_ERR_ALIAS (a
 * written field equal to another field); SHIELD_ERR_TIMESTEP.               */
#define SHIELD_DYCORE_MAX_NK 4096
int shield_apply_physics_to_dycore(const shield_mp_shape* shape, const shield_dycore_fields* f, void* stream);

/* ------------------------------------------------------------ utilities */
const char* shield_strerror(int code);
/* Number of HIP devices visible (0 without a GPU; never fails).            */
int shield_device_count(void);
/* Columns per wavefront and slots per lane the forward kernel picks for nk
 * levels in `mode` with `ntimes` sub-steps on the current device (the
 * busiest-lanes layout when no device is visible).  No reference
 * counterpart: a diagnostic of this implementation.                         */
int shield_mp_layout(int64_t nk, int mode, int ntimes, int* cols_per_wave, int* slots);

#ifdef __cplusplus
}
#endif
#endif /* SHIELD_MP_H */
