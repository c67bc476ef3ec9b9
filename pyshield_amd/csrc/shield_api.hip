// shield_api.hip -- C ABI entry points (include/shield_mp.h): host-side
// set-up of the microphysics scalars, argument validation and launches.
#include <math.h>
#include <string.h>

#include "mp_common.h"

namespace shield {
hipError_t launch_forward(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f,
                          const shield_update_fields& u, int mode, hipStream_t stream);
struct LayoutChoice {
  int S, P;
};
LayoutChoice choose_layout(int64_t nz, int mode, int ntimes, bool fused);
int64_t adjoint_tape_doubles(int64_t nk, int ntimes, bool sedi_w);
int adjoint_cols_per_wave(int64_t nk);
hipError_t launch_adjoint_noalloc(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f,
                                  const shield_mp_adjoint_fields& a, hipStream_t stream);
hipError_t launch_adjoint(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f,
                          const shield_mp_adjoint_fields& a, void* ws, size_t ws_bytes, hipStream_t stream);
}  // namespace shield

using namespace shield;

namespace {

// Microphysics._set_timestep (microphysics.py:1964-1979).  Python's round()
// rounds half to even, as nearbyint does in the default rounding mode.
void set_timestep(shield_mp_params* p, double timestep) {
  const shield_mp_namelist& n = p->nml;
  p->timestep = timestep;
  p->mpdt = timestep < n.mp_time ? timestep : n.mp_time;
  p->rdt = 1.0 / timestep;
  p->ntimes = (int32_t)nearbyint(timestep / p->mpdt);
  p->dts = timestep / p->ntimes;
  p->dt_rain = p->dts * 0.5;
  p->rdts = 1.0 / p->dts;
  p->dt_evap = n.fast_sat_adj ? 0.5 * p->dts : p->dts;
  p->fac_i2s = 1.0 - exp(-p->dts / n.tau_i2s);
  p->fac_g2v = 1.0 - exp(-p->dts / n.tau_g2v);
  p->fac_v2g = 1.0 - exp(-p->dts / n.tau_v2g);
  p->fac_imlt = 1.0 - exp(-0.5 * p->dts / n.tau_imlt);
  p->fac_l2v = 1.0 - exp(-p->dt_evap / n.tau_l2v);
}

// Microphysics.__init__ scalars (microphysics.py:1643-1659,1712-1713) and
// gfdl_cloud_microphys_init / setupm (:1806-1954): collection, evaporation,
// sublimation and melting coefficients of Lin et al. (1983).
void setupm(shield_mp_params* p) {
  const shield_mp_namelist& n = p->nml;
  p->c_air = CP_AIR;
  p->c_vap = CP_VAP;
  p->d0_vap = p->c_vap - C_LIQ;
  p->lv00 = HLV - p->d0_vap * TICE;
  p->do_sedi_w = n.hydrostatic ? 0 : n.do_sedi_w;
  p->cpaut = n.c_paut * 0.104 * GRAV / 1.717e-5;
  p->use_ccn = n.prog_ccn ? 0 : 1;
  p->so3 = 7.0 / 3.0;
  p->zs = 0.0;
  p->log_10 = log(10.0);
  p->tice0 = n.tice - 0.01;
  p->t_wfr = n.tice - 40.0;

  // gamma-function values and intercept / density parameters
  const double gam263 = 1.456943, gam275 = 1.608355, gam290 = 1.827363;
  const double gam325 = 2.54925, gam350 = 3.323363, gam380 = 4.694155;
  const double rnzs = 3.0e6, rnzr = 8.0e6, rnzg = 4.0e6;
  const double rhos = 1.0e2, rhog = 4.0e2, rhor = RHOR;
  const double acc[3] = {5.0, 2.0, 0.5};
  const double pie = 4.0 * atan(1.0);
  p->fac_rc = (4.0 / 3.0) * pie * rhor * pow(n.rthresh, 3.0);
  const double vdifu = 2.11e-5, tcond = 2.36e-2, visk = 1.259e-5;
  const double hlts = 2.8336e6, hltc = 2.5e6, hltf = 3.336e5, ch2o = 4.1855e3;
  const double pisq = pie * pie;
  const double scm3 = pow(visk / vdifu, 1.0 / 3.0);
  p->cracs = pisq * rnzr * rnzs * rhos;
  p->csacr = pisq * rnzr * rnzs * rhor;
  p->cgacr = pisq * rnzr * rnzg * rhor;
  p->cgacs = pisq * rnzg * rnzs * rhos;
  p->cgacs = p->cgacs * n.c_pgacs;
  // act[2k], act[2k+1]: the (collector, collected) slope factors of the four
  // accretion pairs (snow-rain, rain-snow, graupel-rain, graupel-snow)
  const double act_s = pie * rnzs * rhos, act_r = pie * rnzr * rhor, act_g = pie * rnzg * rhog;
  const double act[8] = {act_s, act_r, act_r, act_s, act_r, act_g, act_s, act_g};
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 4; ++k)
      p->acco[i][k] = acc[i] / (pow(act[2 * k], (6 - i) * 0.25) * pow(act[2 * k + 1], (i + 1) * 0.25));
  const double gcon = 40.74 * sqrt(SFCRHO);
  p->csacw = pie * rnzs * n.clin * gam325 / (4.0 * pow(act_s, 0.8125));
  const double craci = pie * rnzr * n.alin * gam380 / (4.0 * pow(act_r, 0.95));
  p->csaci = p->csacw * n.c_psaci;
  p->cgacw = pie * rnzg * gam350 * gcon / (4.0 * pow(act_g, 0.875));
  p->cgaci = p->cgacw * 0.05;
  p->cracw = n.c_cracw * craci;
  p->cssub[0] = 2.0 * pie * vdifu * tcond * RVGAS * rnzs;
  p->cssub[1] = 0.78 / sqrt(act_s);
  p->cssub[2] = 0.31 * scm3 * gam263 * sqrt(n.clin / visk) / pow(act_s, 0.65625);
  p->cssub[3] = tcond * RVGAS;
  p->cssub[4] = (hlts * hlts) * vdifu;
  p->cgsub[0] = 2.0 * pie * vdifu * tcond * RVGAS * rnzg;
  p->cgsub[1] = 0.78 / sqrt(act_g);
  p->cgsub[2] = 0.31 * scm3 * gam275 * sqrt(gcon / visk) / pow(act_g, 0.6875);
  p->cgsub[3] = p->cssub[3];
  p->cgsub[4] = p->cssub[4];
  p->crevp[0] = 2.0 * pie * vdifu * tcond * RVGAS * rnzr;
  p->crevp[1] = 0.78 / sqrt(act_r);
  p->crevp[2] = 0.31 * scm3 * gam290 * sqrt(n.alin / visk) / pow(act_r, 0.725);
  p->crevp[3] = p->cssub[3];
  p->crevp[4] = hltc * hltc * vdifu;
  p->cgfr[0] = 20.0e2 * pisq * rnzr * rhor / pow(act_r, 1.75);
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
  p->ces0 = (RDGAS / RVGAS) * es0;
}

bool shape_ok(const shield_mp_shape* s, int64_t nk_min, int64_t nk_max) {
  if (!s) return false;
  if (s->ni < 0 || s->nj < 0) return false;
  if (s->nk < nk_min || s->nk > nk_max) return false;
  if (s->ni * s->nj > 0 && (s->si < 0 || s->sj < 0 || s->si2 < 0 || s->sj2 < 0)) return false;
  return true;
}

}  // namespace

extern "C" {

int shield_mp_setup(const shield_mp_namelist* nml, double timestep, shield_mp_params* out) {
  if (!nml || !out) return SHIELD_ERR_NULL;
  if (!(timestep > 0.0) || !isfinite(timestep)) return SHIELD_ERR_TIMESTEP;
  memset(out, 0, sizeof(*out));
  out->nml = *nml;
  setupm(out);
  set_timestep(out, timestep);
  out->accumulate = 1;
  if (nml->use_ppm || nml->prog_ccn) return SHIELD_ERR_UNSUPPORTED;
  return SHIELD_OK;
}

int shield_mp_update_timestep(shield_mp_params* p, double timestep) {
  if (!p) return SHIELD_ERR_NULL;
  if (!(timestep > 0.0) || !isfinite(timestep)) return SHIELD_ERR_TIMESTEP;
  if (timestep != p->timestep) set_timestep(p, timestep);
  return SHIELD_OK;
}

int shield_mp_layout(int64_t nk, int mode, int ntimes, int* cols_per_wave, int* slots) {
  if (mode != SHIELD_MP_FP64 && mode != SHIELD_MP_MIXED) return SHIELD_ERR_MODE;
  const LayoutChoice lc = choose_layout(nk, mode, ntimes, false);
  if (cols_per_wave) *cols_per_wave = lc.P;
  if (slots) *slots = lc.S;
  return lc.S == 0 ? SHIELD_ERR_SHAPE : SHIELD_OK;
}

int shield_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* shield_strerror(int code) {
  switch (code) {
    case SHIELD_OK: return "success";
    case SHIELD_ERR_NULL: return "a required pointer is NULL";
    case SHIELD_ERR_SHAPE: return "shape or strides out of range (1 <= nk <= 320 for the microphysics kernel)";
    case SHIELD_ERR_UNSUPPORTED: return "use_ppm and prog_ccn are not supported (both are broken in the reference)";
    case SHIELD_ERR_MODE: return "unknown precision mode";
    case SHIELD_ERR_ALIAS: return "an output buffer aliases another field";
    case SHIELD_ERR_LAUNCH: return "HIP launch failed";
    case SHIELD_ERR_TIMESTEP: return "timestep must be positive and finite";
    case SHIELD_ERR_NOGPU: return "no HIP device is visible";
    case SHIELD_ERR_ADJ_SWITCH:
      return "the adjoint does not differentiate do_sedi_w, or de_ice on accumulated tendencies";
    case SHIELD_ERR_WORKSPACE: return "adjoint workspace smaller than one wavefront's tape";
    default: return "unknown error";
  }
}

// Checks of the switches, shape and input fields shared by the forward and
// adjoint entry points (shape, p and f non-NULL); *empty = 1 when the domain
// is empty (nothing to launch).
static int check_inputs(const shield_mp_shape* shape, const shield_mp_params* p, const shield_mp_fields* f,
                        int* empty) {
  *empty = 0;
  if (p->nml.use_ppm || p->nml.prog_ccn) return SHIELD_ERR_UNSUPPORTED;
  if (!(p->timestep > 0.0) || p->ntimes < 1) return SHIELD_ERR_TIMESTEP;
  if (!shape_ok(shape, 2, 320)) return SHIELD_ERR_SHAPE;
  if (shape->ni * shape->nj == 0) {
    *empty = 1;
    return SHIELD_OK;
  }
  // the kernel addresses cells with 32-bit element offsets
  const int64_t last3 = (shape->ni - 1) * shape->si + (shape->nj - 1) * shape->sj + (shape->nk - 1);
  const int64_t last2 = (shape->ni - 1) * shape->si2 + (shape->nj - 1) * shape->sj2;
  if (last3 > INT32_MAX || last2 > INT32_MAX) return SHIELD_ERR_SHAPE;
  const void* in[] = {f->pt, f->qv, f->ql, f->qr, f->qi, f->qs, f->qg, f->qa, f->ua, f->va, f->delp, f->delz,
                      f->land, f->area};
  for (const void* q : in)
    if (!q) return SHIELD_ERR_NULL;
  return SHIELD_OK;
}

// Checks shared by shield_mp_forward and shield_mp_forward_update; 1 when
// the domain is empty (nothing to launch).
static int check_forward(const shield_mp_shape* shape, const shield_mp_params* p, const shield_mp_fields* f,
                         int mode, int* empty) {
  *empty = 0;
  if (!shape || !p || !f) return SHIELD_ERR_NULL;
  if (mode != SHIELD_MP_FP64 && mode != SHIELD_MP_MIXED) return SHIELD_ERR_MODE;
  const int rc = check_inputs(shape, p, f, empty);
  if (rc != SHIELD_OK || *empty) return rc;
  const void* in[] = {f->pt, f->qv, f->ql, f->qr, f->qi, f->qs, f->qg, f->qa, f->ua, f->va, f->delp, f->delz,
                      f->land, f->area};
  const void* out[] = {f->pt_dt, f->qv_dt, f->ql_dt, f->qr_dt, f->qi_dt, f->qs_dt,
                       f->qg_dt, f->qa_dt, f->udt,   f->vdt,   f->qaz};
  for (const void* q : out)
    if (!q) return SHIELD_ERR_NULL;
  if (p->do_sedi_w && !f->w) return SHIELD_ERR_NULL;
  // outputs must be distinct from each other and from every input
  const void* all_out[] = {f->pt_dt, f->qv_dt, f->ql_dt, f->qr_dt, f->qi_dt, f->qs_dt, f->qg_dt, f->qa_dt,
                           f->udt,   f->vdt,   f->qaz,   f->rain,  f->snow,  f->ice,   f->graupel};
  const int nout = (int)(sizeof(all_out) / sizeof(all_out[0]));
  for (int a = 0; a < nout; ++a) {
    if (!all_out[a]) continue;
    for (int b = a + 1; b < nout; ++b)
      if (all_out[a] == all_out[b]) return SHIELD_ERR_ALIAS;
    for (const void* q : in)
      if (all_out[a] == q) return SHIELD_ERR_ALIAS;
    if (f->w && all_out[a] == f->w) return SHIELD_ERR_ALIAS;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SHIELD_ERR_NOGPU;
  return SHIELD_OK;
}

int shield_mp_forward(const shield_mp_shape* shape, const shield_mp_params* p, const shield_mp_fields* f, int mode,
                      void* stream) {
  int empty = 0;
  const int rc = check_forward(shape, p, f, mode, &empty);
  if (rc != SHIELD_OK || empty) return rc;
  shield_update_fields none;
  memset(&none, 0, sizeof(none));
  const hipError_t e = launch_forward(*shape, *p, *f, none, mode, (hipStream_t)stream);
  return e == hipSuccess ? SHIELD_OK : SHIELD_ERR_LAUNCH;
}

size_t shield_mp_adjoint_workspace_bytes(const shield_mp_shape* shape, const shield_mp_params* p, int64_t waves) {
  if (!shape || !p || p->ntimes < 1 || waves < 1 || !shape_ok(shape, 2, 128)) return 0;
  const int64_t ncol = shape->ni * shape->nj, P = adjoint_cols_per_wave(shape->nk);
  const int64_t groups = (ncol + P - 1) / P;
  if (waves > groups) waves = groups < 1 ? 1 : groups;
  return (size_t)waves * (size_t)(adjoint_tape_doubles(shape->nk, p->ntimes, p->do_sedi_w != 0) + SHIELD_MP_NPARAMS) * sizeof(double);
}

// Checks of shield_mp_adjoint and shield_mp_adjoint_noalloc (ws may be NULL
// for the latter); *empty = 1 when there is nothing to launch.
static int check_adjoint(const shield_mp_shape* shape, const shield_mp_params* p, const shield_mp_fields* f,
                         const shield_mp_adjoint_fields* a, const void* ws, int* empty) {
  *empty = 0;
  if (!shape || !p || !f || !a) return SHIELD_ERR_NULL;
  const int rc = check_inputs(shape, p, f, empty);
  if (rc != SHIELD_OK) return rc;
  if (!f->qaz && !p->nml.reset_qaz) return SHIELD_ERR_NULL;
  if (shape->nk > 128) return SHIELD_ERR_SHAPE;
  if (p->nml.de_ice && p->accumulate) return SHIELD_ERR_ADJ_SWITCH;
  if (*empty) return SHIELD_OK;
  const void* need[] = {a->pt_dt, a->qv_dt, a->ql_dt, a->qr_dt, a->qi_dt, a->qs_dt, a->qg_dt, a->qa_dt, a->udt,
                        a->vdt,   a->pt,    a->qv,    a->ql,    a->qr,    a->qi,    a->qs,    a->qg,    a->qa,
                        a->ua,    a->va,    a->delp,  a->delz,  a->params};
  for (const void* q : need)
    if (!q) return SHIELD_ERR_NULL;
  // the written cotangents, the parameter gradient and the workspace are
  // distinct from each other and from everything the kernel reads
  const bool sw = p->do_sedi_w != 0;
  if (sw && !f->w) return SHIELD_ERR_NULL;
  const void* wr[] = {a->pt, a->qv, a->ql, a->qr, a->qi, a->qs, a->qg, a->qa, a->ua, a->va, a->delp, a->delz,
                      a->qaz, a->params, ws, sw ? a->w : nullptr};
  const void* rd[] = {f->pt,    f->qv,    f->ql,    f->qr,    f->qi,    f->qs,    f->qg,    f->qa,
                      f->ua,    f->va,    f->delp,  f->delz,  f->land,  f->area,  f->qaz,   a->pt_dt,
                      a->qv_dt, a->ql_dt, a->qr_dt, a->qi_dt, a->qs_dt, a->qg_dt, a->qa_dt, a->udt,
                      a->vdt,   a->rain,  a->snow,  a->ice,   a->graupel, sw ? f->w : nullptr,
                      sw ? a->w_after : nullptr};
  const int nwr = (int)(sizeof(wr) / sizeof(wr[0]));
  for (int i = 0; i < nwr; ++i) {
    if (!wr[i]) continue;
    for (int j = i + 1; j < nwr; ++j)
      if (wr[i] == wr[j]) return SHIELD_ERR_ALIAS;
    for (const void* q : rd)
      if (q && wr[i] == q) return SHIELD_ERR_ALIAS;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SHIELD_ERR_NOGPU;
  return SHIELD_OK;
}

int shield_mp_adjoint(const shield_mp_shape* shape, const shield_mp_params* p, const shield_mp_fields* f,
                      const shield_mp_adjoint_fields* a, void* ws, size_t ws_bytes, void* stream) {
  if (!ws) return SHIELD_ERR_NULL;
  int empty = 0;
  const int rc = check_adjoint(shape, p, f, a, ws, &empty);
  if (rc != SHIELD_OK || empty) return rc;
  if (ws_bytes < shield_mp_adjoint_workspace_bytes(shape, p, 1)) return SHIELD_ERR_WORKSPACE;
  const hipError_t e = launch_adjoint(*shape, *p, *f, *a, ws, ws_bytes, (hipStream_t)stream);
  return e == hipSuccess ? SHIELD_OK : SHIELD_ERR_LAUNCH;
}

int shield_mp_adjoint_noalloc(const shield_mp_shape* shape, const shield_mp_params* p, const shield_mp_fields* f,
                              const shield_mp_adjoint_fields* a, void* stream) {
  int empty = 0;
  const int rc = check_adjoint(shape, p, f, a, nullptr, &empty);
  if (rc != SHIELD_OK || empty) return rc;
  if (shield_mp_adjoint_workspace_bytes(shape, p, 1) > SHIELD_MP_ADJOINT_SCRATCH_BYTES) return SHIELD_ERR_WORKSPACE;
  const hipError_t e = launch_adjoint_noalloc(*shape, *p, *f, *a, (hipStream_t)stream);
  return e == hipSuccess ? SHIELD_OK : SHIELD_ERR_LAUNCH;
}

int shield_mp_forward_update(const shield_mp_shape* shape, const shield_mp_params* p, const shield_mp_fields* f,
                             const shield_update_fields* u, int mode, void* stream) {
  if (!u) return SHIELD_ERR_NULL;
  for (int n = 0; n < 10; ++n)
    if (!u->updated[n]) return SHIELD_ERR_NULL;
  for (int n = 0; n < 6; ++n)
    if (!u->tracers[n]) return SHIELD_ERR_NULL;
  if (!u->u_dt || !u->v_dt || !u->pt_dt || !u->delp || !u->prsi) return SHIELD_ERR_NULL;
  if (!shape) return SHIELD_ERR_NULL;
  if (u->si_i < 0 || u->sj_i < 0) return SHIELD_ERR_SHAPE;
  if (shape->ni * shape->nj > 0 &&
      (shape->ni - 1) * u->si_i + (shape->nj - 1) * u->sj_i + shape->nk > INT32_MAX)
    return SHIELD_ERR_SHAPE;
  if (!(u->dt > 0.0) || !isfinite(u->dt) || !isfinite(u->rdt)) return SHIELD_ERR_TIMESTEP;
  // the update's outputs are distinct from each other and from every
  // microphysics field
  const void* uo[] = {u->updated[0], u->updated[1], u->updated[2], u->updated[3], u->updated[4],
                      u->updated[5], u->updated[6], u->updated[7], u->updated[8], u->updated[9],
                      u->u_dt,       u->v_dt,       u->pt_dt,      u->tracers[0], u->tracers[1],
                      u->tracers[2], u->tracers[3], u->tracers[4], u->tracers[5], u->delp};
  const int nuo = (int)(sizeof(uo) / sizeof(uo[0]));
  if (f) {
    const void* mp[] = {f->pt,    f->qv,    f->ql,    f->qr,    f->qi,    f->qs,    f->qg,    f->qa,
                        f->ua,    f->va,    f->delp,  f->delz,  f->w,     f->pt_dt, f->qv_dt, f->ql_dt,
                        f->qr_dt, f->qi_dt, f->qs_dt, f->qg_dt, f->qa_dt, f->udt,   f->vdt,   f->qaz};
    for (int a = 0; a < nuo; ++a) {
      for (int b = a + 1; b < nuo; ++b)
        if (uo[a] == uo[b]) return SHIELD_ERR_ALIAS;
      for (const void* q : mp)
        if (q && uo[a] == q) return SHIELD_ERR_ALIAS;
      if (uo[a] == u->prsi) return SHIELD_ERR_ALIAS;
    }
  }
  int empty = 0;
  const int rc = check_forward(shape, p, f, mode, &empty);
  if (rc != SHIELD_OK || empty) return rc;
  const hipError_t e = launch_forward(*shape, *p, *f, *u, mode, (hipStream_t)stream);
  return e == hipSuccess ? SHIELD_OK : SHIELD_ERR_LAUNCH;
}

}  // extern "C"
