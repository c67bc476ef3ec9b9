// prologue.hip -- the column-local part of Physics.__call__ before the
// microphysics (pySHiELD/stencils/physics.py:278-330): atmos_phys_driver_statein,
// get_prs_fv3, get_phi_fv3 and prepare_microphysics in one kernel.
//
// The three stencils are vertical recurrences (geopotential summed from the
// surface, interface pressure summed from the top, geopotential rebuilt from
// del_gz); one thread walks one column, keeping the running sums in
// registers.  The point-wise conversions ride along in the same k-loops.
#include <math.h>

#include "mp_common.h"

using namespace shield;

namespace {

constexpr int kBlock = 64;
constexpr double KAPPA = RDGAS / CP_AIR;
constexpr double P00 = 1.0e5;
constexpr double ZVIR = RVGAS / RDGAS - 1.0;

struct Geo {
  int64_t ni, nj, nk, si, sj, si_i, sj_i;
};

__global__ __launch_bounds__(kBlock) void prologue_kernel(Geo g, shield_prologue_fields f, double pktop,
                                                          double pk0inv) {
  const int64_t col = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (col >= g.ni * g.nj) return;
  const int64_t i = col / g.nj, j = col - i * g.nj;
  const int64_t b = i * g.si + j * g.sj, bi = i * g.si_i + j * g.sj_i;
  const int nk = (int)g.nk;
  double* phii = f.phii + bi;
  double* prsi = f.prsi + bi;
  double* prsik = f.prsik + bi;

  // atmos_phys_driver_statein (physics.py:46-102)
  double z = 0.0;  // geopotential at the surface interface
  phii[nk] = z;
  for (int k = nk - 1; k >= 0; --k) {
    z = z - f.delz[b + k] * GRAV;
    phii[k] = z;
  }
  double p = f.ptop;  // interface pressure above layer k
  prsi[0] = p;
  for (int k = 0; k < nk; ++k) {
    const int64_t o = b + k;
    const double dp = f.delp[o];
    double m[7];
#pragma unroll
    for (int n = 0; n < 7; ++n) m[n] = f.q[n][o] * dp;
    // nwat = 6: condensate leaves the layer mass (qvapor stays)
    const double dry = dp - m[1] - m[2] - m[3] - m[4] - m[5];
    const double pn = p + dry;
    prsi[k + 1] = pn;
    prsik[k] = log(p);
#pragma unroll
    for (int n = 0; n < 7; ++n) f.q[n][o] = m[n] / dry;
    f.qsgs_tke[o] = f.qsgs_tke[o] / dry;
    const double qv = m[0] / dry;
    const double qgrs_rad = mmax(1.0e-10, qv);
    const double rtv = RDGAS * f.pt[o] * (1.0 + ZVIR * qgrs_rad);
    double prsl = dry * rtv / (phii[k] - phii[k + 1]);
    prsl = mmin(prsl, pn - 0.01 * dry);
    prsl = mmax(prsl, p + 0.01 * dry);
    f.delp[o] = prsl;
    f.dm[o] = dry;
    p = pn;
  }
  prsik[nk] = exp(KAPPA * log(p)) * pk0inv;
  prsik[0] = pktop;

  // get_prs_fv3 (get_prs_fv3.py:15-20) and the first half of get_phi_fv3
  // (get_phi_fv3.py:14-15)
  for (int k = 0; k < nk; ++k) {
    const int64_t o = b + k;
    const double t = f.pt[o];
    const double vf = 1.0 + ZVIR * mmax(0.0, f.q[0][o]);
    f.delprsi[o] = prsi[k + 1] - prsi[k];
    const double dgz = (phii[k] - phii[k + 1]) / (t * vf);
    f.del_gz[o] = dgz * t * vf;
  }
  // get_phi_fv3: geopotential rebuilt from the surface (get_phi_fv3.py:17-25)
  z = 0.0;
  phii[nk] = z;
  for (int k = nk - 1; k >= 0; --k) {
    const double d = f.del_gz[b + k];
    f.phil[b + k] = 0.5 * (z + z + d);
    z = z + d;
    phii[k] = z;
  }

  // prepare_microphysics (physics.py:124-143)
  if (f.omga) {
    for (int k = 0; k < nk; ++k) {
      const int64_t o = b + k;
      f.dz[o] = (phii[k + 1] - phii[k]) * RGRAV;
      f.wmp[o] = -f.omga[o] * (1.0 + ZVIR * f.q[0][o]) * f.pt[o] / f.delp[o] * (RDGAS * RGRAV);
#pragma unroll
      for (int n = 0; n < 10; ++n) f.tend[n][o] = 0.0;
    }
  }
}

}  // namespace

extern "C" int shield_physics_prologue(const shield_mp_shape* shape, const shield_prologue_fields* f,
                                       void* stream) {
  if (!shape || !f) return SHIELD_ERR_NULL;
  if (shape->ni < 0 || shape->nj < 0 || shape->nk < 1 || shape->si < 0 || shape->sj < 0 || f->si_i < 0 ||
      f->sj_i < 0)
    return SHIELD_ERR_SHAPE;
  if (!f->delz || !f->delp || !f->qsgs_tke || !f->pt || !f->phii || !f->prsi || !f->prsik || !f->dm ||
      !f->delprsi || !f->del_gz || !f->phil)
    return SHIELD_ERR_NULL;
  for (int n = 0; n < 7; ++n)
    if (!f->q[n]) return SHIELD_ERR_NULL;
  if (f->omga) {
    if (!f->dz || !f->wmp) return SHIELD_ERR_NULL;
    for (int n = 0; n < 10; ++n)
      if (!f->tend[n]) return SHIELD_ERR_NULL;
  }
  if (!(f->ptop > 0.0) || !isfinite(f->ptop)) return SHIELD_ERR_SHAPE;
  const int64_t ncol = shape->ni * shape->nj;
  if (ncol == 0) return SHIELD_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SHIELD_ERR_NOGPU;
  const Geo g{shape->ni, shape->nj, shape->nk, shape->si, shape->sj, f->si_i, f->sj_i};
  const double pktop = pow(f->ptop / P00, KAPPA), pk0inv = pow(1.0 / P00, KAPPA);
  hipLaunchKernelGGL(prologue_kernel, dim3((unsigned)((ncol + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     (hipStream_t)stream, g, *f, pktop, pk0inv);
  return hipGetLastError() == hipSuccess ? SHIELD_OK : SHIELD_ERR_LAUNCH;
}
