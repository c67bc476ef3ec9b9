// prologue.hip -- the column-local part of Physics.__call__ before the
// microphysics (pySHiELD/stencils/physics.py:278-330): atmos_phys_driver_statein,
// get_prs_fv3, get_phi_fv3 and prepare_microphysics in one kernel.
//
// The three stencils hold three vertical recurrences (geopotential summed
// from the surface, interface pressure summed from the top, geopotential
// rebuilt from del_gz); everything else is point-wise.  A block takes `ncb`
// consecutive columns.  Point-wise passes run one thread per cell (or per
// interface), k fastest, so consecutive threads touch consecutive addresses
// of a column; the recurrences run one thread per column over LDS, in the
// reference's order of operations, so every output is the value a
// column-sequential evaluation gives.
#include <math.h>

#include "mp_common.h"

using namespace shield;

namespace {

constexpr int kBlock = 256;
// LDS of the stage: per column phii and prsi (nk + 1 each) and one layer
// field (nk); bounds nk (one column at least)
constexpr int64_t kStageBytes = 40960;
constexpr double KAPPA = RDGAS / CP_AIR;
constexpr double P00 = 1.0e5;
constexpr double ZVIR = RVGAS / RDGAS - 1.0;

struct Geo {
  int64_t ni, nj, nk, si, sj, si_i, sj_i;
};

__global__ __launch_bounds__(kBlock) void prologue_kernel(Geo g, shield_prologue_fields f, double pktop,
                                                          double pk0inv, int ncb) {
  extern __shared__ double stage[];
  const int nk = (int)g.nk, nk1 = nk + 1;
  const int64_t ncol = g.ni * g.nj;
  const int64_t c0 = (int64_t)blockIdx.x * ncb;
  const int nc = (int)(ncol - c0 < ncb ? ncol - c0 : ncb);
  double* zs = stage;                 // [ncb][nk + 1]: geopotential at interfaces
  double* ps = stage + ncb * nk1;     // [ncb][nk + 1]: interface pressure
  double* ls = stage + 2 * ncb * nk1; // [ncb][nk]: dry layer mass, then del_gz
  auto base = [&](int cl) {
    const int64_t col = c0 + cl, i = col / g.nj, j = col - i * g.nj;
    return i * g.si + j * g.sj;
  };
  auto ibase = [&](int cl) {
    const int64_t col = c0 + cl, i = col / g.nj, j = col - i * g.nj;
    return i * g.si_i + j * g.sj_i;
  };
  const int ncell = nc * nk, nint = nc * nk1;

  // atmos_phys_driver_statein (physics.py:46-102), point-wise part: the
  // geopotential increments, the dry layer mass (nwat = 6: condensate
  // leaves the layer mass, qvapor stays) and the tracers per dry mass
  for (int e = threadIdx.x; e < ncell; e += kBlock) {
    const int cl = e / nk, k = e - cl * nk;
    const int64_t o = base(cl) + k;
    zs[cl * nk1 + k] = f.delz[o] * GRAV;
    const double dp = f.delp[o];
    double m[7];
#pragma unroll
    for (int n = 0; n < 7; ++n) m[n] = f.q[n][o] * dp;
    const double dry = dp - m[1] - m[2] - m[3] - m[4] - m[5];
#pragma unroll
    for (int n = 0; n < 7; ++n) f.q[n][o] = m[n] / dry;
    f.qsgs_tke[o] = f.qsgs_tke[o] / dry;
    f.dm[o] = dry;
    ls[cl * nk + k] = dry;
  }
  __syncthreads();
  // the recurrences: geopotential from the surface, pressure from the top
  if ((int)threadIdx.x < nc) {
    double* z = zs + threadIdx.x * nk1;
    double* p = ps + threadIdx.x * nk1;
    const double* dry = ls + threadIdx.x * nk;
    double zz = 0.0;
    z[nk] = zz;
    for (int k = nk - 1; k >= 0; --k) {
      zz = zz - z[k];
      z[k] = zz;
    }
    double pp = f.ptop;
    p[0] = pp;
    for (int k = 0; k < nk; ++k) {
      pp = pp + dry[k];
      p[k + 1] = pp;
    }
  }
  __syncthreads();
  // the layer-mean pressure (written into delp, as the reference does),
  // get_prs_fv3 (get_prs_fv3.py:15-20), the first half of get_phi_fv3
  // (get_phi_fv3.py:14-15) and, with omga, wmp and the zeroed tendencies of
  // prepare_microphysics (physics.py:124-143)
  for (int e = threadIdx.x; e < ncell; e += kBlock) {
    const int cl = e / nk, k = e - cl * nk;
    const int64_t o = base(cl) + k;
    const double* z = zs + cl * nk1;
    const double* p = ps + cl * nk1;
    const double dry = ls[cl * nk + k];
    const double t = f.pt[o];
    const double qv = f.q[0][o];
    const double qgrs_rad = mmax(1.0e-10, qv);
    const double rtv = RDGAS * t * (1.0 + ZVIR * qgrs_rad);
    double prsl = dry * rtv / (z[k] - z[k + 1]);
    prsl = mmin(prsl, p[k + 1] - 0.01 * dry);
    prsl = mmax(prsl, p[k] + 0.01 * dry);
    f.delp[o] = prsl;
    f.delprsi[o] = p[k + 1] - p[k];
    const double vf = 1.0 + ZVIR * mmax(0.0, qv);
    const double dgz = (z[k] - z[k + 1]) / (t * vf);
    const double del_gz = dgz * t * vf;
    f.del_gz[o] = del_gz;
    ls[cl * nk + k] = del_gz;
    if (f.omga) {
      f.wmp[o] = -f.omga[o] * (1.0 + ZVIR * qv) * t / prsl * (RDGAS * RGRAV);
#pragma unroll
      for (int n = 0; n < 10; ++n) f.tend[n][o] = 0.0;
    }
  }
  // interface pressure and its log (prsik: log p above the surface, p^kappa
  // / p00^kappa at the surface, the top's from ptop)
  for (int e = threadIdx.x; e < nint; e += kBlock) {
    const int cl = e / nk1, k = e - cl * nk1;
    const int64_t o = ibase(cl) + k;
    const double p = ps[cl * nk1 + k];
    f.prsi[o] = p;
    f.prsik[o] = k == 0 ? pktop : k == nk ? exp(KAPPA * log(p)) * pk0inv : log(p);
  }
  __syncthreads();
  // get_phi_fv3: geopotential rebuilt from the surface (get_phi_fv3.py:17-25)
  if ((int)threadIdx.x < nc) {
    double* z = zs + threadIdx.x * nk1;
    const double* d = ls + threadIdx.x * nk;
    double zz = 0.0;
    z[nk] = zz;
    for (int k = nk - 1; k >= 0; --k) {
      zz = zz + d[k];
      z[k] = zz;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ncell; e += kBlock) {
    const int cl = e / nk, k = e - cl * nk;
    const int64_t o = base(cl) + k;
    const double* z = zs + cl * nk1;
    const double zb = z[k + 1];
    f.phil[o] = 0.5 * (zb + zb + ls[cl * nk + k]);
    if (f.omga) f.dz[o] = (z[k + 1] - z[k]) * RGRAV;
  }
  for (int e = threadIdx.x; e < nint; e += kBlock) {
    const int cl = e / nk1, k = e - cl * nk1;
    f.phii[ibase(cl) + k] = zs[cl * nk1 + k];
  }
}

static_assert(kStageBytes / (8 * (3 * SHIELD_PROLOGUE_MAX_NK + 2)) >= 1, "one column must fit the LDS stage");

}  // namespace

extern "C" int shield_physics_prologue(const shield_mp_shape* shape, const shield_prologue_fields* f,
                                       void* stream) {
  if (!shape || !f) return SHIELD_ERR_NULL;
  if (shape->ni < 0 || shape->nj < 0 || shape->nk < 1 || shape->nk > SHIELD_PROLOGUE_MAX_NK || shape->si < 0 ||
      shape->sj < 0 || f->si_i < 0 || f->sj_i < 0)
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
  // columns per block: up to 64 within kStageBytes of LDS
  const int64_t per_col = 8 * (3 * shape->nk + 2);
  const int ncb = (int)(kStageBytes / per_col < 64 ? kStageBytes / per_col : 64);
  hipLaunchKernelGGL(prologue_kernel, dim3((unsigned)((ncol + ncb - 1) / ncb)), dim3(kBlock),
                     (size_t)ncb * per_col, (hipStream_t)stream, g, *f, pktop, pk0inv, ncb);
  return hipGetLastError() == hipSuccess ? SHIELD_OK : SHIELD_ERR_LAUNCH;
}
