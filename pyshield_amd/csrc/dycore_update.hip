// dycore_update.hip -- the steps of ApplyPhysicsToDycore.__call__
// (pySHiELD/update/fv_update_phys.py:148-195) that pySHiELD holds, for one
// tile on one rank: moist_cv, update_pressure_and_surface_winds and
// AGrid2DGridPhysics (pySHiELD/update/update_dwind_phys.py).
//
// Two kernels and a clear:
//  * a column kernel over blocks of whole columns: the point-wise temperature
//    update and the interface-pressure recurrence summed from the top;
//  * a D-grid kernel, one thread per (i, j, k) face point with k fastest, so
//    a wavefront reads consecutive levels of the same faces: each thread
//    rebuilds the 3D tendency vectors of the two A points beside its face
//    (and, on a tile edge, of its neighbour face), blends along the edge and
//    projects on the face's unit vector;
//  * u_dt and v_dt cleared after the D-grid kernel has read them, as the
//    reference's set_winds_zero (grid_type <= 3).
#include <math.h>

#include "mp_common.h"

using namespace shield;

namespace {

constexpr int kBlock = 256;
// LDS of the column kernel's delp / pe stage; bounds nk (one column at least)
constexpr int64_t kStageBytes = 40960;

struct Dims {
  int64_t ni, nj, nk;
};

// moist_cv (fv_update_phys.py:32-50) and update_pressure_and_surface_winds
// (:53-74).  A block takes `ncb` whole columns, whose cells are one
// contiguous run of ncb * nk values (and their interfaces one run of
// ncb * (nk + 1)), so every pass below reads and writes consecutive
// addresses with consecutive threads:
//  1. moist_cv, one thread per cell; delp staged in LDS;
//  2. the interface pressures, one thread per column summing its delp from
//     the top in the reference's order (pe[k] = pe[k-1] + delp[k-1]) in LDS;
//  3. pe, peln = log pe, pk = exp(kappa peln), one thread per interface.
__global__ __launch_bounds__(kBlock) void dycore_column_kernel(Dims d, shield_dycore_fields f, int ncb) {
  extern __shared__ double stage[];  // [ncb][nk]: delp, then pe of interfaces 1..nk
  const int nk = (int)d.nk;
  const int64_t ncol = d.ni * d.nj;
  const int64_t c0 = (int64_t)blockIdx.x * ncb;
  const int nc = (int)(ncol - c0 < ncb ? ncol - c0 : ncb);
  const int ncell = nc * nk;
  const int64_t b = c0 * nk;
  for (int e = threadIdx.x; e < ncell; e += kBlock) {
    const int64_t o = b + e;
    const double qv = f.qvapor[o];
    const double ql = f.qliquid[o] + f.qrain[o];
    const double qs = f.qice[o] + f.qsnow[o] + f.qgraupel[o];
    const double gz = ql + qs;
    const double cvm = (1.0 - (qv + gz)) * CV_AIR + qv * CV_VAP + ql * C_LIQ + qs * C_ICE;
    f.pt[o] = f.pt[o] + f.t_dt[o] * f.dt * f.con_cp / cvm;
    f.t_dt[o] = 0.0;
    stage[e] = f.delp[o];
  }
  __syncthreads();
  if ((int)threadIdx.x < nc) {
    const int64_t col = c0 + threadIdx.x;
    double* s = stage + threadIdx.x * nk;
    double pe = f.pe[col * (nk + 1)];
    for (int k = 0; k < nk; ++k) {
      pe = pe + s[k];
      s[k] = pe;
    }
    f.ps[col] = pe;
    f.u_srf[col] = f.ua[col * nk + nk - 1];
    f.v_srf[col] = f.va[col * nk + nk - 1];
  }
  __syncthreads();
  const int nint = nc * (nk + 1);
  const int64_t bi = c0 * (nk + 1);
  for (int e = threadIdx.x; e < nint; e += kBlock) {
    const int cl = e / (nk + 1), k = e - cl * (nk + 1);
    if (k == 0) continue;  // the top interface is the caller's
    const double pe = stage[cl * nk + k - 1];
    const double peln = log(pe);
    f.pe[bi + e] = pe;
    f.peln[bi + e] = peln;
    f.pk[bi + e] = exp(f.kappa * peln);
  }
}

struct Wind {
  int64_t ni, nj, nk;
  double dt5;
  int im2, jm2;  // edge split points in compute coordinates
  // the 3D tendency vector component c at A point (ia, ja) and level k
  __device__ __forceinline__ double v3(const shield_dycore_fields& f, int c, int64_t ia, int64_t ja,
                                       int64_t k) const {
    const int64_t a = (ia + 1) * (nj + 2) + (ja + 1);
    const double* lon = c == 0 ? f.vlon1 : c == 1 ? f.vlon2 : f.vlon3;
    const double* lat = c == 0 ? f.vlat1 : c == 1 ? f.vlat2 : f.vlat3;
    return f.u_dt[a * nk + k] * lon[a] + f.v_dt[a * nk + k] * lat[a];
  }
};

// AGrid2DGridPhysics.__call__ (update_dwind_phys.py:490-701)
__global__ __launch_bounds__(kBlock) void dycore_dwind_kernel(Wind w, shield_dycore_fields f, int grid_type) {
  const int64_t n = (w.ni + 1) * (w.nj + 1) * w.nk;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  const int64_t k = t % w.nk, ij = t / w.nk;
  const int64_t i = ij / (w.nj + 1), j = ij - i * (w.nj + 1);
  const int64_t ni = w.ni, nj = w.nj, nk = w.nk;
  if (grid_type > 3) {
    // doubly_periodic_wind_update (:153-163), compute domain
    if (i < ni && j < nj) {
      auto A = [&](int64_t ia, int64_t ja) { return ((ia + 1) * (nj + 2) + (ja + 1)) * nk + k; };
      f.u[(i * (nj + 1) + j) * nk + k] += w.dt5 * (f.u_dt[A(i, j - 1)] + f.u_dt[A(i, j)]);
      f.v[(i * nj + j) * nk + k] += w.dt5 * (f.v_dt[A(i - 1, j)] + f.v_dt[A(i, j)]);
    }
    return;
  }
  if (i < ni) {
    // u face (i, j), j in [0, nj]: ue = v3 of the A points south and north
    // (update_dwind_prep_stencil, :21-46)
    double e[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) e[c] = w.v3(f, c, i, j - 1, k) + w.v3(f, c, i, j, k);
    if (j == 0 || j == nj) {
      // south and north tile edges: blend with the face west (i + 1 side
      // below im2) or east (above), :49-106 and the ranges of :236-433
      const bool north = j == nj;
      const double* ev = north ? f.edge_vect_n : f.edge_vect_s;
      const bool lower = (north ? 0 < w.im2 : 0 <= w.im2) && i <= (w.im2 < ni - 1 ? w.im2 : ni - 1);
      const bool upper = (north ? ni - 1 >= w.im2 : ni - 1 > w.im2) && i >= (w.im2 + 1 > 0 ? w.im2 + 1 : 0);
      if (lower || upper) {
        const int64_t nb = lower ? i + 1 : i - 1;
        const double g = ev[i];
#pragma unroll
        for (int c = 0; c < 3; ++c)
          e[c] = g * (w.v3(f, c, nb, j - 1, k) + w.v3(f, c, nb, j, k)) + (1.0 - g) * e[c];
      }
    }
    // update_uwind_stencil (:123-135)
    const int64_t fu = i * (nj + 1) + j;
    f.u[fu * nk + k] += w.dt5 * (e[0] * f.es1_1[fu] + e[1] * f.es1_2[fu] + e[2] * f.es1_3[fu]);
  }
  if (j < nj) {
    // v face (i, j), i in [0, ni]: ve = v3 of the A points west and east
    double e[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) e[c] = w.v3(f, c, i - 1, j, k) + w.v3(f, c, i, j, k);
    if (i == 0 || i == ni) {
      // west and east tile edges: blend with the face north (j + 1 side
      // below jm2) or south (above)
      const double* ev = i == 0 ? f.edge_vect_w : f.edge_vect_e;
      const bool lower = 0 <= w.jm2 && j <= (w.jm2 < nj - 1 ? w.jm2 : nj - 1);
      const bool upper = nj - 1 > w.jm2 && j >= (w.jm2 + 1 > 0 ? w.jm2 + 1 : 0);
      if (lower || upper) {
        const int64_t nb = lower ? j + 1 : j - 1;
        const double g = ev[j];
#pragma unroll
        for (int c = 0; c < 3; ++c)
          e[c] = g * (w.v3(f, c, i - 1, nb, k) + w.v3(f, c, i, nb, k)) + (1.0 - g) * e[c];
      }
    }
    // update_vwind_stencil (:138-150)
    const int64_t fv = i * nj + j;
    f.v[fv * nk + k] += w.dt5 * (e[0] * f.ew2_1[fv] + e[1] * f.ew2_2[fv] + e[2] * f.ew2_3[fv]);
  }
}

// set_winds_zero (:12-18) over the tendencies and their halo
__global__ __launch_bounds__(kBlock) void dycore_clear_kernel(double* u_dt, double* v_dt, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  u_dt[t] = 0.0;
  v_dt[t] = 0.0;
}

}  // namespace

static_assert(kStageBytes / (8 * SHIELD_DYCORE_MAX_NK) >= 1, "one column of delp must fit the LDS stage");

namespace shield {

hipError_t launch_dycore_update(const shield_mp_shape& sh, const shield_dycore_fields& f, hipStream_t stream) {
  const Dims d{sh.ni, sh.nj, sh.nk};
  const int64_t ncol = sh.ni * sh.nj;
  // columns per block: up to 64, within kStageBytes of LDS
  const int ncb = (int)(kStageBytes / (8 * sh.nk) < 64 ? kStageBytes / (8 * sh.nk) : 64);
  hipLaunchKernelGGL(dycore_column_kernel, dim3((unsigned)((ncol + ncb - 1) / ncb)), dim3(kBlock),
                     (size_t)ncb * sh.nk * sizeof(double), stream, d, f, ncb);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  Wind w;
  w.ni = sh.ni;
  w.nj = sh.nj;
  w.nk = sh.nk;
  w.dt5 = 0.5 * f.dt;
  // the reference's im2 = (npx - 1) / 2 + 2 in local indices with a 3-point
  // halo (update_dwind_phys.py:187-188)
  w.im2 = (f.npx - 1) / 2 + 2 - 3;
  w.jm2 = (f.npy - 1) / 2 + 2 - 3;
  const int64_t nw = (sh.ni + 1) * (sh.nj + 1) * sh.nk;
  hipLaunchKernelGGL(dycore_dwind_kernel, dim3((unsigned)((nw + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream, w,
                     f, (int)f.grid_type);
  e = hipGetLastError();
  if (e != hipSuccess || f.grid_type > 3) return e;
  const int64_t nh = (sh.ni + 2) * (sh.nj + 2) * sh.nk;
  hipLaunchKernelGGL(dycore_clear_kernel, dim3((unsigned)((nh + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream,
                     f.u_dt, f.v_dt, nh);
  return hipGetLastError();
}

}  // namespace shield

extern "C" int shield_apply_physics_to_dycore(const shield_mp_shape* shape, const shield_dycore_fields* f,
                                              void* stream) {
  if (!shape || !f) return SHIELD_ERR_NULL;
  const void* need[] = {f->qvapor, f->qliquid, f->qrain, f->qsnow, f->qice, f->qgraupel, f->pt, f->t_dt,
                        f->delp, f->pe, f->peln, f->pk, f->ps, f->ua, f->va, f->u_srf, f->v_srf,
                        f->u, f->v, f->u_dt, f->v_dt};
  for (const void* q : need)
    if (!q) return SHIELD_ERR_NULL;
  const void* grid[] = {f->vlon1, f->vlon2, f->vlon3, f->vlat1, f->vlat2, f->vlat3, f->es1_1, f->es1_2,
                        f->es1_3, f->ew2_1, f->ew2_2, f->ew2_3, f->edge_vect_w, f->edge_vect_e,
                        f->edge_vect_s, f->edge_vect_n};
  if (f->grid_type <= 3)
    for (const void* q : grid)
      if (!q) return SHIELD_ERR_NULL;
  // every written field distinct from the others and from every input
  const void* out[] = {f->pt, f->t_dt, f->pe, f->peln, f->pk, f->ps, f->u_srf, f->v_srf,
                       f->u, f->v, f->u_dt, f->v_dt};
  const void* in[] = {f->qvapor, f->qliquid, f->qrain, f->qsnow, f->qice, f->qgraupel, f->delp, f->ua, f->va};
  for (size_t a = 0; a < sizeof(out) / sizeof(out[0]); ++a) {
    for (size_t b = a + 1; b < sizeof(out) / sizeof(out[0]); ++b)
      if (out[a] == out[b]) return SHIELD_ERR_ALIAS;
    for (const void* q : in)
      if (out[a] == q) return SHIELD_ERR_ALIAS;
    for (const void* q : grid)
      if (q && out[a] == q) return SHIELD_ERR_ALIAS;
  }
  if (shape->ni < 0 || shape->nj < 0 || shape->nk < 1 || shape->nk > SHIELD_DYCORE_MAX_NK) return SHIELD_ERR_SHAPE;
  if (shape->ni == 0 || shape->nj == 0) return SHIELD_OK;
  // one tile on one rank: the edge split points of the wind update lie
  // inside the tile
  if (shape->ni < 2 || shape->nj < 2) return SHIELD_ERR_SHAPE;
  if (f->grid_type <= 3 && (f->npx - 1 != shape->ni || f->npy - 1 != shape->nj)) return SHIELD_ERR_SHAPE;
  if (!(f->dt > 0.0) || !isfinite(f->dt)) return SHIELD_ERR_TIMESTEP;
  if (!isfinite(f->con_cp) || !isfinite(f->kappa)) return SHIELD_ERR_SHAPE;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SHIELD_ERR_NOGPU;
  return launch_dycore_update(*shape, *f, (hipStream_t)stream) == hipSuccess ? SHIELD_OK : SHIELD_ERR_LAUNCH;
}
