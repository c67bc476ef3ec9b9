// dycore_update.hip -- the steps of ApplyPhysicsToDycore.__call__
// (pySHiELD/update/fv_update_phys.py:148-195) that pySHiELD holds, for one
// tile on one rank: moist_cv, update_pressure_and_surface_winds and
// AGrid2DGridPhysics (pySHiELD/update/update_dwind_phys.py).
//
// Two kernels and a clear:
//  * a column kernel over blocks of whole columns: the point-wise temperature
//    update and the interface-pressure recurrence summed from the top;
//  * a D-grid kernel over tiles of faces and levels: the block builds the 3D
//    tendency vectors of the A points its faces read in LDS, then each
//    thread averages the two beside its face (and, on a tile edge, those of
//    its neighbour face), blends along the edge and projects on the face's
//    unit vector (the doubly periodic form: a thread per point, no tiles);
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
};

// The D-grid kernel's block: TI x TJ faces (i, j) by KC levels, k fastest so
// a wavefront's lanes read 16-level runs of 4 faces.  The block first builds
// the 3D tendency vector u_dt vlon + v_dt vlat of each of the (TI + 2) x
// (TJ + 2) A points its faces and their edge neighbours read (A points
// i0 - 1 .. i0 + TI, j0 - 1 .. j0 + TJ) in LDS, once per point and level.
constexpr int TI = 4, TJ = 4, KC = 16, AI = TI + 2, AJ = TJ + 2;
static_assert(TI * TJ * KC == kBlock, "one thread per face and level of the tile");

// Workgroups are dispatched to the XCDs round robin (block b to XCD b % 8);
// the (tile, level chunk) pairs are numbered with the chunk fastest and each
// XCD takes one contiguous range of them, so a tile's chunks and its
// neighbours, which read the same halo A points and cache lines, share an L2.
constexpr int kXcds = 8;

// AGrid2DGridPhysics.__call__ for grid_type <= 3 (update_dwind_phys.py:490-701)
__global__ __launch_bounds__(kBlock) void dycore_dwind_kernel(Wind w, shield_dycore_fields f, int64_t per_xcd) {
  __shared__ double v3s[3][AI * AJ * KC];
  __shared__ double met[6][AI * AJ];  // vlon1..3, vlat1..3 of the tile's A points
  const int64_t ni = w.ni, nj = w.nj, nk = w.nk;
  const int64_t nkc = (nk + KC - 1) / KC, ntj = (nj + 1 + TJ - 1) / TJ, nti = (ni + 1 + TI - 1) / TI;
  const int64_t job = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (job >= nti * ntj * nkc) return;  // the whole block
  const int64_t k0 = (job % nkc) * KC, tile = job / nkc;
  const int64_t i0 = (tile / ntj) * TI, j0 = (tile % ntj) * TJ;
  // the grid metrics do not depend on k: one load per A point
  if (threadIdx.x < AI * AJ) {
    const int p = threadIdx.x;
    const int64_t ia = i0 - 1 + p / AJ, ja = j0 - 1 + p % AJ;
    const bool in = ia <= ni && ja <= nj;
    const int64_t a = in ? (ia + 1) * (nj + 2) + (ja + 1) : 0;
    met[0][p] = in ? f.vlon1[a] : 0.0;
    met[1][p] = in ? f.vlon2[a] : 0.0;
    met[2][p] = in ? f.vlon3[a] : 0.0;
    met[3][p] = in ? f.vlat1[a] : 0.0;
    met[4][p] = in ? f.vlat2[a] : 0.0;
    met[5][p] = in ? f.vlat3[a] : 0.0;
  }
  __syncthreads();
  // update_dwind_prep_stencil (:21-46): v3 at the tile's A points; points
  // outside the tendencies' one-point halo are never read by a face
  for (int e = threadIdx.x; e < AI * AJ * KC; e += kBlock) {
    const int ek = e % KC, p = e / KC;
    const int64_t ia = i0 - 1 + p / AJ, ja = j0 - 1 + p % AJ, k = k0 + ek;
    double x0 = 0.0, x1 = 0.0, x2 = 0.0;
    if (ia <= ni && ja <= nj && k < nk) {
      const int64_t a = (ia + 1) * (nj + 2) + (ja + 1);
      const double ud = f.u_dt[a * nk + k], vd = f.v_dt[a * nk + k];
      x0 = ud * met[0][p] + vd * met[3][p];
      x1 = ud * met[1][p] + vd * met[4][p];
      x2 = ud * met[2][p] + vd * met[5][p];
    }
    v3s[0][e] = x0;
    v3s[1][e] = x1;
    v3s[2][e] = x2;
  }
  __syncthreads();
  const int kk = threadIdx.x % KC, r = threadIdx.x / KC;
  const int ti = r / TJ, tj = r % TJ;
  const int64_t i = i0 + ti, j = j0 + tj, k = k0 + kk;
  if (k >= nk || i > ni || j > nj) return;
  // the vector component c at A point (i0 - 1 + pi, j0 - 1 + pj)
  auto V = [&](int c, int pi, int pj) { return v3s[c][(pi * AJ + pj) * KC + kk]; };
  if (i < ni) {
    // u face (i, j), j in [0, nj]: the A points south (j - 1) and north (j)
    double e[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) e[c] = V(c, ti + 1, tj) + V(c, ti + 1, tj + 1);
    if (j == 0 || j == nj) {
      // south and north tile edges: blend with the face at i + 1 below im2,
      // at i - 1 above (:49-106 and the ranges of :236-433)
      const bool north = j == nj;
      const double* ev = north ? f.edge_vect_n : f.edge_vect_s;
      const bool lower = (north ? 0 < w.im2 : 0 <= w.im2) && i <= (w.im2 < ni - 1 ? w.im2 : ni - 1);
      const bool upper = (north ? ni - 1 >= w.im2 : ni - 1 > w.im2) && i >= (w.im2 + 1 > 0 ? w.im2 + 1 : 0);
      if (lower || upper) {
        const int pb = lower ? ti + 2 : ti;
        const double g = ev[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) e[c] = g * (V(c, pb, tj) + V(c, pb, tj + 1)) + (1.0 - g) * e[c];
      }
    }
    // update_uwind_stencil (:123-135)
    const int64_t fu = i * (nj + 1) + j;
    f.u[fu * nk + k] += w.dt5 * (e[0] * f.es1_1[fu] + e[1] * f.es1_2[fu] + e[2] * f.es1_3[fu]);
  }
  if (j < nj) {
    // v face (i, j), i in [0, ni]: the A points west (i - 1) and east (i)
    double e[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) e[c] = V(c, ti, tj + 1) + V(c, ti + 1, tj + 1);
    if (i == 0 || i == ni) {
      // west and east tile edges: blend with the face at j + 1 below jm2,
      // at j - 1 above
      const double* ev = i == 0 ? f.edge_vect_w : f.edge_vect_e;
      const bool lower = 0 <= w.jm2 && j <= (w.jm2 < nj - 1 ? w.jm2 : nj - 1);
      const bool upper = nj - 1 > w.jm2 && j >= (w.jm2 + 1 > 0 ? w.jm2 + 1 : 0);
      if (lower || upper) {
        const int pb = lower ? tj + 2 : tj;
        const double g = ev[j];
#pragma unroll
        for (int c = 0; c < 3; ++c) e[c] = g * (V(c, ti, pb) + V(c, ti + 1, pb)) + (1.0 - g) * e[c];
      }
    }
    // update_vwind_stencil (:138-150)
    const int64_t fv = i * nj + j;
    f.v[fv * nk + k] += w.dt5 * (e[0] * f.ew2_1[fv] + e[1] * f.ew2_2[fv] + e[2] * f.ew2_3[fv]);
  }
}

// doubly_periodic_wind_update (:153-163) over the compute domain, one thread
// per (i, j, k) with k fastest
__global__ __launch_bounds__(kBlock) void dycore_periodic_kernel(Wind w, shield_dycore_fields f) {
  const int64_t ni = w.ni, nj = w.nj, nk = w.nk;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= ni * nj * nk) return;
  const int64_t k = t % nk, ij = t / nk;
  const int64_t i = ij / nj, j = ij - i * nj;
  auto A = [&](int64_t ia, int64_t ja) { return ((ia + 1) * (nj + 2) + (ja + 1)) * nk + k; };
  f.u[(i * (nj + 1) + j) * nk + k] += w.dt5 * (f.u_dt[A(i, j - 1)] + f.u_dt[A(i, j)]);
  f.v[(i * nj + j) * nk + k] += w.dt5 * (f.v_dt[A(i - 1, j)] + f.v_dt[A(i, j)]);
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
  if (f.grid_type > 3) {
    const int64_t np = sh.ni * sh.nj * sh.nk;
    hipLaunchKernelGGL(dycore_periodic_kernel, dim3((unsigned)((np + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream,
                       w, f);
    return hipGetLastError();
  }
  const int64_t njob = ((sh.ni + 1 + TI - 1) / TI) * ((sh.nj + 1 + TJ - 1) / TJ) * ((sh.nk + KC - 1) / KC);
  const int64_t per_xcd = (njob + kXcds - 1) / kXcds;
  hipLaunchKernelGGL(dycore_dwind_kernel, dim3((unsigned)(per_xcd * kXcds)), dim3(kBlock), 0, stream, w, f,
                     per_xcd);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
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
