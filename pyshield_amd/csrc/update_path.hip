// update_path.hip -- the column-local state-update path around the
// microphysics: prepare_microphysics and update_physics_state_with_tendencies
// (pySHiELD/stencils/physics.py:105-194), fill_gfs_delp and
// prepare_tendencies_and_update_tracers
// (pySHiELD/update/update_atmos_state.py:21-92).
//
// The point-wise stencils run one thread per cell over the flattened
// (column, k) index, so dense [X, Y, Z] storage is read fully coalesced.
// fill_gfs_delp is a pair of non-linear vertical recurrences; it stages
// blocks of whole columns in LDS and runs one thread per column there.
#include <math.h>

#include "mp_common.h"

using namespace shield;

namespace {

constexpr int kBlock = 256;
// LDS of fill_gfs_delp's stage (q and delp of the block's columns)
constexpr int64_t kFillStageBytes = 40960;
static_assert(kFillStageBytes / (16 * SHIELD_FILL_MAX_NK) >= 1, "one column must fit the LDS stage");

struct Idx {
  int64_t ni, nj, nk, si, sj;
  int64_t si_i, sj_i;  // strides of the nk+1-level interface field
  __device__ __forceinline__ int64_t off(int64_t cell) const {
    const int64_t col = cell / nk, k = cell - col * nk;
    const int64_t i = col / nj, j = col - i * nj;
    return i * si + j * sj + k;
  }
  __device__ __forceinline__ int64_t off_i(int64_t cell) const {
    const int64_t col = cell / nk, k = cell - col * nk;
    const int64_t i = col / nj, j = col - i * nj;
    return i * si_i + j * sj_i + k;
  }
};

__global__ __launch_bounds__(kBlock) void prepare_mp_kernel(Idx ix, double* dz, const double* phii, double* wmp,
                                                            const double* omga, const double* qv, const double* pt,
                                                            const double* delp, double* t0, double* t1, double* t2,
                                                            double* t3, double* t4, double* t5, double* t6,
                                                            double* t7, double* t8, double* t9) {
  const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (cell >= ix.ni * ix.nj * ix.nk) return;
  const int64_t o = ix.off(cell), oi = ix.off_i(cell);
  dz[o] = (phii[oi + 1] - phii[oi]) * RGRAV;
  const double zvir = RVGAS / RDGAS - 1.0;
  wmp[o] = -omga[o] * (1.0 + zvir * qv[o]) * pt[o] / delp[o] * (RDGAS * RGRAV);
  t0[o] = 0.0; t1[o] = 0.0; t2[o] = 0.0; t3[o] = 0.0; t4[o] = 0.0;
  t5[o] = 0.0; t6[o] = 0.0; t7[o] = 0.0; t8[o] = 0.0; t9[o] = 0.0;
}

struct Ten {
  const double* p[10];
};
struct TenW {
  double* p[10];
};

__global__ __launch_bounds__(kBlock) void forward_euler_kernel(Idx ix, Ten in, Ten tend, TenW out, double dt) {
  const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (cell >= ix.ni * ix.nj * ix.nk) return;
  const int64_t o = ix.off(cell);
#pragma unroll
  for (int n = 0; n < 10; ++n) out.p[n][o] = in.p[n][o] + tend.p[n][o] * dt;
}

// fill_gfs_delp (update_atmos_state.py:21-37) over a block of `ncb`
// consecutive columns: q and delp are staged in LDS with consecutive threads
// on consecutive levels, one thread per column runs the vertical passes there
// (every element gets the reference's operations in the reference's order),
// and q is written back the same way.
__global__ __launch_bounds__(kBlock) void fill_gfs_delp_kernel(Idx ix, const double* delp, double* q, double q_min,
                                                               int ncb) {
  extern __shared__ double stage[];  // [ncb][nk] of q, then [ncb][nk] of delp
  const int nk = (int)ix.nk;
  const int64_t ncol = ix.ni * ix.nj;
  const int64_t c0 = (int64_t)blockIdx.x * ncb;
  const int nc = (int)(ncol - c0 < ncb ? ncol - c0 : ncb);
  double* qs = stage;
  double* ds = stage + ncb * nk;
  auto base = [&](int cl) {
    const int64_t col = c0 + cl, i = col / ix.nj, j = col - i * ix.nj;
    return i * ix.si + j * ix.sj;
  };
  for (int e = threadIdx.x; e < nc * nk; e += kBlock) {
    const int cl = e / nk, k = e - cl * nk;
    const int64_t o = base(cl) + k;
    qs[e] = q[o];
    ds[e] = delp[o];
  }
  __syncthreads();
  if ((int)threadIdx.x < nc) shield::fill_column(qs + threadIdx.x * nk, ds + threadIdx.x * nk, nk, q_min);
  __syncthreads();
  for (int e = threadIdx.x; e < nc * nk; e += kBlock) {
    const int cl = e / nk, k = e - cl * nk;
    q[base(cl) + k] = qs[e];
  }
}

struct TracerArgs {
  double *u_dt, *v_dt, *pt_dt;
  const double *u_t1, *v_t1, *pt_t1, *qv_t1, *ql_t1, *qr_t1, *qs_t1, *qi_t1, *qg_t1;
  const double *u_t0, *v_t0, *pt_t0;
  double *qv_t0, *ql_t0, *qr_t0, *qs_t0, *qi_t0, *qg_t0;
  const double* prsi;
  double* delp;
};

__global__ __launch_bounds__(kBlock) void update_tracers_kernel(Idx ix, TracerArgs a, double rdt) {
  const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (cell >= ix.ni * ix.nj * ix.nk) return;
  const int64_t o = ix.off(cell), oi = ix.off_i(cell);
  a.u_dt[o] += (a.u_t1[o] - a.u_t0[o]) * rdt;
  a.v_dt[o] += (a.v_t1[o] - a.v_t0[o]) * rdt;
  a.pt_dt[o] += (a.pt_t1[o] - a.pt_t0[o]) * rdt;
  const double dp = a.prsi[oi + 1] - a.prsi[oi];
  const double qwat_qv = dp * a.qv_t1[o];
  const double qwat_ql = dp * a.ql_t1[o];
  const double qwat_qr = dp * a.qr_t1[o];
  const double qwat_qs = dp * a.qs_t1[o];
  const double qwat_qi = dp * a.qi_t1[o];
  const double qwat_qg = dp * a.qg_t1[o];
  const double qt = qwat_qv + qwat_ql + qwat_qr + qwat_qs + qwat_qi + qwat_qg;
  const double q_sum = a.qv_t0[o] + a.ql_t0[o] + a.qr_t0[o] + a.qs_t0[o] + a.qi_t0[o] + a.qg_t0[o];
  const double q0 = a.delp[o] * (1.0 - q_sum) + qt;
  a.delp[o] = q0;
  a.qv_t0[o] = qwat_qv / q0;
  a.ql_t0[o] = qwat_ql / q0;
  a.qr_t0[o] = qwat_qr / q0;
  a.qs_t0[o] = qwat_qs / q0;
  a.qi_t0[o] = qwat_qi / q0;
  a.qg_t0[o] = qwat_qg / q0;
}

bool shape_ok(const shield_mp_shape* s) {
  return s && s->ni >= 0 && s->nj >= 0 && s->nk >= 1 && s->si >= 0 && s->sj >= 0 && s->si2 >= 0 && s->sj2 >= 0;
}

unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

int finish() { return hipGetLastError() == hipSuccess ? SHIELD_OK : SHIELD_ERR_LAUNCH; }

}  // namespace

extern "C" {

int shield_prepare_microphysics(const shield_mp_shape* shape, double* dz, const double* phii, double* wmp,
                                const double* omga, const double* qvapor, const double* pt, const double* delp,
                                double* const tend[10], void* stream) {
  if (!shape_ok(shape)) return SHIELD_ERR_SHAPE;
  if (!dz || !phii || !wmp || !omga || !qvapor || !pt || !delp || !tend) return SHIELD_ERR_NULL;
  for (int n = 0; n < 10; ++n)
    if (!tend[n]) return SHIELD_ERR_NULL;
  const int64_t n = shape->ni * shape->nj * shape->nk;
  if (n == 0) return SHIELD_OK;
  const Idx ix{shape->ni, shape->nj, shape->nk, shape->si, shape->sj, shape->si2, shape->sj2};
  hipLaunchKernelGGL(prepare_mp_kernel, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream, ix, dz, phii, wmp,
                     omga, qvapor, pt, delp, tend[0], tend[1], tend[2], tend[3], tend[4], tend[5], tend[6], tend[7],
                     tend[8], tend[9]);
  return finish();
}

int shield_update_physics_state_with_tendencies(const shield_mp_shape* shape, const double* const in[10],
                                                const double* const tend[10], double* const out[10], double dt,
                                                void* stream) {
  if (!shape_ok(shape)) return SHIELD_ERR_SHAPE;
  if (!in || !tend || !out) return SHIELD_ERR_NULL;
  Ten a, b;
  TenW c;
  for (int n = 0; n < 10; ++n) {
    if (!in[n] || !tend[n] || !out[n]) return SHIELD_ERR_NULL;
    a.p[n] = in[n];
    b.p[n] = tend[n];
    c.p[n] = out[n];
  }
  const int64_t n = shape->ni * shape->nj * shape->nk;
  if (n == 0) return SHIELD_OK;
  const Idx ix{shape->ni, shape->nj, shape->nk, shape->si, shape->sj, shape->si2, shape->sj2};
  hipLaunchKernelGGL(forward_euler_kernel, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream, ix, a, b, c, dt);
  return finish();
}

int shield_fill_gfs_delp(const shield_mp_shape* shape, const double* delp, double* q, double q_min, void* stream) {
  if (!shape_ok(shape) || shape->nk > SHIELD_FILL_MAX_NK) return SHIELD_ERR_SHAPE;
  if (!delp || !q) return SHIELD_ERR_NULL;
  const int64_t ncol = shape->ni * shape->nj;
  if (ncol == 0) return SHIELD_OK;
  const Idx ix{shape->ni, shape->nj, shape->nk, shape->si, shape->sj, shape->si2, shape->sj2};
  // columns per block: up to 64 within kFillStageBytes of LDS
  const int64_t per_col = 16 * shape->nk;
  const int ncb = (int)(kFillStageBytes / per_col < 64 ? kFillStageBytes / per_col : 64);
  hipLaunchKernelGGL(fill_gfs_delp_kernel, dim3((unsigned)((ncol + ncb - 1) / ncb)), dim3(kBlock),
                     (size_t)ncb * per_col, (hipStream_t)stream, ix, delp, q, q_min, ncb);
  return finish();
}

int shield_prepare_tendencies_and_update_tracers(
    const shield_mp_shape* shape, double* u_dt, double* v_dt, double* pt_dt, const double* u_t1, const double* v_t1,
    const double* pt_t1, const double* qv_t1, const double* ql_t1, const double* qr_t1, const double* qs_t1,
    const double* qi_t1, const double* qg_t1, const double* u_t0, const double* v_t0, const double* pt_t0,
    double* qv_t0, double* ql_t0, double* qr_t0, double* qs_t0, double* qi_t0, double* qg_t0, const double* prsi,
    double* delp, double rdt, void* stream) {
  if (!shape_ok(shape)) return SHIELD_ERR_SHAPE;
  const TracerArgs a{u_dt,  v_dt,  pt_dt, u_t1,  v_t1,  pt_t1, qv_t1, ql_t1, qr_t1, qs_t1, qi_t1, qg_t1,
                     u_t0,  v_t0,  pt_t0, qv_t0, ql_t0, qr_t0, qs_t0, qi_t0, qg_t0, prsi,  delp};
  const void* ptrs[] = {u_dt,  v_dt,  pt_dt, u_t1,  v_t1,  pt_t1, qv_t1, ql_t1, qr_t1, qs_t1, qi_t1, qg_t1,
                        u_t0,  v_t0,  pt_t0, qv_t0, ql_t0, qr_t0, qs_t0, qi_t0, qg_t0, prsi,  delp};
  for (const void* p : ptrs)
    if (!p) return SHIELD_ERR_NULL;
  const int64_t n = shape->ni * shape->nj * shape->nk;
  if (n == 0) return SHIELD_OK;
  const Idx ix{shape->ni, shape->nj, shape->nk, shape->si, shape->sj, shape->si2, shape->sj2};
  hipLaunchKernelGGL(update_tracers_kernel, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream, ix, a, rdt);
  return finish();
}

}  // extern "C"
