"""Oracle (test infrastructure only): NumPy restatement of the reference's
ApplyPhysicsToDycore steps that pySHiELD itself holds, for one tile on one
rank (layout 1 x 1, so all four tile edges are local):

* moist_cv (pySHiELD/update/fv_update_phys.py:32-50): pt += t_dt dt cp / cvm,
  then t_dt = 0;
* update_pressure_and_surface_winds (fv_update_phys.py:53-74): pe summed down
  from the top, peln = log pe and pk = exp(kappa peln) below the top
  interface, ps the surface interface, u_srf / v_srf the lowest layer's winds;
* AGrid2DGridPhysics (pySHiELD/update/update_dwind_phys.py:166-701): the A-grid
  wind tendencies as 3D vectors (vlon, vlat), averaged to the D-grid faces,
  blended along the tile edges (edge_vect_*), projected on the face unit
  vectors (es1, ew2) and added with dt/2; the doubly periodic form for
  grid_type > 3.

Arrays are in compute coordinates: i in [0, ni), j in [0, nj), k = 0 the
top.  u_dt, v_dt, vlon*, vlat* carry a one-point halo ([ni + 2, nj + 2, ...],
compute point (i, j) at [i + 1, j + 1]); u is [ni, nj + 1, nk], v
[ni + 1, nj, nk]; es1_* [ni, nj + 1]; ew2_* [ni + 1, nj]; edge_vect_w / _e
[nj], edge_vect_s / _n [ni].  The reference's halo'd local indices are
compute indices + 3 (n_halo), so its im2 = (npx - 1) / 2 + 2 is
(npx - 1) / 2 - 1 here.
"""

import numpy as np

CV_AIR = 1004.6 - 287.05
CV_VAP = 3.0 * 461.50
C_LIQ = 4185.5
C_ICE = 1972.0
N_HALO = 3


def moist_cv(q, pt, t_dt, con_cp, dt):
    """fv_update_phys.py:32-50; q: dict of qvapor, qliquid, qrain, qsnow,
    qice, qgraupel.  Returns (pt, t_dt)."""
    ql = q["qliquid"] + q["qrain"]
    qs = q["qice"] + q["qsnow"] + q["qgraupel"]
    gz = ql + qs
    cvm = (1.0 - (q["qvapor"] + gz)) * CV_AIR + q["qvapor"] * CV_VAP + ql * C_LIQ + qs * C_ICE
    return pt + t_dt * dt * con_cp / cvm, np.zeros_like(t_dt)


def pressure_and_surface_winds(pe, delp, peln, pk, kappa, ua, va):
    """fv_update_phys.py:53-74.  pe, peln, pk: [ni, nj, nk + 1]."""
    pe, peln, pk = pe.copy(), peln.copy(), pk.copy()
    nk = delp.shape[2]
    for k in range(1, nk + 1):
        pe[:, :, k] = pe[:, :, k - 1] + delp[:, :, k - 1]
    peln[:, :, 1:] = np.log(pe[:, :, 1:])
    pk[:, :, 1:] = np.exp(kappa * peln[:, :, 1:])
    return pe, peln, pk, pe[:, :, nk].copy(), ua[:, :, nk - 1].copy(), va[:, :, nk - 1].copy()


def _ranges(n, m2, lower_strict):
    """The lower / upper index ranges of an edge split at m2 (compute
    coordinates), with the reference's conditions (update_dwind_phys.py:236-
    433): lower = [0, min(m2, n - 1)] when 0 <= m2 (0 < m2 for the north edge,
    lower_strict), upper = [max(m2 + 1, 0), n - 1] when n - 1 > m2 (>= for the
    north edge)."""
    lower = range(0, min(m2, n - 1) + 1) if (0 < m2 if lower_strict else 0 <= m2) else range(0)
    upper = range(max(m2 + 1, 0), n) if ((n - 1 >= m2) if lower_strict else (n - 1 > m2)) else range(0)
    return lower, upper


def agrid_to_dgrid(u, v, u_dt, v_dt, grid, dt_atmos, npx, npy, grid_type=0):
    """update_dwind_phys.py:490-701.  Returns (u, v, u_dt, v_dt)."""
    u, v, u_dt, v_dt = u.copy(), v.copy(), u_dt.copy(), v_dt.copy()
    dt5 = 0.5 * dt_atmos
    ni, nj = u.shape[0], v.shape[1]
    if grid_type > 3:
        # compute domain only; u_dt[0, -1] is the point south of (i, j)
        u[:, :nj] = u[:, :nj] + dt5 * (u_dt[1:ni + 1, 0:nj] + u_dt[1:ni + 1, 1:nj + 1])
        v[:ni, :] = v[:ni, :] + dt5 * (v_dt[0:ni, 1:nj + 1] + v_dt[1:ni + 1, 1:nj + 1])
        return u, v, u_dt, v_dt
    v3 = [u_dt * grid["vlon"][c][:, :, None] + v_dt * grid["vlat"][c][:, :, None] for c in range(3)]
    # ue at (i, j), j in [0, nj]: the A points (i, j - 1) and (i, j)
    ue = [x[1:ni + 1, 0:nj + 1] + x[1:ni + 1, 1:nj + 2] for x in v3]
    # ve at (i, j), i in [0, ni]: the A points (i - 1, j) and (i, j)
    ve = [x[0:ni + 1, 1:nj + 1] + x[1:ni + 2, 1:nj + 1] for x in v3]
    u_dt[:] = 0.0
    v_dt[:] = 0.0
    im2 = (npx - 1) // 2 + 2 - N_HALO
    jm2 = (npy - 1) // 2 + 2 - N_HALO
    ue0 = [x.copy() for x in ue]
    ve0 = [x.copy() for x in ve]
    # west (i = 0) and east (i = ni) edges: ve blended along j
    for i_edge, ev in ((0, grid["edge_vect_w"]), (ni, grid["edge_vect_e"])):
        lower, upper = _ranges(nj, jm2, False)
        for c in range(3):
            for j in lower:
                ve[c][i_edge, j] = ev[j] * ve0[c][i_edge, j + 1] + (1.0 - ev[j]) * ve0[c][i_edge, j]
            for j in upper:
                ve[c][i_edge, j] = ev[j] * ve0[c][i_edge, j - 1] + (1.0 - ev[j]) * ve0[c][i_edge, j]
    # south (j = 0) and north (j = nj) edges: ue blended along i
    for j_edge, ev, strict in ((0, grid["edge_vect_s"], False), (nj, grid["edge_vect_n"], True)):
        lower, upper = _ranges(ni, im2, strict)
        for c in range(3):
            for i in lower:
                ue[c][i, j_edge] = ev[i] * ue0[c][i + 1, j_edge] + (1.0 - ev[i]) * ue0[c][i, j_edge]
            for i in upper:
                ue[c][i, j_edge] = ev[i] * ue0[c][i - 1, j_edge] + (1.0 - ev[i]) * ue0[c][i, j_edge]
    es, ew = grid["es1"], grid["ew2"]
    u = u + dt5 * (ue[0] * es[0][:, :, None] + ue[1] * es[1][:, :, None] + ue[2] * es[2][:, :, None])
    v = v + dt5 * (ve[0] * ew[0][:, :, None] + ve[1] * ew[1][:, :, None] + ve[2] * ew[2][:, :, None])
    return u, v, u_dt, v_dt
