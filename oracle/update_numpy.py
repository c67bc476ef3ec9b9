"""TEST INFRASTRUCTURE -- NumPy restatement of the column-local state-update
stencils: prepare_microphysics and update_physics_state_with_tendencies
(pySHiELD/stencils/physics.py:105-194), fill_gfs_delp and
prepare_tendencies_and_update_tracers (pySHiELD/update/update_atmos_state.py
:21-92).  Arrays are [..., nk] with k = 0 at the model top; interface fields
(phii, prsi) have nk + 1 levels."""

import numpy as np

GRAV = 9.80665
RGRAV = 1.0 / GRAV
RDGAS = 287.05
RVGAS = 461.50
ZVIR = RVGAS / RDGAS - 1.0


def prepare_microphysics(phii, omga, qvapor, pt, delp):
    """physics.py:105-143 -> (dz, wmp); the ten tendencies become zero."""
    dz = (phii[..., 1:] - phii[..., :-1]) * RGRAV
    wmp = -omga * (1.0 + ZVIR * qvapor) * pt / delp * (RDGAS * RGRAV)
    return dz, wmp


def update_physics_state_with_tendencies(state, tend, dt):
    """physics.py:146-194: q_t0 + q_dt * dt for each field."""
    return [s + t * dt for s, t in zip(state, tend)]


def fill_gfs_delp(delp, q, q_min):
    """update_atmos_state.py:21-37 over the nk layers of an nk+1-level domain."""
    q = np.array(q, dtype=np.float64, copy=True)
    nk = q.shape[-1]
    for k in range(nk - 2, -1, -1):
        c = q[..., k + 1] < q_min
        q[..., k] = np.where(c, q[..., k] + (q[..., k + 1] - q_min) * delp[..., k + 1] / delp[..., k], q[..., k])
    for k in range(1, nk):
        q[..., k] = np.where(q[..., k] < q_min, q_min, q[..., k])
    for k in range(1, nk):
        c = q[..., k - 1] < 0.0
        q[..., k] = np.where(c, q[..., k] + q[..., k - 1] * (delp[..., k - 1]) / (delp[..., k]), q[..., k])
    for k in range(0, nk):
        q[..., k] = np.where(q[..., k] < 0.0, 0.0, q[..., k])
    return q


def prepare_tendencies_and_update_tracers(u_dt, v_dt, pt_dt, u_t1, v_t1, pt_t1, qv_t1, ql_t1, qr_t1,
                                          qs_t1, qi_t1, qg_t1, u_t0, v_t0, pt_t0, qv_t0, ql_t0, qr_t0,
                                          qs_t0, qi_t0, qg_t0, prsi, delp, rdt):
    """update_atmos_state.py:40-92; returns the updated (u_dt, v_dt, pt_dt,
    qv_t0, ql_t0, qr_t0, qs_t0, qi_t0, qg_t0, delp)."""
    u_dt = u_dt + (u_t1 - u_t0) * rdt
    v_dt = v_dt + (v_t1 - v_t0) * rdt
    pt_dt = pt_dt + (pt_t1 - pt_t0) * rdt
    dp = prsi[..., 1:] - prsi[..., :-1]
    qw = [dp * x for x in (qv_t1, ql_t1, qr_t1, qs_t1, qi_t1, qg_t1)]
    qt = qw[0] + qw[1] + qw[2] + qw[3] + qw[4] + qw[5]
    q_sum = qv_t0 + ql_t0 + qr_t0 + qs_t0 + qi_t0 + qg_t0
    q0 = delp * (1.0 - q_sum) + qt
    return (u_dt, v_dt, pt_dt) + tuple(x / q0 for x in qw) + (q0,)
