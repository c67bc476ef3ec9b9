"""Pins of the state-update restatements (oracle/update_numpy.py) by what the
mathematics fixes, not by their own formulas:

* prepare_tendencies_and_update_tracers (update_atmos_state.py:40-92): the
  GFS -> FV3 conversion must keep the dry air mass of every layer and give
  each tracer the mass dp * q_t1; the tendencies are linear in rdt;
* fill_gfs_delp (update_atmos_state.py:21-37): column tracer mass is
  conserved when the column holds enough, interior levels end >= q_min and
  no level ends negative; two columns worked by hand;
* prepare_microphysics (physics.py:105-143): layer thickness of a uniform
  geopotential profile; w = -omega R T_v / (g p)-type scaling and sign;
* update_physics_state_with_tendencies (physics.py:146-194): identity at
  dt = 0 and linearity in dt.
"""

import numpy as np
import pytest

from oracle import update_numpy as U

RDGAS, RVGAS, GRAV = 287.05, 461.50, 9.80665


def _tracers(rng, shape):
    t1 = [rng.uniform(0, 2e-3, shape) for _ in range(6)]
    t0 = [rng.uniform(0, 2e-3, shape) for _ in range(6)]
    return t1, t0


def test_tracer_update_keeps_dry_mass_and_tracer_mass():
    rng = np.random.default_rng(11)
    sh = (4, 5, 30)
    t1, t0 = _tracers(rng, sh)
    prsi = np.cumsum(np.concatenate([np.full(sh[:2] + (1,), 200.0), rng.uniform(200, 3000, sh)], -1), -1)
    dp = prsi[..., 1:] - prsi[..., :-1]
    delp = dp * rng.uniform(0.98, 1.02, sh)
    zeros = np.zeros(sh)
    out = U.prepare_tendencies_and_update_tracers(zeros, zeros, zeros, zeros, zeros, zeros, *t1,
                                                  zeros, zeros, zeros, *t0, prsi, delp, 1.0)
    q_new, delp_new = out[3:9], out[9]
    # dry mass of the old FV3 state is kept: delp (1 - sum q) is invariant
    np.testing.assert_allclose(delp_new * (1.0 - sum(q_new)), delp * (1.0 - sum(t0)), rtol=1e-13)
    # each tracer carries the mass of the physics state: q_new delp_new = dp q_t1
    # (the argument order of the species is qv, ql, qr, qs, qi, qg)
    for qn, q1 in zip(q_new, t1):
        np.testing.assert_allclose(qn * delp_new, dp * q1, rtol=1e-13)


def test_tracer_update_tendencies_are_linear_in_rdt():
    rng = np.random.default_rng(12)
    sh = (3, 3, 10)
    t1, t0 = _tracers(rng, sh)
    prsi = np.cumsum(np.full(sh[:2] + (11,), 500.0), -1)
    delp = np.full(sh, 500.0)
    u0, u1 = rng.normal(0, 5, sh), rng.normal(0, 5, sh)
    v0, v1 = rng.normal(0, 5, sh), rng.normal(0, 5, sh)
    p0, p1 = rng.uniform(200, 300, sh), rng.uniform(200, 300, sh)
    base = [rng.normal(0, 1, sh) for _ in range(3)]

    def tend(rdt):
        return U.prepare_tendencies_and_update_tracers(*base, u1, v1, p1, *t1, u0, v0, p0, *t0, prsi,
                                                       delp, rdt)[:3]

    a, b = tend(1.0 / 150.0), tend(2.0 / 150.0)
    for x0, xa, xb, (hi, lo) in zip(base, a, b, ((u1, u0), (v1, v0), (p1, p0))):
        np.testing.assert_allclose(xa - x0, (hi - lo) / 150.0, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(xb - x0, 2.0 * (xa - x0), rtol=1e-12, atol=1e-14)


def test_fill_gfs_hand_worked_columns():
    # worked by hand from the stencil's four passes (q_min = 0.1):
    # [1.0, 0.0, 0.5], delp [1, 2, 4]: the top layer pays the deficit of the
    # middle one, (0.1 - 0) * 2 / 1, and the middle one is raised to q_min
    out = U.fill_gfs_delp(np.array([1.0, 2.0, 4.0]), np.array([1.0, 0.0, 0.5]), 0.1)
    np.testing.assert_allclose(out, [0.8, 0.1, 0.5], rtol=1e-15)
    # a column with negative total mass ends empty
    out = U.fill_gfs_delp(np.ones(3), np.array([0.05, 0.0, -0.1]), 0.1)
    np.testing.assert_array_equal(out, [0.0, 0.0, 0.0])
    # nothing below q_min: untouched
    q = np.array([0.3, 0.2, 0.4, 0.5])
    np.testing.assert_array_equal(U.fill_gfs_delp(np.ones(4), q, 0.1), q)


def test_fill_gfs_conserves_column_mass():
    rng = np.random.default_rng(13)
    sh = (50, 79)
    delp = rng.uniform(100, 2000, sh)
    q = rng.uniform(-2e-9, 4e-6, sh)
    q[:, 40:45] = -1e-7
    out = U.fill_gfs_delp(delp, q, 1.0e-9)
    np.testing.assert_allclose((out * delp).sum(-1), (q * delp).sum(-1), rtol=1e-12)
    assert np.all(out[:, 1:] >= 1.0e-9)
    assert np.all(out >= 0.0)


def test_prepare_microphysics_closed_forms():
    sh = (2, 3, 8)
    h = 250.0
    # geopotential falling by g h per layer downward (k = 0 at the top)
    phii = GRAV * h * np.arange(sh[-1] + 1, 0, -1, dtype=np.float64) * np.ones(sh[:2] + (1,))
    omga = np.full(sh, -2.0)
    qv = np.zeros(sh)
    pt = np.full(sh, 300.0)
    delp = np.full(sh, 1000.0)
    dz, w = U.prepare_microphysics(phii, omga, qv, pt, delp)
    np.testing.assert_allclose(dz, -h, rtol=1e-14)
    assert np.all(w > 0.0)  # rising air (omega < 0) has w > 0
    # w scales with T, 1/delp and omega, and with the virtual-temperature
    # factor 1 + (Rv/Rd - 1) qv
    _, w2 = U.prepare_microphysics(phii, 2 * omga, qv, 2 * pt, 4 * delp)
    np.testing.assert_allclose(w2, w, rtol=1e-14)
    _, wq = U.prepare_microphysics(phii, omga, np.full(sh, 0.01), pt, delp)
    np.testing.assert_allclose(wq / w, 1.0 + (RVGAS / RDGAS - 1.0) * 0.01, rtol=1e-14)
    # dimensional value: omega RT / (g p) with p -> delp
    np.testing.assert_allclose(w, 2.0 * RDGAS * 300.0 / (GRAV * 1000.0), rtol=1e-14)


@pytest.mark.parametrize("dt", [0.0, 150.0])
def test_forward_euler_identity_and_linearity(dt):
    rng = np.random.default_rng(14)
    st = [rng.uniform(0, 1, (3, 4)) for _ in range(10)]
    td = [rng.normal(0, 1e-3, (3, 4)) for _ in range(10)]
    out = U.update_physics_state_with_tendencies(st, td, dt)
    out2 = U.update_physics_state_with_tendencies(st, td, 2 * dt)
    for s, t, a, b in zip(st, td, out, out2):
        if dt == 0.0:
            np.testing.assert_array_equal(a, s)
        np.testing.assert_allclose(b - s, 2.0 * (a - s), rtol=1e-12, atol=1e-18)
