"""Pins of the ApplyPhysicsToDycore restatement (oracle/dycore_numpy.py):
closed forms of moist_cv and the pressure recurrence, the edge index ranges
against the reference's halo'd-index conditions and against the Fortran
update_dwinds_phys rule, and hand-derived responses of the D-grid wind update
to a tendency at a single A point."""

import numpy as np
import pytest

from oracle import dycore_numpy as D

N_HALO = 3


def test_moist_cv_closed_forms():
    # cvm is the heat capacity of the one species present
    shape = (2, 3, 4)
    pt = np.full(shape, 280.0)
    t_dt = np.full(shape, 1e-3)
    zero = np.zeros(shape)
    names = ("qvapor", "qliquid", "qrain", "qsnow", "qice", "qgraupel")
    cases = {None: 1004.6 - 287.05, "qvapor": 3.0 * 461.5, "qliquid": 4185.5, "qrain": 4185.5,
             "qsnow": 1972.0, "qice": 1972.0, "qgraupel": 1972.0}
    for sp, cv in cases.items():
        q = {n: (np.ones(shape) if n == sp else zero) for n in names}
        out, tdt = D.moist_cv(q, pt, t_dt, 1004.6, 150.0)
        np.testing.assert_allclose(out, 280.0 + 1e-3 * 150.0 * 1004.6 / cv, rtol=1e-15)
        assert not tdt.any()


def test_pressure_recurrence_closed_form():
    ni, nj, nk = 2, 2, 6
    pe = np.zeros((ni, nj, nk + 1))
    pe[..., 0] = 300.0
    delp = np.full((ni, nj, nk), 1000.0)
    ua = np.arange(ni * nj * nk, dtype=float).reshape(ni, nj, nk)
    kappa = 287.05 / 1004.6
    pe2, peln, pk, ps, us, vs = D.pressure_and_surface_winds(pe, delp, pe.copy(), pe.copy(), kappa, ua, -ua)
    np.testing.assert_array_equal(pe2[0, 0], 300.0 + 1000.0 * np.arange(nk + 1))
    np.testing.assert_allclose(pk[..., 1:], pe2[..., 1:] ** kappa, rtol=1e-14)
    np.testing.assert_array_equal(peln[..., 0], 300.0)  # the top interface is left as it was
    np.testing.assert_array_equal(ps, 300.0 + 1000.0 * nk)
    np.testing.assert_array_equal(us, ua[..., -1])
    np.testing.assert_array_equal(vs, -ua[..., -1])


def _reference_ranges(n, edge):
    """The reference's edge ranges in its halo'd local indices (one rank per
    tile: local == global), update_dwind_phys.py:236-433, shifted to compute
    indices."""
    s, e = N_HALO, N_HALO + n - 1
    m2 = int(n / 2) + 2  # int((npx - 1) / 2) + 2 with npx = n + 1
    lower, upper = [], []
    if edge in ("w", "e", "s"):
        if s <= m2 and min(m2, e) - s + 1 > 0:
            lower = list(range(s, min(m2, e) + 1))
        if e > m2:
            upper = list(range(max(m2 + 1, s), e + 1))
    else:  # north
        if s < m2 and min(m2, e) - s + 1 > 0:
            lower = list(range(s, min(m2, e) + 1))
        if e >= m2 and e - max(m2 + 1, s) + 1 > 0:
            upper = list(range(max(m2 + 1, s), e + 1))
    return [i - N_HALO for i in lower], [i - N_HALO for i in upper]


@pytest.mark.parametrize("n", list(range(2, 41)) + [96, 384])
def test_edge_ranges_match_reference_conditions(n):
    m2 = n // 2 + 2 - N_HALO
    for edge in "wesn":
        lo, up = D._ranges(n, m2, edge == "n")
        ref_lo, ref_up = _reference_ranges(n, edge)
        assert list(lo) == ref_lo and list(up) == ref_up, (n, edge)
        # the Fortran rule: face i (1-based) blends with i + 1 when i <= im2,
        # with i - 1 above, im2 = (npx - 1) / 2; the reference's north edge
        # drops the lower range when it starts at im2 (n = 2, 3)
        f_lo = [i for i in range(n) if i + 1 <= n // 2]
        f_up = [i for i in range(n) if i + 1 > n // 2]
        assert list(up) == f_up
        if edge == "n" and n // 2 == 1:
            assert list(lo) == []
        else:
            assert list(lo) == f_lo


def _grid(ni, nj, rng, unit=False):
    """Grid metrics for a one-tile test: with ``unit`` the lon/lat vectors are
    the first two axes and the face vectors the first axis everywhere."""
    g = {}
    if unit:
        g["vlon"] = [np.full((ni + 2, nj + 2), v) for v in (1.0, 0.0, 0.0)]
        g["vlat"] = [np.full((ni + 2, nj + 2), v) for v in (0.0, 1.0, 0.0)]
        g["es1"] = [np.full((ni, nj + 1), v) for v in (1.0, 0.0, 0.0)]
        g["ew2"] = [np.full((ni + 1, nj), v) for v in (1.0, 0.0, 0.0)]
    else:
        g["vlon"] = [rng.uniform(-1, 1, (ni + 2, nj + 2)) for _ in range(3)]
        g["vlat"] = [rng.uniform(-1, 1, (ni + 2, nj + 2)) for _ in range(3)]
        g["es1"] = [rng.uniform(-1, 1, (ni, nj + 1)) for _ in range(3)]
        g["ew2"] = [rng.uniform(-1, 1, (ni + 1, nj)) for _ in range(3)]
    g["edge_vect_w"] = rng.uniform(0, 1, nj)
    g["edge_vect_e"] = rng.uniform(0, 1, nj)
    g["edge_vect_s"] = rng.uniform(0, 1, ni)
    g["edge_vect_n"] = rng.uniform(0, 1, ni)
    return g


def test_uniform_tendency_is_blended_unchanged():
    # a uniform 3D tendency vector V: every face average is 2 V, the edge
    # blends keep it, so u += dt (V . es1) and v += dt (V . ew2)
    ni, nj, nk, dt = 6, 5, 3, 150.0
    rng = np.random.default_rng(0)
    g = _grid(ni, nj, rng)
    a, b = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
    g["vlon"] = [np.full((ni + 2, nj + 2), x) for x in a]
    g["vlat"] = [np.full((ni + 2, nj + 2), x) for x in b]
    u_dt = np.full((ni + 2, nj + 2, nk), 0.7)
    v_dt = np.full((ni + 2, nj + 2, nk), -0.3)
    vec = 0.7 * a - 0.3 * b
    u0 = rng.normal(size=(ni, nj + 1, nk))
    v0 = rng.normal(size=(ni + 1, nj, nk))
    u, v, udt, vdt = D.agrid_to_dgrid(u0, v0, u_dt, v_dt, g, dt, ni + 1, nj + 1)
    du = dt * sum(vec[c] * g["es1"][c] for c in range(3))
    dv = dt * sum(vec[c] * g["ew2"][c] for c in range(3))
    np.testing.assert_allclose(u - u0, np.repeat(du[:, :, None], nk, 2), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(v - v0, np.repeat(dv[:, :, None], nk, 2), rtol=1e-12, atol=1e-13)
    assert not udt.any() and not vdt.any()


def _delta_response(ni, nj, ia, ja, g, which="u"):
    """u, v increments / dt5 for a unit tendency at A point (ia, ja), compute
    coordinates (-1 and ni / nj are the halo)."""
    nk = 2
    u_dt = np.zeros((ni + 2, nj + 2, nk))
    v_dt = np.zeros((ni + 2, nj + 2, nk))
    (u_dt if which == "u" else v_dt)[ia + 1, ja + 1, :] = 1.0
    u0 = np.zeros((ni, nj + 1, nk))
    v0 = np.zeros((ni + 1, nj, nk))
    u, v, _, _ = D.agrid_to_dgrid(u0, v0, u_dt, v_dt, g, 2.0, ni + 1, nj + 1)
    return u[..., 0], v[..., 0]


def _expect(shape, entries):
    x = np.zeros(shape)
    for idx, val in entries.items():
        x[idx] = val
    return x


def test_single_point_responses():
    # ni = nj = 8: im2 = jm2 = 3 (compute), lower ranges [0, 3], upper [4, 7].
    # Unit vectors: the A-grid u tendency is the first 3D component and the
    # face vectors pick it, so a face's increment / dt5 is its blended average.
    ni = nj = 8
    rng = np.random.default_rng(3)
    g = _grid(ni, nj, rng, unit=True)
    ws, wn, ww, we = g["edge_vect_s"], g["edge_vect_n"], g["edge_vect_w"], g["edge_vect_e"]
    us, vs = (ni, nj + 1), (ni + 1, nj)
    # interior point: the two u faces south/north and two v faces west/east
    u, v = _delta_response(ni, nj, 4, 5, g)
    np.testing.assert_allclose(u, _expect(us, {(4, 5): 1, (4, 6): 1}), atol=1e-15)
    np.testing.assert_allclose(v, _expect(vs, {(4, 5): 1, (5, 5): 1}), atol=1e-15)
    # south row, lower range: face (i, 0) takes g_i of face (i + 1, 0)
    u, v = _delta_response(ni, nj, 2, 0, g)
    np.testing.assert_allclose(u, _expect(us, {(1, 0): ws[1], (2, 0): 1 - ws[2], (2, 1): 1}), atol=1e-15)
    np.testing.assert_allclose(v, _expect(vs, {(2, 0): 1, (3, 0): 1}), atol=1e-15)
    # south row, upper range: face (i, 0) takes g_i of face (i - 1, 0)
    u, v = _delta_response(ni, nj, 5, 0, g)
    np.testing.assert_allclose(u, _expect(us, {(5, 0): 1 - ws[5], (6, 0): ws[6], (5, 1): 1}), atol=1e-15)
    # the split: face 3 (lower) reads face 4, face 4 (upper) reads face 3
    u, v = _delta_response(ni, nj, 3, 0, g)
    np.testing.assert_allclose(u, _expect(us, {(2, 0): ws[2], (3, 0): 1 - ws[3], (4, 0): ws[4], (3, 1): 1}),
                               atol=1e-15)
    # north halo row: only the north edge faces see it
    u, v = _delta_response(ni, nj, 2, nj, g)
    np.testing.assert_allclose(u, _expect(us, {(1, nj): wn[1], (2, nj): 1 - wn[2]}), atol=1e-15)
    assert not v.any()
    # west halo column (the v tendency is the second component: use the u one
    # with ew2 = first axis) -- v faces at i = 0 blend along j
    u, v = _delta_response(ni, nj, -1, 2, g)
    np.testing.assert_allclose(v, _expect(vs, {(0, 1): ww[1], (0, 2): 1 - ww[2]}), atol=1e-15)
    assert not u.any()
    # east halo column, upper range
    u, v = _delta_response(ni, nj, ni, 5, g)
    np.testing.assert_allclose(v, _expect(vs, {(ni, 5): 1 - we[5], (ni, 6): we[6]}), atol=1e-15)
    # the v tendency maps to the second 3D component, which es1 / ew2 ignore
    u, v = _delta_response(ni, nj, 4, 5, g, which="v")
    assert not u.any() and not v.any()


def test_north_edge_lower_range_dropped_on_tiny_tile():
    # n = 3: im2 = 0 (compute); the south face 0 blends with face 1, the north
    # face 0 is left as the plain average (the reference's strict condition)
    ni = nj = 3
    g = _grid(ni, nj, np.random.default_rng(5), unit=True)
    u, _ = _delta_response(ni, nj, 0, 0, g)
    assert u[0, 0] == pytest.approx(1 - g["edge_vect_s"][0])
    u, _ = _delta_response(ni, nj, 0, nj, g)
    assert u[0, nj] == pytest.approx(1.0)
    assert u[1, nj] == pytest.approx(g["edge_vect_n"][1])


def test_doubly_periodic_form():
    ni, nj, nk, dt = 5, 4, 2, 10.0
    u_dt = np.zeros((ni + 2, nj + 2, nk))
    v_dt = np.zeros((ni + 2, nj + 2, nk))
    u_dt[2 + 1, 1 + 1] = 3.0
    u_dt[2 + 1, nj - 1 + 1] = 5.0
    v_dt[1 + 1, 2 + 1] = 7.0
    u0 = np.zeros((ni, nj + 1, nk))
    v0 = np.zeros((ni + 1, nj, nk))
    u, v, udt, vdt = D.agrid_to_dgrid(u0, v0, u_dt, v_dt, {}, dt, 0, 0, grid_type=4)
    dt5 = 0.5 * dt
    eu = np.zeros_like(u0)
    eu[2, 1] = eu[2, 2] = 3.0 * dt5
    eu[2, nj - 1] = 5.0 * dt5  # the face row j = nj is outside the compute domain
    ev = np.zeros_like(v0)
    ev[1, 2] = ev[2, 2] = 7.0 * dt5
    np.testing.assert_array_equal(u, eu)
    np.testing.assert_array_equal(v, ev)
    np.testing.assert_array_equal(udt, u_dt)  # the tendencies are kept
    np.testing.assert_array_equal(vdt, v_dt)
