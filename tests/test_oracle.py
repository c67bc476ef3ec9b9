"""Pins of the FP64 oracle that do not depend on retyping its formulas:
textbook saturation pressures, the Clausius-Clapeyron form of the qsat
gradient, the continuity of the ice/water switch, the bidiagonal sedimentation
solve against a dense linear solve, total-water conservation, and special
cases with known answers."""

import math

import numpy as np
import pytest

import oracle
from pyshield_amd import constants as C
from pyshield_amd._config import PhysicsConfig

from helpers import (PRECIP, TEND, WATER, cold_top_state, dry_state, low_pressure_top_state,
                     run_oracle, small_state, water_imbalance)


def es_water(t, den=1.0):
    return oracle.wqs1(t, den) * C.RVGAS * t * den


def es_ice(t, den=1.0):
    return oracle.iqs1(t, den) * C.RVGAS * t * den


@pytest.mark.parametrize("t,es_ref", [(273.16, 611.655), (283.15, 1228.1), (300.0, 3536.8),
                                      (313.15, 7384.9)])
def test_saturation_pressure_over_water(t, es_ref):
    # reference values: Murphy & Koop (2005) over liquid water
    # the closed form integrates Clausius-Clapeyron with constant heat capacities
    assert es_water(t) == pytest.approx(es_ref, rel=6e-3)


@pytest.mark.parametrize("t,es_ref", [(263.15, 259.9), (253.15, 103.26), (233.15, 12.84),
                                      (213.15, 1.080)])
def test_saturation_pressure_over_ice(t, es_ref):
    # reference values: Murphy & Koop (2005) over ice
    assert es_ice(t) == pytest.approx(es_ref, rel=1.5e-2)


def test_triple_point_and_switches():
    # es(TICE) = E00 exactly on both branches (the exponent vanishes)
    assert es_water(C.TICE) == pytest.approx(C.E00, rel=1e-15)
    assert oracle.iqs1(C.TICE - 1e-9, 1.0) == pytest.approx(oracle.wqs1(C.TICE, 1.0), rel=1e-9)
    # below T_SAT_MIN iqs1 is clamped to its value at T_SAT_MIN
    assert oracle.iqs1(C.T_SAT_MIN - 5.0, 1.0) == pytest.approx(
        oracle.iqs1(C.T_SAT_MIN, 1.0) * C.T_SAT_MIN / C.T_SAT_MIN, rel=1e-12)
    # above TICE + 102 the water branch is clamped
    assert oracle.iqs1(C.TICE + 110.0, 1.0) == oracle.iqs1(C.TICE + 102.0, 1.0)
    # the saturation mixing ratio scales as 1 / density
    assert oracle.wqs1(280.0, 0.5) == pytest.approx(2.0 * oracle.wqs1(280.0, 1.0), rel=1e-15)


@pytest.mark.parametrize("t", [200.0, 240.0, 265.0, 280.0, 300.0])
def test_qsat_gradient_is_clausius_clapeyron(t):
    """wqs2/iqs2 return q * dln(es)/dT (microphysics_funcs.py:68-72,115-145):
    compare with a centred difference of ln es."""
    h = 1e-3
    for fq, fes, f2 in ((oracle.wqs1, es_water, oracle.wqs2), (oracle.iqs1, es_ice, oracle.iqs2)):
        if f2 is oracle.iqs2 and t >= C.TICE:
            continue
        q, dq = f2(t, 0.9)
        dlnes = (math.log(fes(t + h)) - math.log(fes(t - h))) / (2 * h)
        assert q == fq(t, 0.9)
        assert dq == pytest.approx(q * dlnes, rel=1e-7)


def test_implicit_fall_matches_dense_solve():
    """The upwind implicit fall solves (dz_k + dt v_k) rho_k - dt v_{k-1} rho_{k-1}
    = m_k; check it (and the fluxes dt v_k rho_k) against numpy.linalg.solve."""
    rng = np.random.default_rng(3)
    nz = 30
    dz1 = -rng.uniform(50.0, 800.0, nz)
    dp1 = rng.uniform(500.0, 3000.0, nz)
    vt = rng.uniform(0.5, 9.0, nz)
    q0 = np.where(rng.uniform(size=nz) < 0.6, rng.lognormal(np.log(1e-4), 1.0, nz), 0.0)
    dt = 75.0
    q, flux, surf = oracle.fall_column(dt, dz1, dp1, vt, q0)
    dz = -dz1
    A = np.diag(dz + dt * vt) - np.diag(dt * vt[:-1], -1)
    rho = np.linalg.solve(A, q0 * dp1)
    np.testing.assert_allclose(q, rho * dz / dp1, rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(flux, dt * vt * rho, rtol=1e-10, atol=1e-16)
    assert surf == pytest.approx(dt * vt[-1] * rho[-1], rel=1e-10)
    # mass: what stays plus what left through the surface
    assert (q * dp1).sum() + surf == pytest.approx((q0 * dp1).sum(), rel=1e-14)


NAMELISTS = [
    {},
    {"z_slope_ice": True, "irain_f": 1},
    {"de_ice": True, "fix_negative": True},
    {"const_vr": True, "const_vs": True, "const_vg": True, "const_vi": True},
    {"do_sedi_w": True, "sedi_transport": False, "do_qa": True},
    {"fast_sat_adj": True, "rad_snow": False, "rad_rain": False, "z_slope_liq": False},
    {"vi_fac": 0.0, "do_sedi_heat": False},
]


@pytest.mark.parametrize("overrides", NAMELISTS)
@pytest.mark.parametrize("dt", [150.0, 225.0])
def test_water_conservation(overrides, dt):
    """I1: sum_k dp0 sum_species q_dt dt + surface precipitation = 0 per column."""
    st = small_state(6, 5, 79, seed=31)
    cfg = PhysicsConfig(dt_atmos=int(dt), **overrides)
    out = run_oracle(st, cfg, dt)
    if overrides.get("de_ice"):
        # de_ice moves ice to snow in the input state and its tendencies first
        st = dict(st, qi=out["qi"], qs=out["qs"])
    for n in TEND:
        assert np.all(np.isfinite(out[n])), n
    assert water_imbalance(st, out, dt).max() < 1e-13


def test_dry_column_is_unchanged():
    """No condensate and vapour well below ice and water saturation: every
    water tendency is exactly zero, so is the precipitation and qaz; pt_dt is
    round-off only (the sedimentation heating divides cvn tz by cvn)."""
    st = dry_state()
    st["qa"] = np.full_like(st["qa"], 0.3)
    out = run_oracle(st, PhysicsConfig(), 150.0)
    for q in WATER:
        assert np.all(out[f"{q}_dt"] == 0.0), q
    for p in PRECIP:
        assert np.all(out[p] == 0.0), p
    assert np.all(out["qaz"] == 0.0)
    np.testing.assert_allclose(out["qa_dt"], -0.3 / 150.0, rtol=1e-15)
    assert np.max(np.abs(out["pt_dt"])) < 1e-14


def test_freeze_dry_below_t_min():
    """tz < T_MIN: sink = dim(1e-7, qv) moves vapour to ice (B16: the Python's
    argument order, so a nearly dry cell *loses* 1e-7 - qv), cloud cover 1."""
    qv_top = 1.0e-8
    st = cold_top_state(qv_top=qv_top)
    out = run_oracle(st, PhysicsConfig(), 150.0)
    omq = 1.0 - qv_top                     # dp1 / dp0
    sink = 1.0e-7 - qv_top / omq           # dim(1e-7, qv_dry)
    assert out["qv_dt"][..., 0] == pytest.approx(-sink / 150.0 * omq, rel=1e-9)
    assert out["qi_dt"][..., 0] == pytest.approx(sink / 150.0 * omq, rel=1e-9)
    assert out["qa_dt"][..., 0] == pytest.approx(1.0 / 150.0, rel=1e-12)
    assert out["pt_dt"][..., 0] > 0.0
    # with plenty of vapour the sink vanishes
    out2 = run_oracle(cold_top_state(qv_top=1e-6), PhysicsConfig(), 150.0)
    assert out2["qv_dt"][..., 0] == 0.0


def test_low_pressure_bypasses_ice_processes():
    """p1 < P_MIN: icloud_main and subgrid_z_proc do nothing (B27); the same
    supersaturated cell at normal pressure deposits vapour onto ice."""
    thin = run_oracle(low_pressure_top_state(thin=True), PhysicsConfig(), 150.0)
    assert thin["qv_dt"][..., 0] == 0.0 and thin["qi_dt"][..., 0] == 0.0
    thick = run_oracle(low_pressure_top_state(thin=False), PhysicsConfig(), 150.0)
    assert thick["qv_dt"][..., 0] < 0.0 and thick["qi_dt"][..., 0] > 0.0


def test_rejects_broken_switches():
    st = small_state(1, 1, 79)
    for bad in ({"use_ppm": True}, {"prog_ccn": True}):
        with pytest.raises(RuntimeError):
            run_oracle(st, PhysicsConfig(**bad), 150.0)


def test_columns_are_independent():
    st = small_state(4, 5, 79, seed=41)
    a = run_oracle(st, PhysicsConfig(), 150.0)
    rev = {k: v[::-1, ::-1].copy() for k, v in st.items()}
    b = run_oracle(rev, PhysicsConfig(), 150.0)
    for n in TEND + PRECIP:
        assert np.array_equal(b[n][::-1, ::-1], a[n]), n
