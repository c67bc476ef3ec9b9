"""Pins of the Physics-prologue restatement (oracle/prologue_numpy.py) by
closed forms and invariants, not by its own formulas:

* an isothermal column whose thicknesses follow the hypsometric relation:
  interface pressures are ptop + the summed dry masses, the geopotential is
  R T ln(p_s / p), the layer pressure the statein formula produces is the
  log-mean pressure (p_k+1 - p_k) / ln(p_k+1 / p_k), phil sits halfway
  between the interfaces, get_phi_fv3 rebuilds the geopotential of statein,
  and prsik ends as Exner functions (p / p00)^kappa at the top and surface;
* tracer-mass bookkeeping with condensate: q_out dm = q_in delp for all seven
  tracers, dm = delp - condensate mass, qsgs_tke divided by dm;
* the layer pressure scales with the virtual-temperature factor and is
  clamped inside its layer; vapour below 1e-10 counts as 1e-10.
"""

import numpy as np
import pytest

from oracle import prologue_numpy as PR

R, G, KAPPA, P00, ZVIR = PR.RDGAS, PR.GRAV, PR.KAPPA, PR.P00, PR.ZVIR


def _isothermal(nk=30, t=250.0, ptop=300.0, ps=1.0e5):
    p = np.geomspace(ptop, ps, nk + 1)
    delp = np.diff(p)
    delz = -(R * t / G) * np.log(p[1:] / p[:-1])
    q = {n: np.zeros(nk) for n in PR.SPECIES}
    return p, delp, delz, q, np.full(nk, t)


def test_isothermal_column_closed_forms():
    p, delp, delz, q, pt = _isothermal()
    out = PR.prologue(delz, delp, q, np.zeros_like(delp), pt, p[0])
    np.testing.assert_allclose(out["prsi"], p, rtol=1e-13)
    np.testing.assert_allclose(out["phii"], R * 250.0 * np.log(p[-1] / p), rtol=1e-12, atol=1e-9)
    # statein floors the vapour at 1e-10 in the virtual temperature
    logmean = (p[1:] - p[:-1]) / np.log(p[1:] / p[:-1]) * (1.0 + ZVIR * 1.0e-10)
    np.testing.assert_allclose(out["delp"], logmean, rtol=1e-12)
    np.testing.assert_allclose(out["phil"], 0.5 * (out["phii"][:-1] + out["phii"][1:]), rtol=1e-13)
    np.testing.assert_allclose(out["delprsi"], delp, rtol=1e-12)
    st = PR.statein(delz, delp, q, np.zeros_like(delp), pt, p[0])
    np.testing.assert_allclose(out["phii"], st[0], rtol=1e-13, atol=1e-9)
    assert out["prsik"][0] == pytest.approx((p[0] / P00) ** KAPPA, rel=1e-14)
    assert out["prsik"][-1] == pytest.approx((p[-1] / P00) ** KAPPA, rel=1e-13)
    np.testing.assert_allclose(out["prsik"][1:-1], np.log(p[1:-1]), rtol=1e-14)


def test_tracer_mass_bookkeeping():
    rng = np.random.default_rng(3)
    nk = 40
    p, delp, delz, q, pt = _isothermal(nk)
    for n in PR.SPECIES:
        q[n] = rng.uniform(0, 2e-3, nk)
    tke = rng.uniform(0, 1, nk)
    out = PR.prologue(delz, delp, q, tke, pt, p[0])
    cond = sum(q[n] * delp for n in ("qliquid", "qrain", "qice", "qsnow", "qgraupel"))
    np.testing.assert_allclose(out["dm"], delp - cond, rtol=1e-14)
    for n in PR.SPECIES:
        np.testing.assert_allclose(out[n] * out["dm"], q[n] * delp, rtol=1e-14)
    np.testing.assert_allclose(out["qsgs_tke"] * out["dm"], tke, rtol=1e-14)
    np.testing.assert_allclose(out["prsi"][-1] - p[0], out["dm"].sum(), rtol=1e-13)


def test_virtual_temperature_and_clamps():
    p, delp, delz, q, pt = _isothermal(20)
    base = PR.prologue(delz, delp, q, np.zeros_like(delp), pt, p[0])["delp"]
    qv = dict(q, qvapor=np.full(20, 0.01))
    moist = PR.prologue(delz, delp, qv, np.zeros_like(delp), pt, p[0])
    # no condensate: dm = delp, qv unchanged, and the layer pressure (unclamped
    # here) carries 1 + (Rv/Rd - 1) qv
    np.testing.assert_allclose(moist["delp"] / base, (1.0 + ZVIR * 0.01) / (1.0 + ZVIR * 1.0e-10), rtol=1e-13)
    # a layer ten times too thick for its mass sits on the clamp above it
    thick = delz.copy()
    thick[5] *= 10.0
    out = PR.prologue(thick, delp, q, np.zeros_like(delp), pt, p[0])
    assert out["delp"][5] == pytest.approx(p[5] + 0.01 * delp[5], rel=1e-14)
    # ten times too thin: the clamp below it
    thin = delz.copy()
    thin[7] *= 0.1
    out = PR.prologue(thin, delp, q, np.zeros_like(delp), pt, p[0])
    assert out["delp"][7] == pytest.approx(p[8] - 0.01 * delp[7], rel=1e-14)
