"""The oracle's process functions against the physics that fixes them
(tests/oracle_pins.py), and a mutation test: plausible transcription
mistakes planted in a copy of the oracle's source (an exponent, a latent
term, a threshold, an index, a coefficient) must each be caught."""

import types

import pytest

import oracle.mp_numpy as M
import oracle_pins as OP

CHECKS = sorted(n for n in dir(OP) if n.startswith("check_"))


@pytest.mark.parametrize("name", CHECKS)
def test_pin(name):
    getattr(OP, name)(M)


def mutated(old, new):
    """A copy of oracle/mp_numpy.py with one expression replaced."""
    src = open(M.__file__).read()
    assert src.count(old) == 1, (old, src.count(old))
    mod = types.ModuleType("mp_numpy_mutated")
    mod.__file__ = M.__file__
    exec(compile(src.replace(old, new), M.__file__ + " (mutated)", "exec"), mod.__dict__)
    return mod


MUTATIONS = [
    # exponents
    ("pidep exponent", "xp.exp(0.875 * xp.log(self.safe(hasi, qiz * den)))",
     "xp.exp(0.8 * xp.log(self.safe(hasi, qiz * den)))", "check_ice_deposition"),
    ("rain fall-speed exponent", "xp.exp(0.2 * xp.log(qden / NORMR))", "xp.exp(0.25 * xp.log(qden / NORMR))",
     "check_fall_speeds"),
    ("rain ventilation exponent", "* xp.exp(0.725 * xp.log(qden)))", "* xp.exp(0.75 * xp.log(qden)))",
     "check_rain_evaporation"),
    ("snow melt power", "xp.power(qsrho, 0.65625)", "xp.power(qsrho, 0.6875)", "check_melting"),
    # latent terms
    ("ice deposition without the heat of fusion", "tz = W(ci, tz + sink * (lhl + lhi) / cvm, tz)",
     "tz = W(ci, tz + sink * lhl / cvm, tz)", "check_energy_subgrid"),
    ("snow melt without its latent heat", "tz = W(sm, tz - sink * lhi / cvm, tz)", "tz = W(sm, tz, tz)",
     "check_energy_icloud"),
    ("rain evaporation with a constant latent heat", "tz = xp.where(ev, tz - evap * lhl / cvm, tz)",
     "tz = xp.where(ev, tz - evap * d[\"lv00\"] / cvm, tz)", "check_energy_revap"),
    ("condensation with the heat of sublimation", "evap = W(pos, xp.minimum(qlz, factor * dq0 / (1.0 + tcp3 * dwsdt)), dq0 / (1.0 + tcp3 * dwsdt))",
     "evap = W(pos, xp.minimum(qlz, factor * dq0 / (1.0 + tcp3 * dwsdt)), dq0 / (1.0 + tcpk * dwsdt))",
     "check_vapour_adjustment"),
    # thresholds
    ("rain evaporation gated at tice instead of t_wfr", "act = xp.logical_and(m, xp.logical_and(tz > d[\"t_wfr\"], qr > QRMIN))",
     "act = xp.logical_and(m, xp.logical_and(tz > self.n[\"tice\"], qr > QRMIN))", "check_rain_evaporation"),
    ("cloud fraction gated by qi0_crt instead of qc_crt", "q_cond > n[\"qc_crt\"]", "q_cond > n[\"qi0_crt\"]",
     "check_cloud_fraction"),
    ("ice autoconversion threshold", "qim = n[\"qi0_crt\"] / den", "qim = n[\"qs0_crt\"] / den",
     "check_snow_melt_and_ice_autoconversion"),
    # indices and coefficients
    ("psacr with the snow-by-rain acco column", "d[\"csacr\"], acco[0, 1], acco[1, 1], acco[2, 1], den), 0.0)",
     "d[\"csacr\"], acco[0, 0], acco[1, 1], acco[2, 1], den), 0.0)", "check_snow_collects_rain"),
    ("Bigg coefficient A'", "d[\"cgfr\"] = [20.0e2 * pisq * rnzr * RHOR / act[1] ** 1.75, 0.66]",
     "d[\"cgfr\"] = [20.0e2 * pisq * rnzr * RHOR / act[1] ** 1.75, 0.6]", "check_rain_freezing"),
    ("Gamma(3.25) mistyped", "gam325, gam350, gam380 = 2.54925, 3.323363, 4.694155",
     "gam325, gam350, gam380 = 2.45925, 3.323363, 4.694155", "check_continuous_collection"),
    ("snow ventilation Gamma mistyped", "gam263, gam275, gam290 = 1.456943, 1.608355, 1.827363",
     "gam263, gam275, gam290 = 1.465943, 1.608355, 1.827363", "check_snow_sublimation"),
    ("Manton-Cotton exponent", "d[\"so3\"], d[\"zs\"] = 7.0 / 3.0, 0.0", "d[\"so3\"], d[\"zs\"] = 2.0, 0.0",
     "check_autoconversion"),
    ("acco exponent", "acco[i, k] = acc[i] / (act[2 * k] ** ((6 - i) * 0.25) * act[2 * k + 1] ** ((i + 1) * 0.25))",
     "acco[i, k] = acc[i] / (act[2 * k] ** ((6 - i) * 0.25) * act[2 * k + 1] ** ((i + 2) * 0.25))", "check_acr3d"),
]


@pytest.mark.parametrize("what,old,new,check", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_pin_catches_mutation(what, old, new, check):
    bad = mutated(old, new)
    with pytest.raises(AssertionError):
        getattr(OP, check)(bad)
