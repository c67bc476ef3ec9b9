"""Checks that pin the oracle's process functions by what the physics fixes.

Each ``check_*(M)`` takes the oracle module ``oracle.mp_numpy`` (or a mutated
copy of it, see tests/test_oracle_pins.py) and raises AssertionError when the
oracle disagrees with an independent statement of the same physics.  None of
them retypes the oracle's expressions: the rates are rebuilt from the
per-particle laws of the GFDL scheme's sources, integrated numerically over
the inverse-exponential (Marshall-Palmer) size distributions, and the
point-wise processes are checked against an energy invariant written from
the scheme's thermodynamics.

Sources of the physics (the reference's process library is
pySHiELD/functions/microphysics_funcs.py; its coefficients are set up in
pySHiELD/stencils/microphysics.py:1813-1954):
* Lin, Farley and Orville (1983), J. Clim. Appl. Meteor. 22: size
  distributions N(D) = n0 exp(-lambda D) with rho q = pi rho_x n0 / lambda^4,
  fall laws V = a D^b (rho0/rho)^1/2, collection integrals (their eqs. 19-30),
  vapour growth with ventilation 0.78 + 0.31 Sc^1/3 Re^1/2 (eqs. 31, 52),
  snow/graupel melting (eqs. 32, 47), Bigg freezing of rain (eq. 45).
* Manton and Cotton (1977): autoconversion 0.104 g E rho^4/3 q^7/3 /
  (mu (N rho_w)^1/3), with Klein's threshold of N droplets of radius r.
* Hong, Dudhia and Chen (2004, eq. 9) with Dudhia (1989, B7/B8): cloud-ice
  deposition 4 D_i n_i (S_i - 1) / (A + B), n_i = 5.38e7 (rho q_i)^0.75,
  D_i = 11.9 (rho q_i / n_i)^0.5, K = 0.0243, D_v = 2.26e-5.
* Deng and Mace (2008): ice fall speed
  log10 V[cm/s] = log10(IWC[g/m3]) (a tc^2 + b tc + c) + d tc + e.

Not collected by pytest (no test_ prefix); only tests/ import it.
"""

import math

import numpy as np
from numpy.polynomial.laguerre import laggauss
from scipy import integrate
from scipy.special import gamma

# ---------------------------------------------------------------- physics
# size-distribution intercepts (m^-4) and bulk densities (kg m^-3), Lin et al.
N0R, N0S, N0G = 8.0e6, 3.0e6, 4.0e6
RHO_W, RHO_S, RHO_G = 1000.0, 100.0, 400.0
RHO0 = 1.2                    # surface air density of the fall laws
# transport coefficients and latent heats of the Lin et al. table
K_A = 2.36e-2                 # thermal conductivity of air, J / (m s K)
D_V = 2.11e-5                 # diffusivity of water vapour, m^2 / s
NU = 1.259e-5                 # kinematic viscosity of air, m^2 / s
L_V, L_S, L_F = 2.5e6, 2.8336e6, 3.336e5
C_W = 4.1855e3                # heat capacity of liquid water
R_V = 461.5
# fall laws V = a D^b: rain (alin), snow (clin) and graupel; graupel falls
# with a = 45 (the speed) and is collected/ventilated with 40.74 (rho0/rho)^1/2
B_R, B_S, B_G = 0.8, 0.25, 0.5
A_G_FALL = 45.0
A_G_COLL = 40.74
# tolerance for coefficients the reference builds from 7-digit Gamma values
GAMMA_RTOL = 2.0e-5


def lam(n0, rho_x, q, den):
    """Slope of the inverse-exponential distribution holding q kg/kg."""
    return (math.pi * rho_x * n0 / (den * q)) ** 0.25


def dist_integral(f, lmb):
    """integral_0^inf f(D) exp(-lambda D) dD by adaptive quadrature."""
    val, err = integrate.quad(lambda x: f(x / lmb) * math.exp(-x), 0.0, np.inf, limit=200,
                              epsabs=0.0, epsrel=1e-12)
    return val / lmb


def mass_weighted_speed(a, b, lmb, rho_factor):
    """integral m(D) V(D) N(D) dD / integral m(D) N(D) dD with m ~ D^3."""
    num = dist_integral(lambda D: D ** 3 * a * D ** b, lmb)
    den = dist_integral(lambda D: D ** 3, lmb)
    return rho_factor * num / den


def ventilation_integral(n0, lmb, a, b, rho_factor):
    """n0 integral D (0.78 + 0.31 Sc^1/3 Re^1/2) exp(-lambda D) dD with
    Re = V(D) D / nu and V = rho_factor a D^b."""
    sc3 = (NU / D_V) ** (1.0 / 3.0)
    return n0 * dist_integral(
        lambda D: D * (0.78 + 0.31 * sc3 * math.sqrt(rho_factor * a * D ** b * D / NU)), lmb)


def setup(M, dt=150.0, **overrides):
    from pyshield_amd._config import PhysicsConfig
    nml = dict(PhysicsConfig().mp_namelist(), **overrides)
    d = M.derived(nml, dt)
    return nml, d, M.Proc(M.Ops, nml, d)


def a1(x):
    return np.atleast_1d(np.asarray(x, dtype=np.float64))


def rel(a, b):
    return abs(float(a) - float(b)) / max(abs(float(b)), 1e-300)


# --------------------------------------------------------------- constants
def check_distribution_constants(M):
    """NORM = pi rho n0 and VCON = a Gamma(4 + b) / 6 (the mass-weighted mean
    of a D^b over D^3 exp(-lambda D), times lambda^b)."""
    nml, d, P = setup(M)
    for norm, n0, rho in ((M.NORMR, N0R, RHO_W), (M.NORMS, N0S, RHO_S), (M.NORMG, N0G, RHO_G)):
        assert rel(norm, math.pi * rho * n0) < 1e-13, (norm, math.pi * rho * n0)
    for vcon, a, b in ((M.VCONR, nml["alin"], B_R), (M.VCONS, nml["clin"], B_S), (M.VCONG, A_G_FALL, B_G)):
        assert rel(vcon, a * gamma(4.0 + b) / 6.0) < GAMMA_RTOL, (vcon, a * gamma(4.0 + b) / 6.0)


def check_fall_speeds(M):
    """Rain, snow and graupel fall speeds are the mass-weighted Marshall-Palmer
    fall speeds (quadrature), with the air-density factor capped at 10 and
    the namelist limits; ice follows Deng and Mace (2008) in base 10."""
    nml, d, P = setup(M)
    for q, den in ((1e-4, 1.1), (3e-3, 0.4), (2e-6, 0.9), (5e-4, 0.05)):
        rf = math.sqrt(min(10.0, RHO0 / den))
        vr = P.rain_fspeed(a1(False), a1(q), a1(den))[0]
        want = mass_weighted_speed(nml["alin"], B_R, lam(N0R, RHO_W, q, den), rf)
        want = min(nml["vr_max"], max(M.VR_MIN, want))
        assert rel(vr, want) < GAMMA_RTOL, ("rain", q, den, vr, want)
        vtg, vti, vts = P.fall_speed(a1(q), a1(q), a1(q), a1(250.0), a1(den))
        want = min(nml["vs_max"], max(M.VF_MIN, mass_weighted_speed(nml["clin"], B_S, lam(N0S, RHO_S, q, den), rf)))
        assert rel(vts[0], want) < GAMMA_RTOL, ("snow", q, den, vts[0], want)
        want = min(nml["vg_max"], max(M.VF_MIN, mass_weighted_speed(A_G_FALL, B_G, lam(N0G, RHO_G, q, den), rf)))
        assert rel(vtg[0], want) < GAMMA_RTOL, ("graupel", q, den, vtg[0], want)
    # Deng & Mace: ice water content in g/m^3, the speed in cm/s scaled to m/s
    # by vi_fac * 0.01, times 0.8
    ai, bi, ci, di, ei = -4.14122e-5, -0.00538922, -0.0516344, 0.00216078, 1.9714
    for qi, tk, den in ((1e-5, 230.0, 0.5), (2e-4, 255.0, 0.8), (3e-6, 210.0, 0.3)):
        tc = tk - nml["tice"]
        v_cm = 10.0 ** (math.log10(1000.0 * qi * den) * (ai * tc * tc + bi * tc + ci) + di * tc + ei)
        want = min(nml["vi_max"], max(M.VF_MIN, 0.8 * 0.01 * nml["vi_fac"] * v_cm))
        vti = P.fall_speed(a1(0.0), a1(qi), a1(0.0), a1(tk), a1(den))[1][0]
        assert rel(vti, want) < 1e-12, ("ice", qi, tk, vti, want)
    # below the thresholds every species falls at the minimum speed, and a
    # column without rain falls at VF_MIN
    vtg, vti, vts = P.fall_speed(a1(5e-9), a1(5e-9), a1(5e-9), a1(250.0), a1(1.0))
    assert vtg[0] == vti[0] == vts[0] == M.VF_MIN
    assert P.rain_fspeed(a1(False), a1(5e-9), a1(1.0))[0] == M.VR_MIN
    assert P.rain_fspeed(a1(True), a1(1e-3), a1(1.0))[0] == M.VF_MIN
    # the caps bind for very dense precipitation
    assert P.rain_fspeed(a1(False), a1(0.5), a1(0.05))[0] == nml["vr_max"]


def collection_rate(n0_collector, n0_collected, rho_collected, lam_collector, lam_collected, dv):
    """Lin et al. collection of one precipitating species by another, per
    unit volume: integral integral pi/4 (Dx + Dy)^2 |dV| m(Dy) Nx Ny, with
    the mean fall-speed difference taken out of the integral (their eq. 19).
    Gauss-Laguerre quadrature (exact for polynomial x exponential)."""
    x, w = laggauss(24)
    Dx, Dy = x[:, None] / lam_collector, x[None, :] / lam_collected
    integrand = (math.pi / 4.0) * (Dx + Dy) ** 2 * (math.pi * rho_collected / 6.0) * Dy ** 3
    val = np.sum(w[:, None] * w[None, :] * integrand) / (lam_collector * lam_collected)
    return n0_collector * n0_collected * abs(dv) * val


def check_acr3d(M):
    """acr3d * rho is the collection integral for the four pairs it is used
    for (snow by rain, rain by snow, rain by graupel, snow by graupel), with
    the reference's coefficient and acco entries; c_pgacs scales the last."""
    nml, d, P = setup(M)
    acco = d["acco"]
    # (coefficient, acco column, collected (n0, rho), collector (n0, rho), efficiency)
    pairs = ((d["cracs"], 0, (N0S, RHO_S), (N0R, RHO_W), 1.0),
             (d["csacr"], 1, (N0R, RHO_W), (N0S, RHO_S), 1.0),
             (d["cgacr"], 2, (N0R, RHO_W), (N0G, RHO_G), 1.0),
             (d["cgacs"], 3, (N0S, RHO_S), (N0G, RHO_G), nml["c_pgacs"]))
    for c, k, (n0y, rhoy), (n0x, rhox), eff in pairs:
        for qy, qx, den, v1, v2 in ((1e-4, 2e-4, 1.0, 5.0, 1.0), (3e-6, 1e-3, 0.6, 0.8, 7.5),
                                    (2e-3, 5e-5, 0.3, 2.0, 2.5)):
            got = P.acr3d(a1(v1), a1(v2), a1(qy), a1(qx), c, acco[0, k], acco[1, k], acco[2, k], a1(den))[0]
            want = eff * collection_rate(n0x, n0y, rhoy, lam(n0x, rhox, qx, den), lam(n0y, rhoy, qy, den),
                                         v1 - v2) / den
            assert rel(got, want) < 1e-12, (k, qy, qx, got, want)


def check_continuous_collection(M):
    """Collection of cloud water (and ice) by a precipitating species: per
    unit cloud mixing ratio the rate is pi/4 E n0 integral D^2 V(D) N dD
    (Lin et al. eqs. 21-24); csacw/cracw/cgacw are that integral's prefactor
    of (rho q)^((3 + b) / 4), csaci and cgaci scale it by their efficiencies."""
    nml, d, P = setup(M)
    for q, den in ((1e-4, 1.0), (2e-3, 0.5), (5e-6, 0.8)):
        for coef, n0, rho, a, b, rho_fac, eff in (
                (d["csacw"], N0S, RHO_S, nml["clin"], B_S, 1.0, 1.0),
                (d["cracw"], N0R, RHO_W, nml["alin"], B_R, 1.0, nml["c_cracw"]),
                (d["cgacw"], N0G, RHO_G, A_G_COLL * math.sqrt(RHO0), B_G, 1.0 / math.sqrt(den), 1.0)):
            quad = (math.pi / 4.0) * eff * n0 * dist_integral(lambda D: D ** 2 * a * D ** b,
                                                                  lam(n0, rho, q, den)) * rho_fac
            formula = coef * (q * den) ** ((3.0 + b) / 4.0) * rho_fac
            assert rel(formula, quad) < GAMMA_RTOL, (coef, q, den, formula, quad)
    assert rel(d["csaci"], nml["c_psaci"] * d["csacw"]) < 1e-15
    assert rel(d["cgaci"], 0.05 * d["cgacw"]) < 1e-15


def check_melting(M):
    """smlt and gmlt are Lin et al.'s melting rates (eqs. 32, 47): heat
    conducted to the particles minus the latent heat of the vapour they lose,
    over L_f, with ventilation, plus the heat of collected liquid."""
    nml, d, P = setup(M)
    sc_fac = lambda den: math.sqrt(RHO0 / den)                          # noqa: E731
    for tc, dqs, q, den, acc in ((2.0, 1e-3, 1e-4, 1.0, 0.0), (5.0, -2e-4, 2e-3, 0.7, 3e-7),
                                 (0.5, 5e-4, 5e-6, 1.1, 1e-9)):
        denfac = math.sqrt(RHO0 / den)
        vent = ventilation_integral(N0S, lam(N0S, RHO_S, q, den), nml["clin"], B_S, sc_fac(den))
        want = 2.0 * math.pi / L_F * (K_A * tc / den - L_V * D_V * dqs) * vent + C_W * tc / L_F * acc
        got = P.smlt(a1(tc), a1(dqs), a1(q * den), a1(acc), a1(0.0), d["csmlt"], a1(den), a1(denfac))[0]
        assert rel(got, want) < GAMMA_RTOL, ("smlt", tc, dqs, q, got, want)
        vent = ventilation_integral(N0G, lam(N0G, RHO_G, q, den), A_G_COLL, B_G, sc_fac(den))
        want = 2.0 * math.pi / L_F * (K_A * tc / den - L_V * D_V * dqs) * vent + C_W * tc / L_F * acc
        got = P.gmlt(a1(tc), a1(dqs), a1(q * den), a1(0.0), a1(acc), d["cgmlt"], a1(den))[0]
        assert rel(got, want) < GAMMA_RTOL, ("gmlt", tc, dqs, q, got, want)


def vapour_growth(n0, lmb, a, b, rho_factor, L, T, qsat, den, s_minus_1):
    """Rate (kg/kg/s) of the vapour exchange of a population, Lin et al. eq. 31:
    (1/rho) integral 2 pi D f_v(D) (S - 1) / (A + B) N dD with
    A = L^2 / (K R_v T^2) and B = 1 / (rho q_sat D_v)."""
    A = L * L / (K_A * R_V * T * T)
    B = 1.0 / (den * qsat * D_V)
    return 2.0 * math.pi * ventilation_integral(n0, lmb, a, b, rho_factor) * s_minus_1 / (A + B) / den


def check_rain_evaporation(M):
    """revap_racc: the evaporation of rain into clear subsaturated air is the
    ventilated diffusion law (rain ventilated with alin D^0.8 at sea-level
    density), the subgrid deficit is the mean of (qsat - q)^+ over the
    uniform PDF [qpz - dqh, qpz + dqh], and the accretion of cloud water is
    the implicit (backward-Euler) solution of the continuous collection."""
    nml, d, P = setup(M)
    dt = 75.0
    h_var = 0.1
    for T, den, qr, rh in ((285.0, 1.0, 1e-4, 0.6), (300.0, 1.1, 2e-5, 0.4), (265.0, 0.7, 5e-4, 0.7)):
        qsat = float(P.wqs1(a1(T), a1(den))[0])
        qv = rh * qsat
        ql0 = P.revap_racc(a1(True), dt, a1(h_var), a1(0.0), a1(0.0), a1(0.0), a1(qr), a1(0.0), a1(qv),
                           a1(T), a1(den), a1(math.sqrt(RHO0 / den)))
        _, qr1, qv1, T1 = (float(x[0]) for x in ql0)
        rate = -vapour_growth(N0R, lam(N0R, RHO_W, qr, den), nml["alin"], B_R, 1.0, L_V, T, qsat, den,
                              qv / qsat - 1.0)
        assert rel(qr - qr1, dt * rate) < GAMMA_RTOL, (T, qr, qr - qr1, dt * rate)
        assert rel(qv1 - qv, qr - qr1) < 1e-12
    # the subgrid deficit: qsat inside the PDF's range
    T, den, qr = 285.0, 1.0, 1e-4
    qsat = float(P.wqs1(a1(T), a1(den))[0])
    qv = qsat / (1.0 + 0.5 * h_var)
    dqh = h_var * qv
    mean_deficit = integrate.quad(lambda q: max(qsat - q, 0.0), qv - dqh, qv + dqh, points=[qsat],
                                  epsrel=1e-13)[0] / (2 * dqh)
    assert qv - dqh < qsat < qv + dqh
    out = P.revap_racc(a1(True), dt, a1(h_var), a1(0.0), a1(0.0), a1(0.0), a1(qr), a1(0.0), a1(qv), a1(T),
                       a1(den), a1(1.0))
    # the same deficit in a narrow PDF that lies wholly below qsat
    clear = P.revap_racc(a1(True), dt, a1(0.01), a1(0.0), a1(0.0), a1(0.0), a1(qr), a1(0.0),
                         a1(qsat - mean_deficit), a1(T), a1(den), a1(1.0))
    # same rate law, so the evaporated amounts scale with the deficits
    ev = qr - float(out[1][0])
    ev_clear = qr - float(clear[1][0])
    assert rel(ev, ev_clear) < 1e-9, (ev, ev_clear)
    # accretion of cloud water by rain in saturated air: ql / (1 + k dt)
    T, den, qr, ql = 280.0, 1.0, 2e-4, 5e-4
    qsat = float(P.wqs1(a1(T), a1(den))[0])
    qv = qsat
    k = (math.pi / 4.0) * nml["c_cracw"] * N0R * math.sqrt(RHO0 / den) * dist_integral(
        lambda D: D ** 2 * nml["alin"] * D ** B_R, lam(N0R, RHO_W, qr, den))
    out = P.revap_racc(a1(True), dt, a1(h_var), a1(0.0), a1(0.0), a1(ql), a1(qr), a1(0.0), a1(qv), a1(T),
                       a1(den), a1(math.sqrt(RHO0 / den)))
    assert rel(float(out[0][0]), ql / (1.0 + k * dt)) < GAMMA_RTOL, (float(out[0][0]), ql / (1.0 + k * dt))
    # rain does not evaporate below t_wfr, and does so in supercooled air above it
    cold = P.revap_racc(a1(True), dt, a1(h_var), a1(0.0), a1(0.0), a1(0.0), a1(1e-4), a1(0.0), a1(1e-6),
                        a1(d["t_wfr"] - 0.5), a1(0.5), a1(1.5))
    assert float(cold[1][0]) == 1e-4
    sc = P.revap_racc(a1(True), dt, a1(h_var), a1(0.0), a1(0.0), a1(0.0), a1(1e-4), a1(0.0), a1(1e-6),
                      a1(d["t_wfr"] + 5.0), a1(0.5), a1(1.5))
    assert float(sc[1][0]) < 1e-4


def check_autoconversion(M):
    """Manton-Cotton autoconversion with N_c = ccn * rho droplets per m^3 and
    Klein's threshold (N droplets of radius rthresh); the subgrid form
    multiplies the rate by the probability that the uniform PDF
    [ql - dl, ql + dl] exceeds the threshold."""
    nml, d, P = setup(M)
    g, mu, E = 9.80665, 1.717e-5, nml["c_paut"]
    T = 290.0
    for ql, den, ccn_m3 in ((1e-3, 1.0, 1e8), (3e-3, 0.8, 3e8), (5e-4, 1.1, 5e7)):
        ccn = ccn_m3 / den                       # per kg of air, as fields_init carries it
        c_praut = float(d["cpaut"] * np.power(ccn * M.RHOR, -1.0 / 3.0))
        qc = (4.0 / 3.0) * math.pi * RHO_W * nml["rthresh"] ** 3 * ccn
        rate = 0.104 * g * E * den ** (4.0 / 3.0) * ql ** (7.0 / 3.0) / (mu * (ccn_m3 * RHO_W) ** (1.0 / 3.0))
        ql1, qr1 = P.autoconv_no_subgrid(a1(ql), a1(0.0), a1(T), a1(den), a1(ccn), a1(c_praut))
        want = min(ql - qc, d["dt_rain"] * rate)
        assert rel(float(qr1[0]), want) < 1e-12, (ql, den, float(qr1[0]), want)
        # just below the threshold nothing converts
        ql1, qr1 = P.autoconv_no_subgrid(a1(0.999 * qc), a1(0.0), a1(T), a1(den), a1(ccn), a1(c_praut))
        assert float(qr1[0]) == 0.0
        for frac in (0.25, 0.5, 0.9):
            dl = 0.3 * ql
            q_thr = ql - dl + 2 * dl * (1.0 - frac)          # P(q > q_thr) = frac
            ccn_t = q_thr / ((4.0 / 3.0) * math.pi * RHO_W * nml["rthresh"] ** 3)
            cp_t = float(d["cpaut"] * np.power(ccn_t * M.RHOR, -1.0 / 3.0))
            rate_t = 0.104 * g * E * den ** (4.0 / 3.0) * ql ** (7.0 / 3.0) / (
                mu * (ccn_t * den * RHO_W) ** (1.0 / 3.0))
            ql1, qr1 = P.autoconv_subgrid(a1(ql), a1(0.0), a1(T), a1(den), a1(ccn_t), a1(cp_t), a1(dl))
            assert rel(float(qr1[0]), frac * d["dt_rain"] * rate_t) < 1e-9, (frac, float(qr1[0]))


# --------------------------------------------------- thermodynamic invariant
def energy(M, d, qv, ql, qr, qi, qs, qg, T):
    """Internal energy per unit dry mass of the scheme's thermodynamics:
    constant-volume heat capacities c_air, c_vap, C_LIQ, C_ICE with latent
    heats L_v(T) = lv00 + (c_vap - C_LIQ) T and L_f(T) = LI00 + (C_LIQ -
    C_ICE) T, i.e. E = cvm T + qv lv00 - (qi + qs + qg) LI00."""
    q_liq, q_sol = ql + qr, qi + qs + qg
    cvm = d["c_air"] + qv * d["c_vap"] + q_liq * M.C_LIQ + q_sol * M.C_ICE
    return cvm * T + qv * d["lv00"] - q_sol * M.LI00


def random_cells(M, d, n, seed):
    """Cells spanning every branch of the point-wise processes."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(170.0, 310.0, n)
    den = rng.uniform(0.05, 1.3, n)
    P = M.Proc(M.Ops, {}, d)
    qs_w = P.wqs1(np.clip(T, 200.0, 330.0), den)
    qv = rng.uniform(0.2, 1.3, n) * np.minimum(qs_w, 0.03)
    dry = rng.uniform(size=n) < 0.05
    qv[dry] = rng.uniform(0.0, 2e-7, int(dry.sum()))

    def cond():
        q = np.exp(rng.normal(np.log(1e-4), 1.5, n))
        q[rng.uniform(size=n) < 0.4] = 0.0
        return q
    ql, qr, qi, qs, qg = cond(), cond(), cond(), cond(), cond()
    return T, den, qv, ql, qr, qi, qs, qg, rng


def _energy_report(M, d, before, after, tol=1e-12):
    e0 = energy(M, d, *before)
    e1 = energy(M, d, *after)
    scale = np.abs(e0)
    bad = np.abs(e1 - e0) > tol * scale
    assert not bad.any(), f"{int(bad.sum())} cells break energy, worst {np.max(np.abs(e1 - e0) / scale):.2e}"
    w0 = sum(before[:6])
    w1 = sum(after[:6])
    assert np.all(np.abs(w1 - w0) <= 1e-14 * np.abs(w0) + 1e-22), "water not conserved per cell"


def check_energy_revap(M):
    nml, d, P = setup(M)
    T, den, qv, ql, qr, qi, qs, qg, rng = random_cells(M, d, 20000, 1)
    denfac = np.sqrt(RHO0 / den)
    h_var = rng.uniform(0.01, 0.2, T.size)
    ql1, qr1, qv1, T1 = P.revap_racc(np.ones(T.size, bool), 75.0, h_var, qg, qi, ql, qr, qs, qv, T, den, denfac)
    assert np.any(qr1 < qr) and np.any(ql1 < ql)
    _energy_report(M, d, (qv, ql, qr, qi, qs, qg, T), (qv1, ql1, qr1, qi, qs, qg, T1))


def _subgrid_inputs(M, d, nml, n, seed):
    T, den, qv, ql, qr, qi, qs, qg, rng = random_cells(M, d, n, seed)
    h_var = rng.uniform(0.01, 0.2, n)
    rh_adj = 1.0 - h_var - nml["rh_inc"]
    rh_rain = np.maximum(0.35, rh_adj - nml["rh_inr"])
    p1 = den * M.RDGAS * T
    p1[rng.uniform(size=n) < 0.02] = 50.0
    return T, den, qv, ql, qr, qi, qs, qg, h_var, rh_adj, rh_rain, p1, rng


def check_energy_subgrid(M):
    """subgrid_z_proc conserves the scheme's energy and the water in every
    cell: each transfer carries its latent heat at the current temperature."""
    for ov in ({}, {"fast_sat_adj": True}, {"do_qa": True}):
        nml, d, P = setup(M, **ov)
        T, den, qv, ql, qr, qi, qs, qg, h_var, rh_adj, rh_rain, p1, rng = _subgrid_inputs(M, d, nml, 20000, 2)
        qaz = np.zeros_like(T)
        out = P.subgrid_z_proc(h_var, rh_adj, rh_rain, qaz, qg, qi, ql, qr, qs, qv, T, den,
                               np.sqrt(RHO0 / den), p1)
        _, qg1, qi1, ql1, qr1, qs1, qv1, T1 = out
        for before, after in ((qi, qi1), (qs, qs1), (qg, qg1), (ql, ql1), (qr, qr1)):
            assert np.any(after != before)
        _energy_report(M, d, (qv, ql, qr, qi, qs, qg, T), (qv1, ql1, qr1, qi1, qs1, qg1, T1))


def _icloud_inputs(M, d, nml, P, n, seed):
    T, den, qv, ql, qr, qi, qs, qg, h_var, rh_adj, rh_rain, p1, rng = _subgrid_inputs(M, d, nml, n, seed)
    vtg, _, vts = P.fall_speed(qg, qi, qs, T, den)
    vtr = P.rain_fspeed(np.zeros(n, bool), qr, den)
    di = rng.uniform(0.0, 0.5, n) * qi
    return T, den, qv, ql, qr, qi, qs, qg, h_var, rh_adj, rh_rain, p1, vtg, vtr, vts, di


def _icloud(M, d, nml, P, args):
    T, den, qv, ql, qr, qi, qs, qg, h_var, rh_adj, rh_rain, p1, vtg, vtr, vts, di = args
    q_liq, q_sol = ql + qr, qi + qs + qg
    cvm = d["c_air"] + qv * d["c_vap"] + q_liq * M.C_LIQ + q_sol * M.C_ICE
    return P.icloud_main(h_var, rh_adj, rh_rain, np.zeros_like(T), qg, qi, ql, qr, qs, qv, T, den,
                         np.sqrt(RHO0 / den), vtg, vtr, vts, p1, di, q_liq, q_sol, cvm)


def check_energy_icloud(M):
    """icloud_main (melting, collections, freezing, then subgrid_z_proc)
    conserves energy and water in every cell."""
    nml, d, P = setup(M)
    args = _icloud_inputs(M, d, nml, P, 20000, 3)
    T, den, qv, ql, qr, qi, qs, qg = args[:8]
    _, qg1, qi1, ql1, qr1, qs1, qv1, T1 = _icloud(M, d, nml, P, args)
    _energy_report(M, d, (qv, ql, qr, qi, qs, qg, T), (qv1, ql1, qr1, qi1, qs1, qg1, T1))
    # and without the subgrid step, the melting / collection / freezing part alone
    Pn = _no_subgrid(M, nml, d)
    _, qg1, qi1, ql1, qr1, qs1, qv1, T1 = _icloud(M, d, nml, Pn, args)
    for before, after in ((qi, qi1), (qs, qs1), (qg, qg1), (ql, ql1), (qr, qr1)):
        assert np.any(after != before)
    _energy_report(M, d, (qv, ql, qr, qi, qs, qg, T), (qv1, ql1, qr1, qi1, qs1, qg1, T1))


def _no_subgrid(M, nml, d):
    """Proc whose icloud_main stops before subgrid_z_proc, to isolate the
    melting, collection and freezing processes."""
    class P(M.Proc):
        def subgrid_z_proc(self, h_var, rh_adj, rh_rain, qaz, qgz, qiz, qlz, qrz, qsz, qvz, tz, den, denfac, p1):
            return qaz, qgz, qiz, qlz, qrz, qsz, qvz, tz
    return P(M.Ops, nml, d)


def check_energy_ice_melt(M):
    """The melting of cloud ice before the fall (microphysics.py:702-746)
    conserves energy, and where heat limits it relaxes the temperature toward
    tice by the fraction fac_imlt (cvm taken after the melt)."""
    nml, d, P = setup(M, vi_fac=0.0, do_sedi_heat=False)
    nz = 6
    T = np.array([[280.0, 277.0, 275.0, 274.0, 276.0, 290.0]])
    qi = np.array([[2e-3, 5e-5, 1e-3, 3e-3, 1e-12, 2e-4]])
    ql = np.array([[0.0, 3e-3, 1e-4, 0.0, 0.0, 1e-3]])
    qv = np.full((1, nz), 5e-3)
    z = np.zeros((1, nz))
    den = np.full((1, nz), 1.0)
    dz1 = np.full((1, nz), -200.0)
    dp1 = np.full((1, nz), 2000.0)
    out = M._sedimentation(M.Ops, P, nml, d, z, qi, ql, z, z, qv, T, den, dz1, dp1, z)
    qg1, qi1, ql1, qr1, qs1, qv1, T1 = out[:7]
    _energy_report(M, d, (qv, ql, z, qi, z, z, T), (qv1, ql1, qr1, qi1, qs1, qg1, T1))
    tc, tc1 = T - nml["tice"], T1 - nml["tice"]
    cvm1 = d["c_air"] + qv1 * d["c_vap"] + (ql1 + qr1) * M.C_LIQ + qi1 * M.C_ICE
    cvm0 = d["c_air"] + qv * d["c_vap"] + ql * M.C_LIQ + qi * M.C_ICE
    limited = (qi1 > 0.0) & (qi1 < qi)
    assert limited.sum() >= 2
    np.testing.assert_allclose(tc1[limited], (tc * (1.0 - d["fac_imlt"] * cvm0 / cvm1))[limited], rtol=1e-10)
    # melted water goes to cloud water up to ql_mlt, the rest to rain
    assert np.all(ql1 <= np.maximum(ql, nml["ql_mlt"]) + 1e-18)


def check_vapour_adjustment(M):
    """The cloud water <-> vapour adjustment of supersaturated warm cells is
    one Newton step of qv(T) = qsat(T) along the latent-heat line, with the
    scheme's gradient qsat d ln(es)/dT (Clausius-Clapeyron; the ideal-gas
    -qsat/T term of the full derivative at fixed density is left out):
    (qv' - qv) (1 + L/cvm * qsat d ln es/dT) = qsat - qv, and the cell lands
    on qsat(T') up to that neglected term and the curvature 0.5 qsat'' dT^2."""
    nml, d, P = setup(M)
    n = 64
    rng = np.random.default_rng(5)
    T = rng.uniform(276.0, 305.0, n)
    den = rng.uniform(0.6, 1.2, n)
    qsat = P.wqs1(T, den)
    qv = qsat * rng.uniform(1.005, 1.05, n)
    ql = rng.uniform(1e-5, 1e-3, n)
    z = np.zeros(n)
    h_var = np.full(n, 0.05)
    rh_adj = 1.0 - h_var - nml["rh_inc"]
    out = P.subgrid_z_proc(h_var, rh_adj, np.maximum(0.35, rh_adj - nml["rh_inr"]), z, z, z, ql, z, z, qv, T,
                           den, np.sqrt(RHO0 / den), den * M.RDGAS * T)
    _, _, _, ql1, _, _, qv1, T1 = out
    assert np.all(ql1 > ql)

    def es(t):
        return P.wqs1(t, den) * M.RVGAS * t * den
    h = 1e-3
    dlnes = (np.log(es(T + h)) - np.log(es(T - h))) / (2 * h)
    L = d["lv00"] + d["d0_vap"] * T
    cvm = d["c_air"] + qv * d["c_vap"] + ql * M.C_LIQ
    g = qsat * dlnes
    np.testing.assert_allclose((qv1 - qv) * (1.0 + L / cvm * g), qsat - qv, rtol=1e-7)
    dT = T1 - T
    hh = 0.05
    q2 = (P.wqs1(T1 + hh, den) - 2 * P.wqs1(T1, den) + P.wqs1(T1 - hh, den)) / hh ** 2
    # the neglected -qsat/T of the gradient, carried through the step, and the curvature
    predicted = (qv - qsat) * (L / cvm) * (qsat / T) / (1.0 + L / cvm * g) - 0.5 * q2 * dT ** 2
    resid = qv1 - P.wqs1(T1, den)
    assert np.all(np.abs(resid - predicted) <= 0.05 * np.abs(predicted)), np.max(np.abs(resid / predicted - 1))
    assert np.all(np.abs(resid) < 0.1 * (qv - qsat))


def check_ice_deposition(M):
    """Growth of cloud ice in air supersaturated over ice (subgrid_z_proc's
    pidep) is Hong et al. (2004) eq. 9 with Dudhia's (1989) number and size."""
    nml, d, P = setup(M)
    Ls2 = (2.5e6 + 3.3358e5) ** 2
    for T, den, qi, si in ((240.0, 0.5, 1e-5, 1.05), (255.0, 0.8, 5e-5, 1.02), (225.0, 0.3, 3e-5, 1.1)):
        qsi = float(P.iqs1(a1(T), a1(den))[0])
        qv = si * qsi
        n_i = 5.38e7 * (den * qi) ** 0.75
        D_i = 11.9 * math.sqrt(den * qi / n_i)
        A = Ls2 / (0.0243 * R_V * T * T)
        B = 1.0 / (den * qsi * 2.26e-5)
        want = d["dts"] * 4.0 * D_i * n_i * (si - 1.0) / (A + B) / den
        h_var = 0.05
        rh_adj = 1.0 - h_var - nml["rh_inc"]
        out = P.subgrid_z_proc(a1(h_var), a1(rh_adj), a1(0.3), a1(0.0), a1(0.0), a1(qi), a1(0.0), a1(0.0),
                               a1(0.0), a1(qv), a1(T), a1(den), a1(math.sqrt(RHO0 / den)),
                               a1(den * M.RDGAS * T))
        got = float(out[2][0]) - qi
        assert rel(got, want) < 1e-6, (T, got, want)


def check_snow_sublimation(M):
    """Snow sublimating in air subsaturated over ice: Lin et al. eq. 31 with
    snow's ventilation at the air density (clin D^0.25 (rho0/rho)^1/2), the
    latent heat of sublimation, over dts."""
    nml, d, P = setup(M)
    for T, den, qs, si in ((250.0, 0.7, 2e-4, 0.75), (265.0, 1.0, 1e-5, 0.8), (230.0, 0.4, 1e-3, 0.9)):
        qsi = float(P.iqs1(a1(T), a1(den))[0])
        qv = si * qsi
        want = -d["dts"] * vapour_growth(N0S, lam(N0S, RHO_S, qs, den), nml["clin"], B_S, math.sqrt(RHO0 / den),
                                         L_S, T, qsi, den, si - 1.0)
        h_var = 0.05
        rh_adj = 1.0 - h_var - nml["rh_inc"]
        # no cloud water or ice: only the snow exchanges vapour (the
        # sublimation factor min(1, (T - t_sub) / 5) is 1 here)
        out = P.subgrid_z_proc(a1(h_var), a1(rh_adj), a1(0.3), a1(0.0), a1(0.0), a1(0.0), a1(0.0), a1(0.0),
                               a1(qs), a1(qv), a1(T), a1(den), a1(math.sqrt(RHO0 / den)), a1(den * M.RDGAS * T))
        got = qs - float(out[5][0])
        assert rel(got, want) < GAMMA_RTOL, (T, got, want)


def check_cloud_fraction(M):
    """The cloud fraction added by subgrid_z_proc is the share of the uniform
    total-water PDF [qpz - dq, qpz + dq] above the saturation value qstar."""
    nml, d, P = setup(M)
    T, den, h_var = 290.0, 1.0, 0.1
    z = a1(0.0)
    rh_adj = 1.0 - h_var - nml["rh_inc"]
    for sat, ql in ((1.0, 2e-6), (1.01, 1e-5), (1.03, 1e-4), (0.93, 2e-6)):
        qsw = float(P.wqs1(a1(T), a1(den))[0])
        qv = sat * qsw
        out = P.subgrid_z_proc(a1(h_var), a1(rh_adj), a1(0.3), z, z, z, a1(ql), z, z, a1(qv), a1(T), a1(den),
                               a1(1.0), a1(den * M.RDGAS * T))
        qaz, _, _, ql1, _, _, qv1, T1 = (float(x[0]) for x in out)
        # what the cloud-fraction step sees: the state after the adjustment
        lhl = d["lv00"] + d["d0_vap"] * T1
        cvm = d["c_air"] + (qv1 + ql1) * d["c_vap"]
        qpz = qv1 + ql1
        tin = T1 - lhl / cvm * ql1
        qstar = float(P.wqs1(a1(tin), a1(den))[0])
        dq = h_var * qpz
        frac = integrate.quad(lambda q: 1.0 if q > qstar else 0.0, qpz - dq, qpz + dq,
                              points=[qstar], epsrel=1e-13)[0] / (2 * dq)
        if ql1 <= nml["qc_crt"] and frac < 1.0:
            frac = 0.0
        assert abs(qaz - frac) < 1e-9, (sat, qaz, frac)


def check_rain_freezing(M):
    """Bigg freezing of rain (pgfr): Lin et al. eq. 45, the drop-volume-
    weighted rate B' (exp(A' (T0 - T)) - 1) integrated over the rain drops,
    B' = 100 m^-3 s^-1, A' = 0.66 K^-1."""
    nml, d, P = setup(M)
    Pn = _no_subgrid(M, nml, d)
    for T, den, qr in ((253.0, 1.0, 9e-7), (255.0, 0.8, 8e-7), (251.0, 1.2, 5e-7)):
        tc = T - nml["tice"]
        lmb = lam(N0R, RHO_W, qr, den)
        rate = 100.0 * (math.exp(-0.66 * tc) - 1.0) * N0R * dist_integral(
            lambda D: (math.pi * D ** 3 / 6.0) * (math.pi * RHO_W * D ** 3 / 6.0), lmb) / den
        z = a1(0.0)
        out = _icloud(M, d, nml, Pn, (a1(T), a1(den), a1(1e-4), z, a1(qr), z, z, z, a1(0.05), a1(0.8), a1(0.3),
                                      a1(den * M.RDGAS * T), a1(1.0), a1(1.0), a1(1.0), z))
        got = float(out[1][0])
        assert 1e-8 < d["dts"] * rate < qr          # neither the QRMIN floor nor qr limits it
        assert rel(got, d["dts"] * rate) < 1e-10, (T, got, d["dts"] * rate)


def check_snow_melt_and_ice_autoconversion(M):
    """Snow melting alone in warm air is dts x smlt's ventilated rate; cloud
    ice converts to snow at fac_i2s exp(0.025 tc) times the mean excess over
    qi0_crt / rho of the uniform subgrid PDF [qi - di, qi + di]."""
    nml, d, P = setup(M)
    Pn = _no_subgrid(M, nml, d)
    z = a1(0.0)
    T, den, qs = 273.5, 1.0, 2e-3
    qv = 4.9e-3
    tc = T - nml["tice"]
    dqs = d["ces0"] / (den * M.RDGAS * T) - qv
    vent = ventilation_integral(N0S, lam(N0S, RHO_S, qs, den), nml["clin"], B_S, math.sqrt(RHO0 / den))
    rate = 2.0 * math.pi / L_F * (K_A * tc / den - L_V * D_V * dqs) * vent
    out = _icloud(M, d, nml, Pn, (a1(T), a1(den), a1(qv), z, z, z, a1(qs), z, a1(0.05), a1(0.8), a1(0.3),
                                  a1(den * M.RDGAS * T), a1(1.0), a1(1.0), a1(1.0), z))
    got = qs - float(out[5][0])
    # neither the snow present nor the heat that brings the cell to tice limits it
    assert 0.0 < d["dts"] * rate < min(qs, tc * (d["c_air"] + qv * d["c_vap"] + qs * M.C_ICE)
                                       / (M.LI00 + M.DC_ICE * T))
    assert rel(got, d["dts"] * rate) < GAMMA_RTOL, (got, d["dts"] * rate)
    for qi, di in ((1e-4, 5e-5), (5e-5, 2e-5), (3e-4, 1e-4)):
        T, den = 250.0, 0.8
        tc = T - nml["tice"]
        qim = nml["qi0_crt"] / den
        excess = integrate.quad(lambda q: max(q - qim, 0.0), qi - di, qi + di, points=[qim],
                                epsrel=1e-13)[0] / (2 * di)
        want = min(0.75 * qi, d["fac_i2s"] * math.exp(0.025 * tc) * excess)
        out = _icloud(M, d, nml, Pn, (a1(T), a1(den), a1(1e-4), z, z, a1(qi), z, z, a1(0.05), a1(0.8), a1(0.3),
                                      a1(den * M.RDGAS * T), a1(1.0), a1(1.0), a1(1.0), a1(di)))
        got = float(out[5][0])
        assert abs(got - want) <= 1e-12 * want + 1e-22, (qi, di, got, want)


def check_setup_constants(M):
    """The remaining setupm scalars: the Gamma values of the rate integrals,
    the latent-heat and transport constants and ces0 = eps es(0 C)."""
    nml, d, P = setup(M)
    # ces0: saturation vapour pressure at 0 C (Murphy and Koop 2005: 611.2 Pa
    # at 273.15 K) times Rd/Rv
    assert rel(d["ces0"], M.RDGAS / M.RVGAS * 611.2) < 1.5e-3
    # snow sublimation constants: the diffusion law's pieces
    cs = d["cssub"]
    assert rel(cs[0], 2 * math.pi * D_V * K_A * R_V * N0S) < 1e-12
    assert rel(cs[3], K_A * R_V) < 1e-12 and rel(cs[4], L_S ** 2 * D_V) < 1e-12
    cr = d["crevp"]
    assert rel(cr[0], 2 * math.pi * D_V * K_A * R_V * N0R) < 1e-12
    assert rel(cr[3], K_A * R_V) < 1e-12 and rel(cr[4], L_V ** 2 * D_V) < 1e-12
    # the ventilation pieces are integrals over the distributions
    act_s, act_r = math.pi * N0S * RHO_S, math.pi * N0R * RHO_W
    assert rel(cs[1], 0.78 * dist_integral(lambda D: D, 1.0) / math.sqrt(act_s)) < 1e-12
    assert rel(cs[2], 0.31 * (NU / D_V) ** (1 / 3) * math.sqrt(nml["clin"] / NU)
               * dist_integral(lambda D: D ** ((3 + B_S) / 2), 1.0) / act_s ** ((5 + B_S) / 8)) < GAMMA_RTOL
    assert rel(cr[2], 0.31 * (NU / D_V) ** (1 / 3) * math.sqrt(nml["alin"] / NU)
               * dist_integral(lambda D: D ** ((3 + B_R) / 2), 1.0) / act_r ** ((5 + B_R) / 8)) < GAMMA_RTOL
    # melting: conduction and evaporation prefactors over L_f, the liquid's
    # heat capacity over L_f
    for c, n0 in ((d["csmlt"], N0S), (d["cgmlt"], N0G)):
        assert rel(c[0], 2 * math.pi * K_A * n0 / L_F) < 1e-12
        assert rel(c[1], 2 * math.pi * D_V * n0 * L_V / L_F) < 1e-12
        assert rel(c[4], C_W / L_F) < 1e-12
    assert d["csmlt"][2] == cs[1] and d["csmlt"][3] == cs[2]
    # rain freezing: A' = 0.66
    assert d["cgfr"][1] == 0.66


def check_snow_collects_rain(M):
    """Rain collected by snow in air just below 0 C (psacr in icloud_main's
    cold branch) is dts times the collection integral of rain drops by snow
    particles; Bigg freezing is negligible there."""
    nml, d, P = setup(M)
    Pn = _no_subgrid(M, nml, d)
    z = a1(0.0)
    for T, den, qr, qs, vtr, vts in ((272.66, 1.0, 1e-4, 1e-4, 2.0, 1.0), (272.9, 0.8, 5e-5, 3e-4, 1.1, 0.8)):
        want = d["dts"] * collection_rate(N0S, N0R, RHO_W, lam(N0S, RHO_S, qs, den), lam(N0R, RHO_W, qr, den),
                                          vtr - vts) / den
        out = _icloud(M, d, nml, Pn, (a1(T), a1(den), a1(3e-3), z, a1(qr), z, a1(qs), z, a1(0.05), a1(0.8),
                                      a1(0.3), a1(den * M.RDGAS * T), a1(1.0), a1(vtr), a1(vts), z))
        qg1, qr1, qs1 = float(out[1][0]), float(out[4][0]), float(out[5][0])
        assert want < qr                                # not limited by the rain present
        assert qg1 < 1e-3 * want                        # the frozen rain is negligible
        assert rel(qs1 - qs, want) < 1e-3, (T, qs1 - qs, want)
        assert rel(qr - qr1, qs1 - qs + qg1) < 1e-12
