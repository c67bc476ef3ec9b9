"""TEST INFRASTRUCTURE -- stage-faithful array restatement of the GFDL microphysics.

A second restatement of pySHiELD's ``Microphysics.__call__``
(pySHiELD/stencils/microphysics.py:1981-2263) with its process library
(pySHiELD/functions/microphysics_funcs.py), written the way GT4Py executes
the stencils: every PARALLEL block is one vectorised statement over all cells
[ncol, nz]; every FORWARD/BACKWARD block is an explicit loop over k; every
``if`` of the reference becomes a mask (``where``).  It shares no code with
the C oracle (oracle/gfdl_mp_oracle.c) and is used to cross-check it.

The module is written against a small array namespace ``Ops`` so that the
same code runs under NumPy (the checker) and under torch float64 (an autograd
transcription of the reference column code, used by the adjoint tests).  For
autograd, arguments of log/sqrt/division on masked-out lanes are replaced by
harmless values, so unselected branches never produce NaN gradients.

Only tests/ use this module.
"""

import math

import numpy as np

# ndsl.constants (GFS) and microphysics_funcs.py:7-42
GRAV = 9.80665
RGRAV = 1.0 / GRAV
RDGAS = 287.05
RVGAS = 461.50
HLV = 2.5e6
HLF = 3.3358e5
CP_AIR = 1004.6
TICE = 273.16
C_LIQ = 4185.5
C_ICE = 1972.0
E00 = 611.21
T_MIN = 178.0
CV_AIR = CP_AIR - RDGAS
CV_VAP = 3.0 * RVGAS
CP_VAP = 4.0 * RVGAS
EPS = RDGAS / RVGAS
DC_VAP = CP_VAP - C_LIQ
DC_ICE = C_LIQ - C_ICE
D2ICE = DC_VAP + DC_ICE
LV0 = HLV - DC_VAP * TICE
LI00 = HLF - DC_ICE * TICE
LI2 = LV0 + LI00
T_SAT_MIN = TICE - 160.0
LAT2 = (HLV + HLF) ** 2
VCONS, VCONG = 6.6280504, 87.2382675
NORMS, NORMG = 942477796.076938, 5026548245.74367
VCONR, NORMR = 2503.23638966667, 25132741228.7183
THR = THI = THG = THS = 1.0e-8
AA, BB, CC, DD_FS, EE = -4.14122e-5, -0.00538922, -0.0516344, 0.00216078, 1.9714
VR_MIN, VF_MIN = 1.0e-3, 1.0e-5
P_MIN, DT_FR, SFCRHO = 100.0, 8.0, 1.2
RHOS, RHOG, RHOR = 1.0e2, 4.0e2, 1.0e3
QCMIN, QRMIN, QVMIN = 1.0e-12, 1.0e-8, 1.0e-20


class Ops:
    """NumPy array namespace."""
    exp = staticmethod(np.exp)
    log = staticmethod(np.log)
    sqrt = staticmethod(np.sqrt)
    abs = staticmethod(np.abs)
    minimum = staticmethod(np.minimum)
    maximum = staticmethod(np.maximum)
    where = staticmethod(np.where)

    @staticmethod
    def full_like(x, v):
        return np.full_like(x, v)

    @staticmethod
    def power(x, y):
        return np.power(x, y)

    @staticmethod
    def stack_k(cols):
        return np.stack(cols, axis=-1)

    @staticmethod
    def logical_and(a, b):
        return np.logical_and(a, b)

    @staticmethod
    def logical_or(a, b):
        return np.logical_or(a, b)

    @staticmethod
    def logical_not(a):
        return np.logical_not(a)


class TorchOps:
    """torch namespace (float64 tensors, autograd-safe where-masking).

    minimum, maximum and abs select an operand, as the reference's
    ``a < b ? a : b`` does, so a gradient follows the branch the forward
    takes: at a tie it goes to the second operand (torch.minimum would split
    it), and abs at 0 passes it unchanged.  Values are those of the numpy
    namespace."""

    def __init__(self):
        import torch
        self.t = torch

    def exp(self, x):
        return self.t.exp(x)

    def log(self, x):
        return self.t.log(x)

    def sqrt(self, x):
        return self.t.sqrt(x)

    def abs(self, x):
        return self.t.where(x < 0.0, -x, x)

    def minimum(self, a, b):
        a, b = self._b(a, b)
        return self.t.where(a < b, a, b)

    def maximum(self, a, b):
        a, b = self._b(a, b)
        return self.t.where(a > b, a, b)

    def where(self, c, a, b):
        t = self.t
        if any(isinstance(x, bool) or (isinstance(x, t.Tensor) and x.dtype == t.bool) for x in (a, b)):
            a, b = self._bb(a, b)
            if not isinstance(a, t.Tensor):
                a = t.full_like(c, bool(a))
            if not isinstance(b, t.Tensor):
                b = t.full_like(c, bool(b))
            return t.where(c, a, b)
        a, b = self._b(a, b, like=c)
        return t.where(c, a, b)

    def full_like(self, x, v):
        # a tensor fill value (a differentiated parameter) keeps its graph
        if isinstance(v, self.t.Tensor):
            return self.t.zeros_like(x) + v
        return self.t.full_like(x, v)

    def power(self, x, y):
        return self.t.pow(x, y)

    def stack_k(self, cols):
        return self.t.stack(cols, dim=-1)

    def logical_and(self, a, b):
        return self.t.logical_and(*self._bb(a, b))

    def logical_or(self, a, b):
        return self.t.logical_or(*self._bb(a, b))

    def logical_not(self, a):
        return self.t.logical_not(a)

    def _bb(self, a, b):
        t = self.t
        if not isinstance(a, t.Tensor):
            a = t.as_tensor(bool(a), device=b.device)
        if not isinstance(b, t.Tensor):
            b = t.as_tensor(bool(b), device=a.device)
        return a, b

    def _b(self, a, b, like=None):
        t = self.t
        ref = a if isinstance(a, t.Tensor) else (b if isinstance(b, t.Tensor) else like)
        if not isinstance(a, t.Tensor):
            a = t.full_like(ref, float(a), dtype=t.float64)
        if not isinstance(b, t.Tensor):
            b = t.full_like(ref, float(b), dtype=t.float64)
        return a, b


# ----------------------------------------------------------------- set-up
def derived(nml, timestep):
    """Microphysics.__init__/setupm/_set_timestep (microphysics.py:1643-1979)."""
    d = {}
    d["c_air"], d["c_vap"] = CP_AIR, CP_VAP
    d["d0_vap"] = d["c_vap"] - C_LIQ
    d["lv00"] = HLV - d["d0_vap"] * TICE
    d["do_sedi_w"] = False if nml["hydrostatic"] else bool(nml["do_sedi_w"])
    d["cpaut"] = nml["c_paut"] * 0.104 * GRAV / 1.717e-5
    d["use_ccn"] = not nml["prog_ccn"]
    d["so3"], d["zs"] = 7.0 / 3.0, 0.0
    d["log_10"] = float(np.log(10.0))
    d["tice0"], d["t_wfr"] = nml["tice"] - 0.01, nml["tice"] - 40.0
    gam263, gam275, gam290 = 1.456943, 1.608355, 1.827363
    gam325, gam350, gam380 = 2.54925, 3.323363, 4.694155
    rnzs, rnzr, rnzg = 3.0e6, 8.0e6, 4.0e6
    acc = np.array([5.0, 2.0, 0.5])
    pie = 4.0 * float(np.arctan(1.0))
    d["fac_rc"] = (4.0 / 3.0) * pie * RHOR * nml["rthresh"] ** 3
    vdifu, tcond, visk = 2.11e-5, 2.36e-2, 1.259e-5
    hlts, hltc, hltf, ch2o = 2.8336e6, 2.5e6, 3.336e5, 4.1855e3
    pisq = pie * pie
    scm3 = (visk / vdifu) ** (1.0 / 3.0)
    d["cracs"] = pisq * rnzr * rnzs * RHOS
    d["csacr"] = pisq * rnzr * rnzs * RHOR
    d["cgacr"] = pisq * rnzr * rnzg * RHOR
    d["cgacs"] = pisq * rnzg * rnzs * RHOS * nml["c_pgacs"]
    act = np.empty(8)
    act[0], act[1], act[5] = pie * rnzs * RHOS, pie * rnzr * RHOR, pie * rnzg * RHOG
    act[2], act[3], act[4], act[6], act[7] = act[1], act[0], act[1], act[0], act[5]
    acco = np.empty((3, 4))
    for i in range(3):
        for k in range(4):
            acco[i, k] = acc[i] / (act[2 * k] ** ((6 - i) * 0.25) * act[2 * k + 1] ** ((i + 1) * 0.25))
    d["acco"] = acco
    gcon = 40.74 * float(np.sqrt(SFCRHO))
    d["csacw"] = pie * rnzs * nml["clin"] * gam325 / (4.0 * act[0] ** 0.8125)
    craci = pie * rnzr * nml["alin"] * gam380 / (4.0 * act[1] ** 0.95)
    d["csaci"] = d["csacw"] * nml["c_psaci"]
    d["cgacw"] = pie * rnzg * gam350 * gcon / (4.0 * act[5] ** 0.875)
    d["cgaci"] = d["cgacw"] * 0.05
    d["cracw"] = nml["c_cracw"] * craci
    cssub = [2.0 * pie * vdifu * tcond * RVGAS * rnzs, 0.78 / np.sqrt(act[0]),
             0.31 * scm3 * gam263 * np.sqrt(nml["clin"] / visk) / act[0] ** 0.65625,
             tcond * RVGAS, (hlts ** 2) * vdifu]
    cgsub1 = 0.78 / np.sqrt(act[5])
    cgsub2 = 0.31 * scm3 * gam275 * np.sqrt(gcon / visk) / act[5] ** 0.6875
    d["cssub"] = [float(x) for x in cssub]
    d["crevp"] = [2.0 * pie * vdifu * tcond * RVGAS * rnzr, float(0.78 / np.sqrt(act[1])),
                  float(0.31 * scm3 * gam290 * np.sqrt(nml["alin"] / visk) / act[1] ** 0.725),
                  cssub[3], hltc ** 2 * vdifu]
    d["cgfr"] = [20.0e2 * pisq * rnzr * RHOR / act[1] ** 1.75, 0.66]
    d["csmlt"] = [2.0 * pie * tcond * rnzs / hltf, 2.0 * pie * vdifu * rnzs * hltc / hltf,
                  d["cssub"][1], d["cssub"][2], ch2o / hltf]
    d["cgmlt"] = [2.0 * pie * tcond * rnzg / hltf, 2.0 * pie * vdifu * rnzg * hltc / hltf,
                  float(cgsub1), float(cgsub2), ch2o / hltf]
    d["ces0"] = EPS * 6.107799961e2
    mpdt = min(timestep, nml["mp_time"])
    d["rdt"] = 1.0 / timestep
    d["ntimes"] = int(round(timestep / mpdt))
    d["dts"] = timestep / d["ntimes"]
    d["dt_rain"] = d["dts"] * 0.5
    d["rdts"] = 1.0 / d["dts"]
    d["dt_evap"] = 0.5 * d["dts"] if nml["fast_sat_adj"] else d["dts"]
    d["fac_i2s"] = 1.0 - math.exp(-d["dts"] / nml["tau_i2s"])
    d["fac_g2v"] = 1.0 - math.exp(-d["dts"] / nml["tau_g2v"])
    d["fac_v2g"] = 1.0 - math.exp(-d["dts"] / nml["tau_v2g"])
    d["fac_imlt"] = 1.0 - math.exp(-0.5 * d["dts"] / nml["tau_imlt"])
    d["fac_l2v"] = 1.0 - math.exp(-d["dt_evap"] / nml["tau_l2v"])
    return d


# ------------------------------------------------------- process functions
class Proc:
    """microphysics_funcs.py restated over arrays with masks."""

    def __init__(self, xp, nml, d):
        self.xp, self.n, self.d = xp, nml, d

    def dim(self, x, y):
        return self.xp.maximum(x - y, 0.0)

    def safe(self, cond, x, fill=1.0):
        return self.xp.where(cond, x, fill)

    # f:53-72
    def wqs1(self, ta, den):
        xp = self.xp
        return (E00 * xp.exp((DC_VAP * xp.log(ta / TICE) + LV0 * (ta - TICE) / (ta * TICE)) / RVGAS)) / (
            RVGAS * ta * den)

    def wqs2(self, ta, den):
        tmp = self.wqs1(ta, den)
        return tmp, tmp * (DC_VAP + LV0 / ta) / (RVGAS * ta)

    # f:76-145
    def iqs1(self, ta, den):
        xp = self.xp
        cold = ta < TICE
        tmin = ta >= T_SAT_MIN
        ta_i = self.safe(xp.logical_and(cold, tmin), ta, 250.0)
        ice = (E00 * xp.exp((D2ICE * xp.log(ta_i / TICE) + LI2 * (ta_i - TICE) / (ta_i * TICE)) / RVGAS)) / (
            RVGAS * ta_i * den)
        clamp = (E00 * math.exp((D2ICE * math.log(1.0 - 160.0 / TICE) - LI2 * 160.0 / (T_SAT_MIN * TICE))
                                / RVGAS)) / (RVGAS * T_SAT_MIN * den)
        ta_w = xp.where(ta <= TICE + 102.0, ta, TICE + 102.0)
        ta_w = self.safe(xp.logical_not(cold), ta_w, 300.0)
        water = self.wqs1(ta_w, den)
        return xp.where(cold, xp.where(tmin, ice, clamp), water)

    def iqs2(self, ta, den):
        xp = self.xp
        tmp = self.iqs1(ta, den)
        cold = ta < TICE
        tmin = ta >= T_SAT_MIN
        ta_s = self.safe(xp.logical_and(cold, tmin), ta, 250.0)
        d_ice = tmp * (D2ICE + LI2 / ta_s) / (RVGAS * ta_s)
        d_clamp = tmp * (D2ICE + LI2 / T_SAT_MIN) / (RVGAS * T_SAT_MIN)
        warm = ta <= TICE + 102.0
        ta_w = self.safe(xp.logical_and(xp.logical_not(cold), warm), ta, 300.0)
        d_w = tmp * (DC_VAP + LV0 / ta_w) / (RVGAS * ta_w)
        d_wc = tmp * (DC_VAP + LV0 / (TICE + 102.0)) / (RVGAS * (TICE + 102.0))
        return tmp, xp.where(cold, xp.where(tmin, d_ice, d_clamp), xp.where(warm, d_w, d_wc))

    # f:149-179
    def acr3d(self, v1, v2, q1, q2, c, a0, a1, a2, rho):
        xp = self.xp
        t1 = xp.sqrt(q1 * rho)
        s1 = xp.sqrt(q2 * rho)
        s2 = xp.sqrt(s1)
        return c * xp.abs(v1 - v2) * q1 * s2 * (a0 * t1 + a1 * xp.sqrt(t1) * s2 + a2 * s1)

    def smlt(self, tc, dqs, qsrho, psacw, psacr, c, rho, rhofac):
        xp = self.xp
        return (c[0] * tc / rho - c[1] * dqs) * (c[2] * xp.sqrt(qsrho) + c[3] * xp.power(qsrho, 0.65625)
                                                  * xp.sqrt(rhofac)) + c[4] * tc * (psacw + psacr)

    def gmlt(self, tc, dqs, qgrho, pgacw, pgacr, c, rho):
        xp = self.xp
        return (c[0] * tc / rho - c[1] * dqs) * (c[2] * xp.sqrt(qgrho) + c[3] * xp.power(qgrho, 0.6875)
                                                  / xp.power(rho, 0.25)) + c[4] * tc * (pgacw + pgacr)

    def cvm(self, qv, q_liq, q_sol):
        return self.d["c_air"] + qv * self.d["c_vap"] + q_liq * C_LIQ + q_sol * C_ICE

    # f:183-265
    def revap_racc(self, m, dt, h_var, qg, qi, ql, qr, qs, qv, tz, den, denfac):
        xp, d = self.xp, self.d
        act = xp.logical_and(m, xp.logical_and(tz > d["t_wfr"], qr > QRMIN))
        lhl = d["lv00"] + d["d0_vap"] * tz
        q_liq = ql + qr
        q_sol = qi + qs + qg
        cvm = self.cvm(qv, q_liq, q_sol)
        lcpk = lhl / cvm
        tin = tz - lcpk * ql
        qpz = qv + ql
        qsat, dqsdt = self.wqs2(self.safe(act, tin, 280.0), den)
        dqh = xp.maximum(ql, h_var * xp.maximum(qpz, QCMIN))
        dqh = xp.minimum(dqh, 0.2 * qpz)
        dqv = qsat - qv
        q_minus = qpz - dqh
        q_plus = qpz + dqh
        ev = xp.logical_and(act, xp.logical_and(dqv > QVMIN, qsat > q_minus))
        # the divisor is made safe wherever its branch is not taken (it is zero
        # where vapour and cloud water are), so autograd's gradient of the
        # untaken branch stays finite; the values are the reference's
        dqh_s = self.safe(xp.logical_and(ev, xp.logical_not(qsat > q_plus)), dqh)
        dq = xp.where(qsat > q_plus, qsat - qpz, 0.25 * (q_minus - qsat) ** 2 / dqh_s)
        qden = self.safe(ev, qr * den)
        t2 = tin * tin
        evap = (d["crevp"][0] * t2 * dq * (d["crevp"][1] * xp.sqrt(qden) + d["crevp"][2]
                                          * xp.exp(0.725 * xp.log(qden)))
                / (d["crevp"][3] * t2 + d["crevp"][4] * qsat * den))
        evap = xp.minimum(qr, xp.minimum(dt * evap, dqv / (1.0 + lcpk * dqsdt)))
        qr = xp.where(ev, qr - evap, qr)
        qv = xp.where(ev, qv + evap, qv)
        q_liq = xp.where(ev, q_liq - evap, q_liq)
        cvm = xp.where(ev, self.cvm(qv, q_liq, q_sol), cvm)
        tz = xp.where(ev, tz - evap * lhl / cvm, tz)
        ac = xp.logical_and(act, xp.logical_and(xp.logical_and(qr > QRMIN, ql > 1.0e-6), qsat < q_minus))
        qrd = self.safe(ac, qr * den)
        sink = dt * denfac * d["cracw"] * xp.exp(0.95 * xp.log(qrd))
        sink = sink / (1.0 + sink) * ql
        ql = xp.where(ac, ql - sink, ql)
        qr = xp.where(ac, qr + sink, qr)
        return ql, qr, qv, tz

    # f:269-341
    def fall_speed(self, qg, qi, qs, tk, den):
        xp, n, d = self.xp, self.n, self.d
        rhof = xp.sqrt(xp.minimum(10.0, SFCRHO / den))
        if n["const_vi"]:
            vti = xp.full_like(qi, n["vi_fac"])
        else:
            vi0 = 0.01 * n["vi_fac"]
            big = qi >= THI
            tc = tk - n["tice"]
            v = (3.0 + xp.log(self.safe(big, qi * den)) / d["log_10"]) * (tc * (AA * tc + BB) + CC) + DD_FS * tc + EE
            v = vi0 * xp.exp(d["log_10"] * v) * 0.8
            vti = xp.where(big, xp.minimum(n["vi_max"], xp.maximum(VF_MIN, v)), VF_MIN)
        if n["const_vs"]:
            vts = xp.full_like(qs, n["vs_fac"])
        else:
            big = qs >= THS
            v = n["vs_fac"] * VCONS * rhof * xp.exp(0.0625 * xp.log(self.safe(big, qs * den / NORMS)))
            vts = xp.where(big, xp.minimum(n["vs_max"], xp.maximum(VF_MIN, v)), VF_MIN)
        if n["const_vg"]:
            vtg = xp.full_like(qg, n["vg_fac"])
        else:
            big = qg >= THG
            v = n["vg_fac"] * VCONG * rhof * xp.sqrt(xp.sqrt(xp.sqrt(self.safe(big, qg * den / NORMG))))
            vtg = xp.where(big, xp.minimum(n["vg_max"], xp.maximum(VF_MIN, v)), VF_MIN)
        return vtg, vti, vts

    # f:344-372
    def rain_fspeed(self, no_fall, qrz, den):
        xp, n = self.xp, self.n
        if n["const_vr"]:
            v = xp.full_like(qrz, n["vr_fac"])
        else:
            big = qrz >= THR
            qden = self.safe(big, qrz * den)
            v = n["vr_fac"] * VCONR * xp.sqrt(xp.minimum(10.0, SFCRHO / den)) * xp.exp(0.2 * xp.log(qden / NORMR))
            v = xp.where(big, xp.minimum(n["vr_max"], xp.maximum(VR_MIN, v)), VR_MIN)
        return xp.where(no_fall, VF_MIN, v)

    # f:375-428
    def autoconv_no_subgrid(self, qlz, qrz, tz, den, ccn, c_praut):
        xp, d = self.xp, self.d
        qc0 = d["fac_rc"] * ccn
        qc = qc0 if d["use_ccn"] else qc0 / den
        dq = qlz - qc
        act = xp.logical_and(tz > d["t_wfr"], dq > 0.0)
        sink = xp.minimum(dq, d["dt_rain"] * c_praut * den * xp.exp(d["so3"] * xp.log(self.safe(act, qlz))))
        return xp.where(act, qlz - sink, qlz), xp.where(act, qrz + sink, qrz)

    def autoconv_subgrid(self, qlz, qrz, tz, den, ccn, c_praut, dl):
        xp, d = self.xp, self.d
        qc0 = d["fac_rc"] * ccn
        warm = tz > d["t_wfr"] + DT_FR
        dl = xp.minimum(xp.maximum(1.0e-6, dl), 0.5 * qlz)
        qc = qc0 if d["use_ccn"] else qc0 / den
        dq = 0.5 * (qlz + dl - qc)
        act = xp.logical_and(warm, dq > 0.0)
        sink = (xp.minimum(1.0, dq / self.safe(act, dl)) * d["dt_rain"] * c_praut * den
                * xp.exp(d["so3"] * xp.log(self.safe(act, qlz))))
        return xp.where(act, qlz - sink, qlz), xp.where(act, qrz + sink, qrz)

    # f:431-850
    def subgrid_z_proc(self, h_var, rh_adj, rh_rain, qaz, qgz, qiz, qlz, qrz, qsz, qvz, tz, den, denfac, p1):
        xp, n, d = self.xp, self.n, self.d
        tice, t_wfr, t_sub, dts = n["tice"], d["t_wfr"], n["t_sub"], d["dts"]
        lv00, d0_vap = d["lv00"], d["d0_vap"]
        W, AND, NOT = xp.where, xp.logical_and, xp.logical_not
        lhl = lv00 + d0_vap * tz
        lhi = LI00 + DC_ICE * tz
        q_liq = qlz + qrz
        q_sol = qiz + qsz + qgz
        cvm = self.cvm(qvz, q_liq, q_sol)
        lcpk = lhl / cvm
        icpk = lhi / cvm
        on = p1 >= P_MIN
        # freeze-dry below T_MIN (B16: dim(1e-7, qv))
        fd = AND(on, tz < T_MIN)
        sink = self.dim(1.0e-7, qvz)
        qvz = W(fd, qvz - sink, qvz)
        qiz = W(fd, qiz + sink, qiz)
        q_sol_fd = q_sol + sink
        cvm_fd = self.cvm(qvz, q_liq, q_sol_fd)
        tz = W(fd, tz + sink * (lhl + lhi) / cvm_fd, tz)
        if not n["do_qa"]:
            qaz = W(fd, qaz + 1.0, qaz)
        m = AND(on, NOT(fd))   # the else branch of the T_MIN test (entry temperature)
        lhl = lv00 + d0_vap * tz
        lhi = LI00 + DC_ICE * tz
        lcpk = lhl / cvm
        icpk = lhi / cvm
        tcpk = lcpk + icpk
        tcp3 = lcpk + icpk * xp.minimum(1.0, self.dim(tice, tz) / (tice - t_wfr))
        qpz = qvz + qlz + qiz
        tin = tz - (lhl * (qlz + qiz) + lhi * qiz) / (d["c_air"] + qpz * d["c_vap"] + qrz * C_LIQ
                                                     + (qsz + qgz) * C_ICE)
        hi = tin > t_sub + 6.0
        rh = qpz / self.iqs1(self.safe(AND(m, hi), tin, 280.0), den)
        dry = AND(AND(m, hi), rh < rh_adj)
        tz = W(dry, tin, tz)
        qvz = W(dry, qpz, qvz)
        qlz = W(dry, 0.0, qlz)
        qiz = W(dry, 0.0, qiz)
        m = AND(m, xp.logical_or(AND(hi, rh >= rh_adj), xp.logical_not(hi)))
        # cloud water <-> vapour
        qsw, dwsdt = self.wqs2(self.safe(m, tz, 280.0), den)
        dq0 = qsw - qvz
        pos = dq0 > 0.0
        factor = xp.minimum(1.0, d["fac_l2v"] * (10.0 * dq0 / qsw))
        evap = W(pos, xp.minimum(qlz, factor * dq0 / (1.0 + tcp3 * dwsdt)), dq0 / (1.0 + tcp3 * dwsdt))
        evap = W(m, evap, 0.0)
        qvz = qvz + evap
        qlz = qlz - evap
        q_liq = q_liq - evap
        cvm = W(m, self.cvm(qvz, q_liq, q_sol), cvm)
        tz = W(m, tz - evap * lhl / cvm, tz)
        lhi = LI00 + DC_ICE * tz
        icpk = lhi / cvm
        # complete freezing below t_wfr
        dtmp = t_wfr - tz
        fz = AND(m, AND(dtmp > 0.0, qlz > QCMIN))
        sink = xp.minimum(qlz, xp.minimum(qlz * dtmp * 0.125, dtmp / icpk))
        sink = W(fz, sink, 0.0)
        qlz, qiz = qlz - sink, qiz + sink
        q_liq, q_sol = q_liq - sink, q_sol + sink
        cvm = W(fz, self.cvm(qvz, q_liq, q_sol), cvm)
        tz = W(fz, tz + sink * lhi / cvm, tz)
        lhi = LI00 + DC_ICE * tz
        icpk = lhi / cvm
        # Bigg mechanism
        if n["fast_sat_adj"]:
            dt_pisub = 0.5 * dts
        else:
            dt_pisub = dts
            tc = tice - tz
            bg = AND(m, AND(qlz > QRMIN, tc > 0.0))
            sink = 3.3333e-10 * dts * (xp.exp(0.66 * W(bg, tc, 0.0)) - 1.0) * den * qlz * qlz
            sink = xp.minimum(qlz, xp.minimum(tc / icpk, sink))
            sink = W(bg, sink, 0.0)
            qlz, qiz = qlz - sink, qiz + sink
            q_liq, q_sol = q_liq - sink, q_sol + sink
            cvm = W(bg, self.cvm(qvz, q_liq, q_sol), cvm)
            tz = W(bg, tz + sink * lhi / cvm, tz)
        lhl = lv00 + d0_vap * tz
        lhi = LI00 + DC_ICE * tz
        lcpk = lhl / cvm
        icpk = lhi / cvm
        tcpk = lcpk + icpk
        # ice sublimation / deposition
        ci = AND(m, tz < tice)
        qsi, dqsdt = self.iqs2(self.safe(ci, tz, 250.0), den)
        dq = qvz - qsi
        sink = dq / (1.0 + tcpk * dqsdt)
        hasi = qiz > QRMIN
        tzs = self.safe(ci, tz, 250.0)
        pidep = W(hasi, dt_pisub * dq * 349138.78 * xp.exp(0.875 * xp.log(self.safe(hasi, qiz * den)))
                  / (qsi * den * LAT2 / (0.0243 * RVGAS * tzs ** 2) + 4.42478e4), 0.0)
        tmp = tice - tz
        qi_crt = n["qi_gen"] * xp.minimum(n["qi_lim"], 0.1 * tmp) / den
        s_dep = xp.minimum(sink, xp.minimum(xp.maximum(qi_crt - qiz, pidep), tmp / tcpk))
        pidep_sub = pidep * xp.minimum(1.0, self.dim(tz, t_sub) * 0.2)
        s_sub = xp.maximum(pidep_sub, xp.maximum(sink, -qiz))
        sink = W(ci, W(dq > 0.0, s_dep, s_sub), 0.0)
        qvz, qiz = qvz - sink, qiz + sink
        q_sol = q_sol + sink
        cvm = W(ci, self.cvm(qvz, q_liq, q_sol), cvm)
        tz = W(ci, tz + sink * (lhl + lhi) / cvm, tz)
        lhl = lv00 + d0_vap * tz
        lhi = LI00 + DC_ICE * tz
        lcpk = lhl / cvm
        icpk = lhi / cvm
        tcpk = lcpk + icpk
        # snow sublimation / deposition
        cs = AND(m, qsz > QRMIN)
        qsi, dqsdt = self.iqs2(self.safe(cs, tz, 250.0), den)
        qden = self.safe(cs, qsz * den)
        tmp = xp.exp(0.65625 * xp.log(qden))
        tsq = tz * tz
        dq = (qsi - qvz) / (1.0 + tcpk * dqsdt)
        pssub = (d["cssub"][0] * tsq * (d["cssub"][1] * xp.sqrt(qden) + d["cssub"][2] * tmp * xp.sqrt(denfac))
                 / (d["cssub"][3] * tsq + d["cssub"][4] * qsi * den))
        pssub = (qsi - qvz) * dts * pssub
        p_sub = xp.minimum(pssub * xp.minimum(1.0, self.dim(tz, t_sub) * 0.2), qsz)
        p_dep = W(tz > tice, 0.0, xp.maximum(pssub, xp.maximum(dq, (tz - tice) / tcpk)))
        pssub = W(cs, W(pssub > 0.0, p_sub, p_dep), 0.0)
        qsz, qvz = qsz - pssub, qvz + pssub
        q_sol = q_sol - pssub
        cvm = W(cs, self.cvm(qvz, q_liq, q_sol), cvm)
        tz = W(cs, tz - pssub * (lhl + lhi) / cvm, tz)
        lhl = lv00 + d0_vap * tz
        lhi = LI00 + DC_ICE * tz
        lcpk = lhl / cvm
        icpk = lhi / cvm
        tcpk = lcpk + icpk
        # graupel sublimation / deposition
        cg = AND(m, qgz > QRMIN)
        qsi, dqsdt = self.iqs2(self.safe(cg, tz, 250.0), den)
        dq = (qvz - qsi) / (1.0 + tcpk * dqsdt)
        pgsub = (qvz / qsi - 1.0) * qgz
        g_dep = W(tz > tice, 0.0, xp.minimum(xp.minimum(d["fac_v2g"] * pgsub, 0.2 * dq),
                                             xp.minimum(qlz + qrz, (tice - tz) / tcpk)))
        g_sub = xp.maximum(d["fac_g2v"] * pgsub, dq) * xp.minimum(1.0, self.dim(tz, t_sub) * 0.1)
        pgsub = W(cg, W(pgsub > 0.0, g_dep, g_sub), 0.0)
        qgz, qvz = qgz + pgsub, qvz - pgsub
        q_sol = q_sol + pgsub
        cvm = W(cg, self.cvm(qvz, q_liq, q_sol), cvm)
        tz = W(cg, tz + pgsub * (lhl + lhi) / cvm, tz)
        # minimum rain evaporation
        lhl = lv00 + d0_vap * tz
        lcpk = lhl / cvm
        cr = AND(m, qrz > QCMIN)
        qsw, dqsdt = self.wqs2(self.safe(cr, tz, 280.0), den)
        sink = W(cr, xp.minimum(qrz, self.dim(rh_rain * qsw, qvz) / (1.0 + lcpk * dqsdt)), 0.0)
        qvz, qrz = qvz + sink, qrz - sink
        q_liq = q_liq - sink
        cvm = W(cr, self.cvm(qvz, q_liq, q_sol), cvm)
        tz = W(cr, tz - sink * lhl / cvm, tz)
        lhl = lv00 + d0_vap * tz
        cvm = d["c_air"] + (qvz + q_liq + q_sol) * d["c_vap"]
        lcpk = lhl / cvm
        # cloud fraction
        if not n["do_qa"]:
            q_sol2 = qiz + qsz if n["rad_snow"] else qiz
            q_liq2 = qlz + qrz if n["rad_rain"] else qlz
            q_cond = q_liq2 + q_sol2
            qpz = qvz + q_cond
            tin = tz - (lcpk * q_cond + icpk * q_sol2)
            tin_s = self.safe(m, tin, 250.0)
            qsi = self.iqs1(tin_s, den)
            qsw = self.wqs1(self.safe(m, tin, 280.0), den)
            rqi = W(q_cond > 3.0e-6, q_sol2 / self.safe(q_cond > 3.0e-6, q_cond), (tice - tin) / (tice - t_wfr))
            qstar = W(tin <= t_wfr, qsi, W(tin >= tice, qsw, rqi * qsi + (1.0 - rqi) * qsw))
            dq = xp.maximum(QCMIN, h_var * qpz)
            q_plus = qpz + dq
            q_minus = qpz - dq
            cf = AND(m, qpz > QRMIN)
            full = AND(cf, qstar < q_minus)
            part = AND(AND(cf, xp.logical_not(qstar < q_minus)), AND(qstar < q_plus, q_cond > n["qc_crt"]))
            qaz = W(full, qaz + 1.0, W(part, qaz + (q_plus - qstar) / (dq + dq), qaz))
        return qaz, qgz, qiz, qlz, qrz, qsz, qvz, tz

    # f:853-1260
    def icloud_main(self, h_var, rh_adj, rh_rain, qaz, qgz, qiz, qlz, qrz, qsz, qvz, tz, den, denfac,
                    vtgz, vtrz, vtsz, p1, di, q_liq, q_sol, cvm):
        xp, n, d = self.xp, self.n, self.d
        W, AND, NOT = xp.where, xp.logical_and, xp.logical_not
        tice, dts, rdts = n["tice"], d["dts"], d["rdts"]
        acco = d["acco"]
        lhi = LI00 + DC_ICE * tz
        icpk = lhi / cvm
        on = p1 >= P_MIN
        pgacr = xp.full_like(tz, 0.0)
        pgacw = xp.full_like(tz, 0.0)
        tc = tz - tice
        warm = AND(on, tc >= 0.0)
        cold = AND(on, NOT(tc >= 0.0))
        # --- warm: snow melt
        dqs0 = d["ces0"] / p1 - qvz
        sm = AND(warm, qsz > QCMIN)
        qsd = self.safe(sm, qsz * den)
        factor = denfac * d["csacw"] * xp.exp(0.8125 * xp.log(qsd))
        psacw = W(qlz > QRMIN, factor / (1.0 + dts * factor) * qlz, 0.0)
        hr = qrz > QRMIN
        qr_s = self.safe(AND(sm, hr), qrz)
        qs_s = self.safe(sm, qsz)
        psacr = W(hr, xp.minimum(self.acr3d(vtsz, vtrz, qr_s, qs_s, d["csacr"], acco[0, 1], acco[1, 1],
                                            acco[2, 1], den), qrz * rdts), 0.0)
        pracs = W(hr, self.acr3d(vtrz, vtsz, qs_s, qr_s, d["cracs"], acco[0, 0], acco[1, 0], acco[2, 0], den),
                  0.0)
        psmlt = xp.maximum(0.0, self.smlt(tc, dqs0, qsd, psacw, psacr, d["csmlt"], den, denfac))
        sink = W(sm, xp.minimum(qsz, xp.minimum(dts * (psmlt + pracs), tc / icpk)), 0.0)
        qsz = qsz - sink
        tmp = xp.minimum(sink, self.dim(n["qs_mlt"], qlz))
        qlz = qlz + tmp
        qrz = qrz + sink - tmp
        q_liq, q_sol = q_liq + sink, q_sol - sink
        cvm = W(sm, self.cvm(qvz, q_liq, q_sol), cvm)
        tz = W(sm, tz - sink * lhi / cvm, tz)
        tc = W(sm, tz - tice, tc)
        lhi = W(warm, LI00 + DC_ICE * tz, lhi)
        icpk = W(warm, lhi / cvm, icpk)
        # --- warm: graupel melt
        gm = AND(warm, AND(qgz > QCMIN, tc > 0.0))
        hr = AND(gm, qrz > QRMIN)
        qg_s = self.safe(gm, qgz)
        pgacr = W(hr, xp.minimum(self.acr3d(vtgz, vtrz, self.safe(hr, qrz), qg_s, d["cgacr"], acco[0, 2],
                                            acco[1, 2], acco[2, 2], den), rdts * qrz), pgacr)
        qden = qg_s * den
        factor = d["cgacw"] * qden / xp.sqrt(den * xp.sqrt(xp.sqrt(qden)))
        pgacw = W(AND(gm, qlz > QRMIN), factor / (1.0 + dts * factor) * qlz, pgacw)
        pgmlt = dts * self.gmlt(tc, dqs0, qden, pgacw, pgacr, d["cgmlt"], den)
        pgmlt = W(gm, xp.minimum(xp.maximum(0.0, pgmlt), xp.minimum(qgz, tc / icpk)), 0.0)
        qgz, qrz = qgz - pgmlt, qrz + pgmlt
        q_liq, q_sol = q_liq + pgmlt, q_sol - pgmlt
        cvm = W(gm, self.cvm(qvz, q_liq, q_sol), cvm)
        tz = W(gm, tz - pgmlt * lhi / cvm, tz)
        # --- cold: cloud ice sinks
        ci = AND(cold, qiz > 3.0e-7)
        hs = qsz > 1.0e-7
        factor = dts * denfac * d["csaci"] * xp.exp(0.05 * tc + 0.8125 * xp.log(self.safe(AND(ci, hs), qsz * den)))
        psaci = W(hs, factor / (1.0 + factor) * qiz, 0.0)
        qim = n["qi0_crt"] / den
        tmp = d["fac_i2s"] if n["const_vi"] else d["fac_i2s"] * xp.exp(0.025 * tc)
        di = xp.maximum(di, QRMIN)
        q_plus = qiz + di
        dq = W(qim > qiz - di, (0.25 * (q_plus - qim) ** 2) / di, qiz - qim)
        psaut = W(q_plus > qim + QRMIN, tmp * dq, 0.0)
        sink = W(ci, xp.minimum(0.75 * qiz, psaci + psaut), 0.0)
        qiz, qsz = qiz - sink, qsz + sink
        ga = AND(ci, qgz > 1.0e-6)
        factor = dts * d["cgaci"] * xp.sqrt(den) * qgz
        pgaci = W(ga, factor / (1.0 + factor) * qiz, 0.0)
        qiz, qgz = qiz - pgaci, qgz + pgaci
        # --- cold: rain to snow / graupel
        tc = W(cold, tz - tice, tc)
        cr = AND(cold, AND(qrz > 1e-7, tc < 0.0))
        psacr = W(qsz > 1.0e-7, dts * self.acr3d(vtsz, vtrz, self.safe(cr, qrz), self.safe(AND(cr, qsz > 1e-7), qsz),
                                                 d["csacr"], acco[0, 1], acco[1, 1], acco[2, 1], den), 0.0)
        pgfr = (dts * d["cgfr"][0] / den * (xp.exp(-d["cgfr"][1] * W(cr, tc, 0.0)) - 1.0)
                * xp.exp(1.75 * xp.log(self.safe(cr, qrz * den))))
        sink = psacr + pgfr
        factor = xp.minimum(sink, xp.minimum(qrz, -tc / icpk)) / xp.maximum(sink, QRMIN)
        psacr = W(cr, factor * psacr, 0.0)
        pgfr = W(cr, factor * pgfr, 0.0)
        sink = psacr + pgfr
        qrz, qsz, qgz = qrz - sink, qsz + psacr, qgz + pgfr
        q_liq, q_sol = q_liq - sink, q_sol + sink
        cvm = W(cr, self.cvm(qvz, q_liq, q_sol), cvm)
        tz = W(cr, tz + sink * lhi / cvm, tz)
        lhi = W(cold, LI00 + DC_ICE * tz, lhi)
        icpk = W(cold, lhi / cvm, icpk)
        # --- cold: graupel production from snow
        sg = AND(cold, qsz > 1.0e-7)
        sink = W(qgz > QRMIN, dts * self.acr3d(vtgz, vtsz, self.safe(sg, qsz), self.safe(AND(sg, qgz > QRMIN), qgz),
                                               d["cgacs"], acco[0, 3], acco[1, 3], acco[2, 3], den), 0.0)
        qsm = n["qs0_crt"] / den
        factor = dts * 1.0e-3 * xp.exp(0.09 * (tz - tice))
        sink = W(qsz > qsm, sink + factor / (1.0 + factor) * (qsz - qsm), sink)
        sink = W(sg, xp.minimum(qsz, sink), 0.0)
        qsz, qgz = qsz - sink, qgz + sink
        # --- cold: graupel riming
        gr = AND(cold, AND(qgz > 1.0e-7, tz < d["tice0"]))
        qden = self.safe(gr, qgz * den)
        factor = dts * d["cgacw"] * qden / xp.sqrt(den * xp.sqrt(xp.sqrt(qden)))
        pw = W(qlz > 1.0e-6, factor / (1.0 + factor) * qlz, 0.0)
        pr = W(qrz > 1.0e-6, xp.minimum(dts * self.acr3d(vtgz, vtrz, self.safe(AND(gr, qrz > 1e-6), qrz), self.safe(gr, qgz),
                                                         d["cgacr"], acco[0, 2], acco[1, 2], acco[2, 2], den), qrz), 0.0)
        sink = pr + pw
        factor = xp.minimum(sink, self.dim(tice, tz) / icpk) / xp.maximum(sink, QRMIN)
        pr = W(gr, factor * pr, 0.0)
        pw = W(gr, factor * pw, 0.0)
        sink = pr + pw
        qgz, qrz, qlz = qgz + sink, qrz - pr, qlz - pw
        q_liq, q_sol = q_liq - sink, q_sol + sink
        cvm = W(gr, self.cvm(qvz, q_liq, q_sol), cvm)
        tz = W(gr, tz + sink * lhi / cvm, tz)
        return self.subgrid_z_proc(h_var, rh_adj, rh_rain, qaz, qgz, qiz, qlz, qrz, qsz, qvz, tz, den, denfac, p1)


# --------------------------------------------------------------- stencils
def _col(a, k):
    return a[..., k]


class Column:
    """Fields [ncol, nz] of one call; k-loops index the last axis."""


def _no_fall(xp, q):
    """FORWARD/BACKWARD pair (microphysics.py:349-368): True where no level of
    the column exceeds QRMIN."""
    nz = q.shape[-1]
    nf = [None] * nz
    nf[0] = xp.logical_not(q[..., 0] > QRMIN)
    for k in range(1, nz):
        nf[k] = xp.where(nf[k - 1], xp.logical_not(q[..., k] > QRMIN), False)
    for k in range(nz - 2, -1, -1):
        nf[k] = xp.where(xp.logical_not(nf[k + 1]), nf[k + 1], nf[k])
    return xp.stack_k(nf)


def _fall(xp, q, vt, dt, ze, zs, dp1, falls):
    """Implicit upwind sedimentation blocks of one species (e.g.
    microphysics.py:445-515).  Returns (q, flux, surface flux per column)."""
    nz = q.shape[-1]
    dz = xp.stack_k([ze[..., k] - ze[..., k + 1] for k in range(nz - 1)] + [ze[..., nz - 1] - zs])
    dd = dt * vt
    qq = q * dp1
    qm = [None] * nz
    qm[0] = qq[..., 0] / (dz[..., 0] + dd[..., 0])
    for k in range(1, nz):
        qm[k] = (qq[..., k] + dd[..., k - 1] * qm[k - 1]) / (dz[..., k] + dd[..., k])
    qm = xp.stack_k(qm) * dz
    fl = [None] * nz
    fl[0] = qq[..., 0] - qm[..., 0]
    for k in range(1, nz):
        fl[k] = fl[k - 1] + qq[..., k] - qm[..., k]
    flux = xp.stack_k(fl)
    surf = flux[..., nz - 1]
    f2 = falls[..., 0]
    return (xp.where(falls, qm / dp1, q), xp.where(falls, flux, 0.0), xp.where(f2, surf, 0.0))


def _heights(xp, dz1, zs):
    nz = dz1.shape[-1]
    ze = [None] * nz
    ze[nz - 1] = zs - dz1[..., nz - 1]
    for k in range(nz - 2, -1, -1):
        ze[k] = ze[k + 1] - dz1[..., k]
    return xp.stack_k(ze)


def _heat(xp, tz, flux, dz1, dp1, qv, ql, qr, qi, qs, qg, cq, apply):
    """Sedimentation heating (microphysics.py:516-557 / :1176-1217)."""
    nz = tz.shape[-1]
    dgz = -0.5 * GRAV * dz1
    cvn = dp1 * (CV_AIR + qv * CV_VAP + (qr + ql) * C_LIQ + (qi + qs + qg) * C_ICE)
    t = [None] * nz
    t[0] = tz[..., 0] + flux[..., 0] * dgz[..., 0] / (cvn[..., 0] + flux[..., 0] * cq)
    for k in range(1, nz):
        t[k] = ((cvn[..., k] + cq * (flux[..., k] - flux[..., k - 1])) * tz[..., k]
                + flux[..., k - 1] * cq * t[k - 1] + dgz[..., k] * (flux[..., k - 1] + flux[..., k])) / (
                   cvn[..., k] + cq * flux[..., k])
    if apply is True:
        return xp.stack_k(t)
    return xp.where(apply, xp.stack_k(t), tz)


def _slope(xp, q):
    """Vertical subgrid variability limiter (microphysics.py:595-620)."""
    nz = q.shape[-1]
    dq = [None] + [0.5 * (q[..., k] - q[..., k - 1]) for k in range(1, nz)]
    zero = q[..., 0] * 0.0
    dl = [zero]
    for k in range(1, nz - 1):
        v = 0.5 * xp.minimum(xp.abs(dq[k] + dq[k + 1]), 0.5 * q[..., k])
        v = xp.where(dq[k] * dq[k + 1] <= 0.0, xp.where(dq[k] > 0.0, xp.minimum(v, xp.minimum(dq[k], -dq[k + 1])), 0.0), v)
        dl.append(v)
    dl.append(zero)
    return xp.stack_k(dl)


def run(nml, timestep, state, qaz=None, tend=None, xp=None, derived_overrides=None):
    """One Microphysics call.  ``state`` maps pt qv ql qr qi qs qg qa ua va delp
    delz w ([ncol, nz]) and land area ([ncol]) to arrays of the namespace's
    type.  Returns the ten tendencies, qaz, w and the 2D precipitation.
    ``derived_overrides`` replaces derived scalars (e.g. fac_rc by a tensor
    that carries a gradient under TorchOps)."""
    xp = xp or Ops
    n = nml
    d = derived(nml, timestep)
    if derived_overrides:
        d.update(derived_overrides)
    P = Proc(xp, n, d)
    W, AND, NOT = xp.where, xp.logical_and, xp.logical_not
    if n["use_ppm"] or n["prog_ccn"]:
        raise ValueError("use_ppm and prog_ccn are not supported")
    pt, qv_in = state["pt"], state["qv"]
    nz = pt.shape[-1]
    land = state["land"][..., None]
    area = state["area"][..., None]
    zeros = pt * 0.0
    qaz = zeros if (qaz is None or n.get("reset_qaz")) else qaz
    # ---------------- fields_init (microphysics.py:101-276)
    qi, qs = state["qi"], state["qs"]
    tend = dict(tend) if tend is not None else {}
    tget = {k: tend.get(k, zeros) for k in ("pt_dt", "qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt",
                                             "qa_dt", "udt", "vdt")}
    if n["de_ice"]:
        qio = qi - timestep * tget["qi_dt"]
        qin = xp.maximum(qio, n["qi0_max"])
        big = qi > qin
        qs = W(big, qs + qi - qin, qs)
        dqi = (qin - qio) * d["rdt"]
        tget["qs_dt"] = W(big, tget["qs_dt"] + tget["qi_dt"] - dqi, tget["qs_dt"])
        tget["qi_dt"] = W(big, dqi, tget["qi_dt"])
        qi = W(big, qin, qi)
    t0 = pt
    tz = t0
    dp0 = state["delp"]
    dp1 = dp0 * (1.0 - qv_in)
    omq = dp0 / dp1
    qvz, qlz, qrz = qv_in * omq, state["ql"] * omq, state["qr"] * omq
    qiz, qsz, qgz = qi * omq, qs * omq, state["qg"] * omq
    qa0 = state["qa"]
    dz0 = state["delz"]
    den0 = -dp1 / (GRAV * dz0)
    p1 = den0 * RDGAS * t0
    qv0, ql0, qr0, qi0, qs0, qg0 = qvz, qlz, qrz, qiz, qsz, qgz
    m1 = zeros
    u0, v0 = state["ua"], state["va"]
    w = state.get("w")
    ccn = (n["ccn_l"] * land + n["ccn_o"] * (1.0 - land)) * 1.0e6 + zeros
    if d["use_ccn"]:
        cb = ccn[..., nz - 1] * RDGAS * tz[..., nz - 1] / p1[..., nz - 1]
        ccn = cb[..., None] + zeros
    c_praut = d["cpaut"] * xp.power(ccn * RHOR, -1.0 / 3.0)
    s_leng = xp.sqrt(xp.sqrt(area * 1.0e-10))
    h_var = (n["dw_land"] * s_leng) * land + (n["dw_ocean"] * s_leng) * (1.0 - land)
    h_var = xp.minimum(0.2, xp.maximum(0.01, h_var)) + zeros
    rh_adj = 1.0 - h_var - n["rh_inc"]
    rh_rain = xp.maximum(0.35, rh_adj - n["rh_inr"])
    if n["fix_negative"]:
        cvm = d["c_air"] + qvz * d["c_vap"] + (qrz + qlz) * C_LIQ + (qiz + qsz + qgz) * C_ICE
        lcpk = (d["lv00"] + d["d0_vap"] * tz) / cvm
        icpk = (LI00 + DC_ICE * tz) / cvm
        c = qiz < 0.0
        qsz, qiz = W(c, qsz + qiz, qsz), W(c, 0.0, qiz)
        c = qsz < 0.0
        qgz, qsz = W(c, qgz + qsz, qgz), W(c, 0.0, qsz)
        c = qgz < 0.0
        qrz, tz, qgz = W(c, qrz + qgz, qrz), W(c, tz - qgz * icpk, tz), W(c, 0.0, qgz)
        c = qrz < 0.0
        qlz, qrz = W(c, qlz + qrz, qlz), W(c, 0.0, qrz)
        c = qlz < 0.0
        qvz, tz, qlz = W(c, qvz + qlz, qvz), W(c, tz - qlz * lcpk, tz), W(c, 0.0, qlz)
        cols = [qvz[..., 0]]
        for k in range(1, nz):
            cols.append(W(cols[k - 1] < 0.0, qvz[..., k] + cols[k - 1] * dp1[..., k - 1] / dp1[..., k], qvz[..., k]))
        cols = [W(c_ < 0.0, 0.0, c_) if k < nz - 1 else c_ for k, c_ in enumerate(cols)]
        kb, ka = nz - 1, nz - 2
        flag = AND(cols[kb] < 0.0, cols[ka] > 0.0)
        dq = xp.minimum(-cols[kb] * dp1[..., kb], cols[ka] * dp1[..., ka])
        cols[ka] = W(flag, cols[ka] - dq / dp1[..., ka], cols[ka])
        cols[kb] = W(flag, cols[kb] + dq / dp1[..., kb], cols[kb])
        qvz = xp.stack_k(cols)
    rain = snow = ice = graupel = zeros[..., 0]
    # ---------------- sub-steps (microphysics.py:2045-2220)
    den = denfac = dz1 = vtrz = vtgz = vtsz = m1_sol = zeros
    for _ in range(d["ntimes"]):
        for is_first in (True, False):
            if is_first:
                dz1 = dz0 * tz / t0
                den = den0 * dz0 / dz1
                denfac = xp.sqrt(SFCRHO / den)
            # ---------------- warm_rain (microphysics.py:279-653)
            dt5 = 0.5 * d["dt_rain"]
            nf = _no_fall(xp, qrz)
            vtrz = P.rain_fspeed(nf, qrz, den)
            falls = NOT(nf)
            ze = _heights(xp, dz1, d["zs"])
            qlz, qrz, qvz, tz = P.revap_racc(falls, dt5, h_var, qgz, qiz, qlz, qrz, qsz, qvz, tz, den, denfac)
            if d["do_sedi_w"]:
                dm = dp1 * (1.0 + qvz + qlz + qrz + qiz + qsz + qgz)
            qrz, m1_rain, r1 = _fall(xp, qrz, vtrz, d["dt_rain"], ze, d["zs"], dp1, falls)
            if d["do_sedi_w"]:
                w = _w_transport(xp, w, dm, m1_rain, vtrz, falls)
            if n["do_sedi_heat"]:
                tz = _heat(xp, tz, m1_rain, dz1, dp1, qvz, qlz, qrz, qiz, qsz, qgz, C_LIQ, falls)
            qlz, qrz, qvz, tz = P.revap_racc(falls, dt5, h_var, qgz, qiz, qlz, qrz, qsz, qvz, tz, den, denfac)
            if n["irain_f"] != 0:
                qlz, qrz = P.autoconv_no_subgrid(qlz, qrz, tz, den, ccn, c_praut)
            else:
                if n["z_slope_liq"]:
                    dl = xp.maximum(_slope(xp, qlz), xp.maximum(QVMIN, h_var * qlz))
                else:
                    dl = xp.maximum(QVMIN, h_var * qlz)
                qlz, qrz = P.autoconv_subgrid(qlz, qrz, tz, den, ccn, c_praut, dl)
            rain = rain + r1
            m1 = m1 + m1_rain if is_first else m1 + m1_rain + m1_sol
            if is_first:
                # ---------------- sedimentation
                (qgz, qiz, qlz, qrz, qsz, qvz, tz, w, vtgz, vtsz, m1_sol, i1, s1, g1) = _sedimentation(
                    xp, P, n, d, qgz, qiz, qlz, qrz, qsz, qvz, tz, den, dz1, dp1, w)
                snow, graupel, ice = snow + s1, graupel + g1, ice + i1
        # ---------------- icloud (microphysics.py:1220-1449)
        lhi = LI00 + DC_ICE * tz
        q_liq = qlz + qrz
        q_sol = qiz + qsz + qgz
        cvm = P.cvm(qvz, q_liq, q_sol)
        icpk = lhi / cvm
        tw = n["tice"]
        mlt = AND(tz > tw, qiz > QCMIN)
        melt = xp.minimum(qiz, d["fac_imlt"] * (tz - tw) / icpk)
        tmp = xp.minimum(melt, P.dim(n["ql_mlt"], qlz))
        hom = AND(NOT(mlt), AND(tz < d["t_wfr"], qlz > QCMIN))
        dtmp = d["t_wfr"] - tz
        factor = xp.minimum(1.0, dtmp / DT_FR)
        sink = xp.minimum(qlz * factor, dtmp / icpk)
        qi_crt = n["qi_gen"] * xp.minimum(n["qi_lim"], 0.1 * (tw - tz)) / den
        tmp2 = xp.minimum(sink, P.dim(qi_crt, qiz))
        qlz_n = W(mlt, qlz + tmp, W(hom, qlz - sink, qlz))
        qrz = W(mlt, qrz + melt - tmp, qrz)
        qsz = W(hom, qsz + sink - tmp2, qsz)
        qiz_n = W(mlt, qiz - melt, W(hom, qiz + tmp2, qiz))
        q_liq = W(mlt, q_liq + melt, W(hom, q_liq - sink, q_liq))
        q_sol = W(mlt, q_sol - melt, W(hom, q_sol + sink, q_sol))
        qlz, qiz = qlz_n, qiz_n
        cvm = W(NOT(AND(NOT(mlt), NOT(hom))), P.cvm(qvz, q_liq, q_sol), cvm)
        tz = W(mlt, tz - melt * lhi / cvm, W(hom, tz + sink * lhi / cvm, tz))
        if n["z_slope_ice"]:
            di = xp.maximum(_slope(xp, qiz), xp.maximum(QVMIN, h_var * qiz))
        else:
            di = xp.maximum(QVMIN, h_var * qiz)
        qaz, qgz, qiz, qlz, qrz, qsz, qvz, tz = P.icloud_main(
            h_var, rh_adj, rh_rain, qaz, qgz, qiz, qlz, qrz, qsz, qvz, tz, den, denfac, vtgz, vtrz, vtsz, p1, di,
            q_liq, q_sol, cvm)
    # ---------------- fields_update (microphysics.py:1452-1553)
    rdt = d["rdt"]
    out = {}
    udt, vdt = tget["udt"], tget["vdt"]
    if n["sedi_transport"]:
        u1 = [u0[..., 0]]
        v1 = [v0[..., 0]]
        for k in range(1, nz):
            u1.append((dp0[..., k] * u0[..., k] + m1[..., k - 1] * u1[k - 1]) / (dp0[..., k] + m1[..., k - 1]))
            v1.append((dp0[..., k] * v0[..., k] + m1[..., k - 1] * v1[k - 1]) / (dp0[..., k] + m1[..., k - 1]))
        udt = xp.stack_k([udt[..., 0]] + [udt[..., k] + (u1[k] - u0[..., k]) * rdt for k in range(1, nz)])
        vdt = xp.stack_k([vdt[..., 0]] + [vdt[..., k] + (v1[k] - v0[..., k]) * rdt for k in range(1, nz)])
    omq = dp1 / dp0
    out["qv_dt"] = tget["qv_dt"] + rdt * (qvz - qv0) * omq
    out["ql_dt"] = tget["ql_dt"] + rdt * (qlz - ql0) * omq
    out["qr_dt"] = tget["qr_dt"] + rdt * (qrz - qr0) * omq
    out["qi_dt"] = tget["qi_dt"] + rdt * (qiz - qi0) * omq
    out["qs_dt"] = tget["qs_dt"] + rdt * (qsz - qs0) * omq
    out["qg_dt"] = tget["qg_dt"] + rdt * (qgz - qg0) * omq
    cvm = d["c_air"] + qvz * d["c_vap"] + (qrz + qlz) * C_LIQ + (qiz + qsz + qgz) * C_ICE
    out["pt_dt"] = tget["pt_dt"] + rdt * (tz - t0) * cvm / CP_AIR
    out["qa_dt"] = zeros if n["do_qa"] else tget["qa_dt"] + rdt * (qaz / d["ntimes"] - qa0)
    out["udt"], out["vdt"] = udt, vdt
    convt = 86400.0 * rdt * RGRAV
    out["rain"], out["snow"], out["ice"], out["graupel"] = rain * convt, snow * convt, ice * convt, graupel * convt
    out["qaz"] = qaz
    out["w"] = w
    out["qi"], out["qs"] = qi, qs
    return out


def _w_transport(xp, w, dm, flux, vt, falls):
    """Vertical velocity transport (microphysics.py:499-515)."""
    nz = w.shape[-1]
    cols = [(dm[..., 0] * w[..., 0] + flux[..., 0] * vt[..., 0]) / (dm[..., 0] - flux[..., 0])]
    for k in range(1, nz):
        cols.append((dm[..., k] * w[..., k] - flux[..., k - 1] * vt[..., k - 1] + flux[..., k] * vt[..., k])
                    / (dm[..., k] + flux[..., k - 1] - flux[..., k]))
    return xp.where(falls, xp.stack_k(cols), w)


def _sedimentation(xp, P, n, d, qgz, qiz, qlz, qrz, qsz, qvz, tz, den, dz1, dp1, w):
    W, AND, NOT = xp.where, xp.logical_and, xp.logical_not
    nz = tz.shape[-1]
    vtgz, vtiz, vtsz = P.fall_speed(qgz, qiz, qsz, tz, den)
    lhi = LI00 + DC_ICE * tz
    q_liq = qlz + qrz
    q_sol = qiz + qsz + qgz
    cvm = P.cvm(qvz, q_liq, q_sol)
    icpk = lhi / cvm
    # melting level (:706-724)
    hot = tz > n["tice"]
    sk = [hot[..., 0]]
    for k in range(1, nz - 1):
        sk.append(W(NOT(sk[k - 1]), hot[..., k], True))
    sk.append(xp.logical_or(hot[..., nz - 1], True))
    stop_k = xp.stack_k(sk)
    tc = tz - n["tice"]
    mel = AND(stop_k, AND(qiz > QCMIN, tc > 0.0))
    sink = xp.minimum(qiz, d["fac_imlt"] * tc / icpk)
    tmp = xp.minimum(sink, P.dim(n["ql_mlt"], qlz))
    sink = W(mel, sink, 0.0)
    tmp = W(mel, tmp, 0.0)
    qlz, qrz, qiz = qlz + tmp, qrz + sink - tmp, qiz - sink
    q_liq, q_sol = q_liq + sink, q_sol - sink
    cvm = W(mel, P.cvm(qvz, q_liq, q_sol), cvm)
    tz = W(mel, tz - sink * lhi / cvm, tz)
    ze = _heights(xp, dz1, d["zs"])
    zeros = tz * 0.0
    m1_sol = zeros
    i1 = zeros[..., 0]
    # cloud ice
    nf = _no_fall(xp, qiz)
    falls = NOT(nf)
    if not (n["vi_fac"] < 1.0e-5):
        if d["do_sedi_w"]:
            dm = dp1 * (1.0 + qvz + qlz + qrz + qiz + qsz + qgz)
        qiz, m1_sol, i1 = _fall(xp, qiz, vtiz, d["dts"], ze, d["zs"], dp1, falls)
        if d["do_sedi_w"]:
            w = _w_transport(xp, w, dm, m1_sol, vtiz, falls)
    # snow
    nf = _no_fall(xp, qsz)
    falls = NOT(nf)
    if d["do_sedi_w"]:
        dm = dp1 * (1.0 + qvz + qlz + qrz + qiz + qsz + qgz)
    qsz, fl, s1 = _fall(xp, qsz, vtsz, d["dts"], ze, d["zs"], dp1, falls)
    m1_sol = m1_sol + fl
    if d["do_sedi_w"]:
        w = _w_transport(xp, w, dm, fl, vtsz, falls)
    # graupel
    nf = _no_fall(xp, qgz)
    falls = NOT(nf)
    if d["do_sedi_w"]:
        dm = dp1 * (1.0 + qvz + qlz + qrz + qiz + qsz + qgz)
    qgz, fl, g1 = _fall(xp, qgz, vtgz, d["dts"], ze, d["zs"], dp1, falls)
    m1_sol = m1_sol + fl
    if d["do_sedi_w"]:
        w = _w_transport(xp, w, dm, fl, vtgz, falls)
    if n["do_sedi_heat"]:
        tz = _heat(xp, tz, m1_sol, dz1, dp1, qvz, qlz, qrz, qiz, qsz, qgz, C_ICE, True)
    return qgz, qiz, qlz, qrz, qsz, qvz, tz, w, vtgz, vtsz, m1_sol, i1, s1, g1
