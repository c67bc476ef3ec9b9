"""TEST INFRASTRUCTURE -- NumPy restatement of the Physics driver prologue:
atmos_phys_driver_statein (pySHiELD/stencils/physics.py:26-102), get_prs_fv3
(pySHiELD/stencils/get_prs_fv3.py:7-20) and get_phi_fv3
(pySHiELD/stencils/get_phi_fv3.py:7-25), in the order Physics.__call__ runs
them (physics.py:278-310).  Layer fields are [..., nk] with k = 0 at the model
top; interface fields (phii, prsi, prsik) are [..., nk + 1].  The reference
stores every field with nk + 1 levels; its extra level of a layer field is
padding that no layer result reads, so it is not represented here.
"""

import numpy as np

GRAV = 9.80665
RDGAS = 287.05
RVGAS = 461.50
CP_AIR = 1004.6
ZVIR = RVGAS / RDGAS - 1.0
KAPPA = RDGAS / CP_AIR
P00 = 1.0e5
SPECIES = ("qvapor", "qliquid", "qrain", "qice", "qsnow", "qgraupel", "qo3mr")


def statein(delz, delp, q, qsgs_tke, pt, ptop, nwat=6):
    """physics.py:26-102.  q maps SPECIES to arrays.  Returns phii, prsi, prsik,
    the layer-mean pressure (the reference writes it into delp), the converted
    tracers (dict), qsgs_tke and dm (the dry layer mass)."""
    nk = delp.shape[-1]
    sh = delp.shape[:-1]
    phii = np.zeros(sh + (nk + 1,))
    for k in range(nk - 1, -1, -1):
        phii[..., k] = phii[..., k + 1] - delz[..., k] * GRAV
    mass = {n: q[n] * delp for n in SPECIES}
    dry = delp
    if nwat == 6:
        dry = delp - mass["qliquid"] - mass["qrain"] - mass["qice"] - mass["qsnow"] - mass["qgraupel"]
    prsi = np.empty(sh + (nk + 1,))
    prsi[..., 0] = ptop
    for k in range(1, nk + 1):
        prsi[..., k] = prsi[..., k - 1] + dry[..., k - 1]
    prsik = np.log(prsi)
    prsik[..., 0] = np.log(ptop)
    out_q = {n: mass[n] / dry for n in SPECIES}
    tke = qsgs_tke / dry
    qgrs_rad = np.maximum(1.0e-10, out_q["qvapor"])
    rtv = RDGAS * pt * (1.0 + ZVIR * qgrs_rad)
    dm = dry
    prsl = dm * rtv / (phii[..., :-1] - phii[..., 1:])
    prsl = np.minimum(prsl, prsi[..., 1:] - 0.01 * dm)
    prsl = np.maximum(prsl, prsi[..., :-1] + 0.01 * dm)
    pk0inv = (1.0 / P00) ** KAPPA
    pktop = (ptop / P00) ** KAPPA
    prsik[..., nk] = np.exp(KAPPA * prsik[..., nk]) * pk0inv
    prsik[..., 0] = pktop
    return phii, prsi, prsik, prsl, out_q, tke, dm


def get_prs_fv3(phii, prsi, tgrs, qgrs):
    """get_prs_fv3.py:7-20 -> (del_, del_gz)."""
    del_ = prsi[..., 1:] - prsi[..., :-1]
    del_gz = (phii[..., :-1] - phii[..., 1:]) / (tgrs * (1.0 + ZVIR * np.maximum(0.0, qgrs)))
    return del_, del_gz


def get_phi_fv3(gt0, gq0, del_gz):
    """get_phi_fv3.py:7-25 -> (del_gz, phii, phil)."""
    nk = gt0.shape[-1]
    del_gz = del_gz * gt0 * (1.0 + ZVIR * np.maximum(0.0, gq0))
    phii = np.zeros(gt0.shape[:-1] + (nk + 1,))
    phil = np.empty_like(gt0)
    for k in range(nk - 1, -1, -1):
        phil[..., k] = 0.5 * (phii[..., k + 1] + phii[..., k + 1] + del_gz[..., k])
        phii[..., k] = phii[..., k + 1] + del_gz[..., k]
    return del_gz, phii, phil


def prologue(delz, delp, q, qsgs_tke, pt, ptop):
    """The three stencils in Physics.__call__ order; returns a dict of every
    output (delp is the layer-mean pressure, delprsi the dry layer mass from
    the interface pressures)."""
    phii, prsi, prsik, prsl, out_q, tke, dm = statein(delz, delp, q, qsgs_tke, pt, ptop)
    delprsi, del_gz = get_prs_fv3(phii, prsi, pt, out_q["qvapor"])
    del_gz, phii, phil = get_phi_fv3(pt, out_q["qvapor"], del_gz)
    out = {"phii": phii, "prsi": prsi, "prsik": prsik, "delp": prsl, "qsgs_tke": tke, "dm": dm,
           "delprsi": delprsi, "del_gz": del_gz, "phil": phil}
    out.update(out_q)
    return out
