"""Shared helpers of the parity tests: synthetic inputs, one GPU call through
the C ABI, one oracle call, and the error metrics."""

import numpy as np

import oracle
from pyshield_amd import constants as C
from pyshield_amd._config import PhysicsConfig
from pyshield_amd.synthetic import make_state

TEND = ("pt_dt", "qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt", "qa_dt", "udt", "vdt")
PRECIP = ("rain", "snow", "ice", "graupel")
WATER = ("qv", "ql", "qr", "qi", "qs", "qg")


def small_state(ni=4, nj=6, nz=79, seed=7, **kw):
    return make_state(ni, nj, nz, seed, **kw)


def run_gpu(state, cfg: PhysicsConfig, timestep, mode="fp64", qaz=None, tend=None):
    """One shield_mp_forward call on cuda:0; returns numpy outputs like oracle.run."""
    import torch
    from pyshield_amd import _lib
    from pyshield_amd.microphysics import forward

    dev = torch.device("cuda:0")
    ni, nj, nz = state["pt"].shape

    def t(x):
        return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev).clone()

    f = {n: t(state[n]) for n in ("pt", "qv", "ql", "qr", "qi", "qs", "qg", "qa", "ua", "va",
                                  "delp", "delz", "w", "land", "area")}
    for n in TEND:
        f[n] = t(tend[n]) if (tend is not None and n in tend) else torch.zeros((ni, nj, nz), dtype=torch.float64, device=dev)
    f["qaz"] = t(qaz) if qaz is not None else torch.zeros((ni, nj, nz), dtype=torch.float64, device=dev)
    for n in PRECIP:
        f[n] = torch.zeros((ni, nj), dtype=torch.float64, device=dev)
    params = _lib.setup(cfg.mp_namelist(), timestep)
    forward(f, params, mode)
    torch.cuda.synchronize()
    out = {n: f[n].cpu().numpy() for n in TEND + PRECIP + ("qaz", "qi", "qs", "w")}
    return out


def run_oracle(state, cfg: PhysicsConfig, timestep, qaz=None, tend=None):
    return oracle.run(cfg.mp_namelist(), timestep, state, qaz=qaz, tend=tend)


def norm_err(a, b, floor=0.0):
    """max |a - b| / max(max |b|, floor)"""
    scale = max(float(np.max(np.abs(b))) if b.size else 0.0, floor)
    if scale == 0.0:
        return float(np.max(np.abs(a - b))) if a.size else 0.0
    return float(np.max(np.abs(a - b))) / scale


def updated_state(state, out, timestep):
    """The state after the forward-Euler update (physics.py:146-194)."""
    new = {"pt": state["pt"] + out["pt_dt"] * timestep}
    for q in WATER:
        new[q] = state[q] + out[f"{q}_dt"] * timestep
    return new


def water_imbalance(state, out, timestep):
    """Per column: (sum_k dp0 sum_species q_dt dt + surface precipitation in Pa)
    / total column water.  Zero for an exactly conservative scheme."""
    convt = 86400.0 / timestep / C.GRAV
    wat = sum(out[f"{q}_dt"] for q in WATER)
    col = (state["delp"] * wat * timestep).sum(-1)
    pr = sum(out[p] for p in PRECIP) / convt
    tot = (state["delp"] * sum(state[q] for q in WATER)).sum(-1)
    return np.abs((col + pr) / tot)


def dry_state(ni=3, nj=4, nz=79, seed=5, rh=0.2):
    """No condensate anywhere and vapour at rh x saturation over water."""
    from pyshield_amd.synthetic import qsat_water
    st = make_state(ni, nj, nz, seed)
    for q in ("ql", "qr", "qi", "qs", "qg"):
        st[q] = np.zeros_like(st[q])
    pmid = np.cumsum(st["delp"], axis=-1) - 0.5 * st["delp"] + 300.0
    st["qv"] = rh * qsat_water(st["pt"], pmid / (C.RDGAS * st["pt"]))
    return st


def cold_top_state(qv_top=1.0e-8, t_top=170.0):
    """A dry column whose top layer is colder than T_MIN (178 K)."""
    st = dry_state(1, 1, 79, seed=9)
    st["pt"][..., 0] = t_top
    st["qv"][..., 0] = qv_top
    return st


def overdraw_state(ni=4, nj=6, nz=79, seed=515):
    """Warm layers (>= 280 K) holding 1e-2 .. 3e-2 kg/kg of cloud water in half
    of the columns.  There the subgrid autoconversion of one rain step
    (dt c_praut den ql^(7/3)) exceeds ql, which the reference does not bound:
    ql and, after the saturation adjustment, qv turn negative and the layers
    warm by up to ~180 K."""
    st = make_state(ni, nj, nz, seed)
    hot = (slice(0, ni // 2), slice(None), slice(52, 60))
    st["ql"][hot] = np.linspace(1.0e-2, 3.0e-2, 8)
    st["pt"][hot] = np.maximum(st["pt"][hot], 280.0)
    return st


def low_pressure_top_state(thin=True):
    """A dry column whose top layer is supersaturated over ice at 200 K; with
    thin=True its dry pressure p1 is below P_MIN (100 Pa)."""
    from pyshield_amd.synthetic import qsat_water
    st = dry_state(1, 1, 79, seed=10)
    st["pt"][..., 0] = 200.0
    if thin:
        st["delz"][..., 0] = -20000.0
    den = -st["delp"][..., 0] / (C.GRAV * st["delz"][..., 0])
    st["qv"][..., 0] = 1.2 * qsat_water(np.array(200.0), den)
    return st
