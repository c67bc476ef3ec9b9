"""The fused microphysics + state-update call (shield_mp_forward_update) against
the same sequence run as separate kernels (shield_mp_forward,
update_physics_state_with_tendencies, fill_gfs_delp,
prepare_tendencies_and_update_tracers), and in FP64 against the oracle chain
(the C oracle, then the NumPy restatements of the three update stencils)."""

import numpy as np
import pytest

from oracle import update_numpy as U
from pyshield_amd._config import PhysicsConfig

from helpers import TEND, norm_err, run_oracle, small_state

pytestmark = pytest.mark.gpu

DT = 150.0
# physics-state sources of the ten physics_updated fields, in the order of
# shield_update_fields.updated (qcld is the microphysics' qa)
UPD_SRC = ("qv", "ql", "qr", "qi", "qs", "qg", "qa", "pt", "ua", "va")
UPD_TEND = ("qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt", "qa_dt", "pt_dt", "udt", "vdt")
# physics-state species of the six FV3 tracers, in the order of .tracers
TR_SRC = ("qv", "ql", "qr", "qs", "qi", "qg")


def _dycore(st, seed):
    rng = np.random.default_rng(seed)
    ni, nj, nz = st["pt"].shape
    sh = (ni, nj, nz)
    tracers = [st[s] * rng.uniform(0.98, 1.02, sh) for s in TR_SRC]
    prsi = np.concatenate([np.full((ni, nj, 1), 100.0), 100.0 + np.cumsum(st["delp"], -1)], -1)
    delp = st["delp"] * rng.uniform(0.99, 1.01, sh)
    tend = [rng.normal(0, 1e-4, sh) for _ in range(3)]
    return tracers, prsi, delp, tend


def _run(st, dyn, mode, fused):
    import torch
    from pyshield_amd import _lib
    from pyshield_amd.microphysics import forward, forward_update
    from pyshield_amd.update import (fill_gfs_delp, prepare_tendencies_and_update_tracers,
                                     update_physics_state_with_tendencies)
    dev = torch.device("cuda:0")
    ni, nj, nz = st["pt"].shape

    def t(x):
        return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev).clone()

    def z():
        return torch.zeros((ni, nj, nz), dtype=torch.float64, device=dev)

    f = {n: t(st[n]) for n in ("pt", "qv", "ql", "qr", "qi", "qs", "qg", "qa", "ua", "va",
                               "delp", "delz", "w", "land", "area")}
    for n in TEND + ("qaz",):
        f[n] = z()
    tracers, prsi, delp, (u_dt, v_dt, pt_dt) = dyn
    upd = {"updated": {n: z() for n in _lib.UPDATED_NAMES}, "u_dt": t(u_dt), "v_dt": t(v_dt),
           "pt_dt": t(pt_dt), "tracers": {n: t(x) for n, x in zip(_lib.TRACER_NAMES, tracers)},
           "delp": t(delp), "prsi": t(prsi)}
    params = _lib.setup(PhysicsConfig(dt_atmos=int(DT), npz=nz).mp_namelist(), DT)
    params.accumulate = 0
    if fused:
        forward_update(f, upd, params, dt=DT, rdt=1.0 / DT, mode=mode)
    else:
        forward(f, params, mode)
        u = upd["updated"]
        update_physics_state_with_tendencies(*[f[n] for n in UPD_SRC], *[f[n] for n in UPD_TEND],
                                             *[u[n] for n in _lib.UPDATED_NAMES], DT)
        fill_gfs_delp(upd["delp"], u["qvapor"], 1.0e-9)
        tr = upd["tracers"]
        prepare_tendencies_and_update_tracers(
            upd["u_dt"], upd["v_dt"], upd["pt_dt"], u["ua"], u["va"], u["pt"], u["qvapor"], u["qliquid"],
            u["qrain"], u["qsnow"], u["qice"], u["qgraupel"], f["ua"], f["va"], f["pt"], *[tr[n] for n in
                                                                                          _lib.TRACER_NAMES],
            upd["prsi"], upd["delp"], 1.0 / DT)
    torch.cuda.synchronize()
    out = {n: f[n].cpu().numpy() for n in TEND + ("qaz",)}
    out.update({"upd_" + n: x.cpu().numpy() for n, x in upd["updated"].items()})
    out.update({"tr_" + n: x.cpu().numpy() for n, x in upd["tracers"].items()})
    for n in ("u_dt", "v_dt", "pt_dt", "delp"):
        out[n] = upd[n].cpu().numpy()
    return out


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_fused_update_matches_the_separate_kernels(mode):
    st = small_state(12, 12, 79, seed=1000)
    dyn = _dycore(st, 3)
    a = _run(st, dyn, mode, fused=True)
    b = _run(st, dyn, mode, fused=False)
    # the fused instance may deal columns to slots differently (scan
    # association), so agreement is to round-off
    bad = {n: norm_err(a[n], b[n], floor=1e-30) for n in a}
    print({k: f"{v:.1e}" for k, v in bad.items()}, flush=True)
    assert all(v <= 1e-12 for v in bad.values()), bad
    assert np.all(a["upd_qvapor"] >= 0.0)


def test_fused_update_matches_the_oracle_chain():
    st = small_state(12, 12, 79, seed=2024)
    dyn = _dycore(st, 5)
    got = _run(st, dyn, "fp64", fused=True)
    cfg = PhysicsConfig(dt_atmos=int(DT), npz=79)
    ora = run_oracle(st, cfg, DT)
    upd = U.update_physics_state_with_tendencies([st[n] for n in UPD_SRC], [ora[n] for n in UPD_TEND], DT)
    upd[0] = U.fill_gfs_delp(dyn[2], upd[0], 1.0e-9)
    tracers, prsi, delp, (u_dt, v_dt, pt_dt) = dyn
    by = dict(zip(UPD_SRC, upd))
    ref = U.prepare_tendencies_and_update_tracers(
        u_dt, v_dt, pt_dt, by["ua"], by["va"], by["pt"], by["qv"], by["ql"], by["qr"], by["qs"], by["qi"],
        by["qg"], st["ua"], st["va"], st["pt"], *tracers, prsi, delp, 1.0 / DT)
    names = ("u_dt", "v_dt", "pt_dt") + tuple("tr_" + n for n in ("qvapor", "qliquid", "qrain", "qsnow", "qice",
                                                                   "qgraupel")) + ("delp",)
    errs = {n: norm_err(got[n], r, floor=1e-30) for n, r in zip(names, ref)}
    for n, src in zip(("qvapor", "qliquid", "qrain", "qice", "qsnow", "qgraupel", "qcld", "pt", "ua", "va"),
                      UPD_SRC):
        errs["upd_" + n] = norm_err(got["upd_" + n], by[src], floor=1e-30)
    print({k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert all(v <= 1e-9 for v in errs.values()), errs
