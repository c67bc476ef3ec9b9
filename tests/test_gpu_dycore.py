"""shield_apply_physics_to_dycore against the NumPy restatement
(oracle/dycore_numpy.py) on synthetic one-tile cases: moist_cv, the pressure
recurrence and surface winds, and the D-grid wind update (grid_type 0 with
all four tile edges, and the doubly periodic form).  Every output is a short
FP64 expression of its inputs, so the bar is 1e-13 relative, with an absolute
floor of 1e-14 of the field's largest magnitude for values a sum cancels to
near zero (log / exp and fused multiply-adds differ from NumPy by a few
ulp)."""

import types

import numpy as np
import pytest

from oracle import dycore_numpy as D

pytestmark = pytest.mark.gpu

KAPPA = 287.05 / 1004.6
Q = ("qvapor", "qliquid", "qrain", "qsnow", "qice", "qgraupel")


def _case(ni, nj, nk, seed):
    rng = np.random.default_rng(seed)
    s = {n: rng.uniform(0, 2e-3, (ni, nj, nk)) for n in Q}
    s["qvapor"] = rng.uniform(1e-6, 2e-2, (ni, nj, nk))
    s["pt"] = rng.uniform(200, 310, (ni, nj, nk))
    s["t_dt"] = rng.normal(0, 1e-4, (ni, nj, nk))
    s["delp"] = rng.uniform(50, 2000, (ni, nj, nk))
    pe = np.zeros((ni, nj, nk + 1))
    pe[..., 0] = 300.0
    s["pe"], s["peln"], s["pk"] = pe, rng.normal(size=pe.shape), rng.normal(size=pe.shape)
    s["ps"] = np.zeros((ni, nj))
    s["ua"], s["va"] = rng.normal(0, 10, (ni, nj, nk)), rng.normal(0, 10, (ni, nj, nk))
    s["u"], s["v"] = rng.normal(0, 10, (ni, nj + 1, nk)), rng.normal(0, 10, (ni + 1, nj, nk))
    s["u_dt"] = rng.normal(0, 1e-3, (ni + 2, nj + 2, nk))
    s["v_dt"] = rng.normal(0, 1e-3, (ni + 2, nj + 2, nk))
    g = {"vlon": [rng.uniform(-1, 1, (ni + 2, nj + 2)) for _ in range(3)],
         "vlat": [rng.uniform(-1, 1, (ni + 2, nj + 2)) for _ in range(3)],
         "es1": [rng.uniform(-1, 1, (ni, nj + 1)) for _ in range(3)],
         "ew2": [rng.uniform(-1, 1, (ni + 1, nj)) for _ in range(3)],
         "edge_vect_w": rng.uniform(0, 1, nj), "edge_vect_e": rng.uniform(0, 1, nj),
         "edge_vect_s": rng.uniform(0, 1, ni), "edge_vect_n": rng.uniform(0, 1, ni)}
    return s, g


def _flat_grid(g):
    out = {f"vlon{c + 1}": g["vlon"][c] for c in range(3)}
    out.update({f"vlat{c + 1}": g["vlat"][c] for c in range(3)})
    out.update({f"es1_{c + 1}": g["es1"][c] for c in range(3)})
    out.update({f"ew2_{c + 1}": g["ew2"][c] for c in range(3)})
    out.update({n: g[n] for n in ("edge_vect_w", "edge_vect_e", "edge_vect_s", "edge_vect_n")})
    return out


def _t(x):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda:0").clone()


def _oracle(s, g, dt, npx, npy, grid_type):
    pt, t_dt = D.moist_cv({n: s[n] for n in Q}, s["pt"], s["t_dt"], 1004.6, dt)
    pe, peln, pk, ps, us, vs = D.pressure_and_surface_winds(s["pe"], s["delp"], s["peln"], s["pk"], KAPPA,
                                                           s["ua"], s["va"])
    u, v, udt, vdt = D.agrid_to_dgrid(s["u"], s["v"], s["u_dt"], s["v_dt"], g, dt, npx, npy, grid_type)
    return dict(pt=pt, t_dt=t_dt, pe=pe, peln=peln, pk=pk, ps=ps, u_srf=us, v_srf=vs, u=u, v=v, u_dt=udt,
                v_dt=vdt)


def _check(got, want, where):
    for n, w in want.items():
        g = got[n].cpu().numpy()
        if n == "peln" or n == "pk":
            g, w = g[..., 1:], w[..., 1:]   # the top interface is not written
        # relative to the value, and to the field's scale where a sum cancels
        np.testing.assert_allclose(g, w, rtol=1e-13, atol=1e-14 * float(np.abs(w).max(initial=0.0)),
                                   err_msg=f"{where}: {n}")


@pytest.mark.parametrize("ni,nj,nk,grid_type", [(8, 8, 79, 0), (7, 10, 5, 0), (3, 3, 4, 0), (2, 2, 1, 0),
                                                (48, 48, 79, 0), (6, 5, 7, 4), (9, 7, 300, 0)])
def test_apply_physics_to_dycore_matches_oracle(ni, nj, nk, grid_type):
    import torch
    from pyshield_amd.dycore import apply_physics_to_dycore
    print(f"dycore case ni={ni} nj={nj} nk={nk} grid_type={grid_type}", flush=True)
    s, g = _case(ni, nj, nk, seed=ni * 100 + nj)
    dt = 225.0
    npx, npy = (ni + 1, nj + 1) if grid_type <= 3 else (0, 0)
    want = _oracle(s, g, dt, npx, npy, grid_type)
    fields = {n: _t(x) for n, x in s.items()}
    fields["u_srf"] = torch.zeros((ni, nj), dtype=torch.float64, device="cuda:0")
    fields["v_srf"] = torch.zeros_like(fields["u_srf"])
    grid = {n: _t(x) for n, x in _flat_grid(g).items()} if grid_type <= 3 else {}
    apply_physics_to_dycore(fields, grid, dt, npx, npy, grid_type)
    torch.cuda.synchronize()
    _check(fields, want, f"{ni}x{nj}x{nk} grid_type {grid_type}")
    # the top interface of peln and pk is left alone
    np.testing.assert_array_equal(fields["peln"][..., 0].cpu().numpy(), s["peln"][..., 0])


def test_apply_physics_to_dycore_class_and_repeat():
    # the ApplyPhysicsToDycore mirror on a state namespace, twice in a row:
    # the second call sees cleared t_dt, u_dt, v_dt and changes nothing but
    # the (recomputed) pressures
    import torch
    from pyshield_amd.dycore import ApplyPhysicsToDycore
    ni, nj, nk = 12, 12, 16
    s, g = _case(ni, nj, nk, seed=11)
    want = _oracle(s, g, 150.0, ni + 1, nj + 1, 0)
    state = types.SimpleNamespace(**{n: _t(s[n]) for n in Q + ("pt", "delp", "pe", "peln", "pk", "ps", "ua",
                                                                "va", "u", "v")})
    grid_info = types.SimpleNamespace(grid_type=0, **{n: _t(x) for n, x in _flat_grid(g).items()})
    seen = []
    op = ApplyPhysicsToDycore(types.SimpleNamespace(npx=ni + 1, npy=nj + 1, dt_atmos=150.0), grid_info,
                              cubed_to_latlon=lambda u, v, ua, va: seen.append((u.data_ptr(), v.data_ptr())))
    u_dt, v_dt, t_dt = _t(s["u_dt"]), _t(s["v_dt"]), _t(s["t_dt"])
    op(state, u_dt, v_dt, t_dt, 150.0)
    torch.cuda.synchronize()
    got = {n: getattr(state, n) for n in ("pt", "pe", "peln", "pk", "ps", "u", "v")}
    got.update(t_dt=t_dt, u_dt=u_dt, v_dt=v_dt, u_srf=op.u_srf, v_srf=op.v_srf)
    _check(got, want, "class call")
    assert seen == [(state.u.data_ptr(), state.v.data_ptr())]
    before = {n: getattr(state, n).clone() for n in ("pt", "u", "v", "pe")}
    op(state, u_dt, v_dt, t_dt, 150.0)
    torch.cuda.synchronize()
    for n, x in before.items():
        torch.testing.assert_close(getattr(state, n), x, rtol=0, atol=0)


def test_apply_physics_to_dycore_rejects_wrong_shapes():
    import torch
    from pyshield_amd.dycore import apply_physics_to_dycore
    s, g = _case(4, 4, 3, seed=1)
    fields = {n: _t(x) for n, x in s.items()}
    fields["u_srf"] = torch.zeros((4, 4), dtype=torch.float64, device="cuda:0")
    fields["v_srf"] = torch.zeros_like(fields["u_srf"])
    grid = {n: _t(x) for n, x in _flat_grid(g).items()}
    bad = dict(fields, u=_t(np.zeros((4, 4, 3))))
    with pytest.raises(ValueError):
        apply_physics_to_dycore(bad, grid, 150.0, 5, 5)
    bad = dict(fields, pt=fields["pt"].float())
    with pytest.raises(TypeError):
        apply_physics_to_dycore(bad, grid, 150.0, 5, 5)
    from pyshield_amd._lib import ShieldError
    with pytest.raises(ShieldError):
        apply_physics_to_dycore(fields, grid, 150.0, 9, 5)   # npx - 1 != ni


def _halo(x, fill):
    """[ni, nj, nk] -> [ni + 2, nj + 2, nk]: the edges of the halo from
    ``fill(h)`` (h the halo'd array with the interior set), corners 0."""
    ni, nj, nk = x.shape
    h = np.zeros((ni + 2, nj + 2, nk))
    h[1:-1, 1:-1] = x
    fill(h)
    return h


def _periodic(h):
    h[0, 1:-1], h[-1, 1:-1] = h[-2, 1:-1], h[1, 1:-1]
    h[1:-1, 0], h[1:-1, -1] = h[1:-1, -2], h[1:-1, 1]


@pytest.mark.parametrize("grid_type", [0, 4])
def test_update_atmosphere_state_applies_tendencies(grid_type):
    # UpdateAtmosphereState with apply_tendencies: fill_gfs_delp,
    # prepare_tendencies_and_update_tracers, then ApplyPhysicsToDycore on the
    # updated tracers, delp and tendencies; compute-domain u_dt / v_dt whose
    # halo comes from a callable (cubed sphere: planted neighbour values) or
    # is filled periodically (grid_type 4)
    import torch
    from oracle import update_numpy as U
    from pyshield_amd.update import UpdateAtmosphereState
    ni, nj, nk, dt = 8, 8, 6, 225.0
    print(f"update+dycore chain grid_type={grid_type}", flush=True)
    s, g = _case(ni, nj, nk, seed=21 + grid_type)
    rng = np.random.default_rng(5)
    prsi = np.concatenate([np.full((ni, nj, 1), 300.0), 300.0 + np.cumsum(s["delp"], axis=2)], axis=2)
    upd = {n: s[n] * rng.uniform(0.9, 1.1, s[n].shape) for n in Q}
    upd["qvapor"][0, 0, 2] = -1e-7            # one cell for fill_gfs_delp to fill
    upd_pt, upd_ua, upd_va = s["pt"] + rng.normal(0, 0.5, s["pt"].shape), s["ua"] + 0.1, s["va"] - 0.1
    u_dt0, v_dt0 = s["u_dt"][1:-1, 1:-1].copy(), s["v_dt"][1:-1, 1:-1].copy()
    pt_dt0 = s["t_dt"]
    edge = {"w": rng.normal(0, 1e-3, (2, nj, nk)), "e": rng.normal(0, 1e-3, (2, nj, nk)),
            "s": rng.normal(0, 1e-3, (2, ni, nk)), "n": rng.normal(0, 1e-3, (2, ni, nk))}

    def planted(h, c):
        h[0, 1:-1], h[-1, 1:-1] = edge["w"][c], edge["e"][c]
        h[1:-1, 0], h[1:-1, -1] = edge["s"][c], edge["n"][c]

    # the oracle chain
    qv1 = U.fill_gfs_delp(s["delp"], upd["qvapor"], 1e-9)
    r = U.prepare_tendencies_and_update_tracers(
        u_dt0, v_dt0, pt_dt0, upd_ua, upd_va, upd_pt, qv1, upd["qliquid"], upd["qrain"], upd["qsnow"],
        upd["qice"], upd["qgraupel"], s["ua"], s["va"], s["pt"], s["qvapor"], s["qliquid"], s["qrain"],
        s["qsnow"], s["qice"], s["qgraupel"], prsi, s["delp"], 1.0 / dt)
    u_dt1, v_dt1, pt_dt1 = r[:3]
    q2 = dict(zip(("qvapor", "qliquid", "qrain", "qsnow", "qice", "qgraupel"), r[3:9]))
    delp2 = r[9]
    if grid_type > 3:
        hu, hv = _halo(u_dt1, _periodic), _halo(v_dt1, _periodic)
    else:
        hu, hv = _halo(u_dt1, lambda h: planted(h, 0)), _halo(v_dt1, lambda h: planted(h, 1))
    s2 = dict(s, delp=delp2, t_dt=pt_dt1, u_dt=hu, v_dt=hv, **q2)
    want = _oracle(s2, g, dt, ni + 1, nj + 1, grid_type)

    # the GPU chain
    names = Q + ("pt", "delp", "pe", "peln", "pk", "ps", "ua", "va", "u", "v")
    dyn = types.SimpleNamespace(**{n: _t(s[n]) for n in names})
    phy = types.SimpleNamespace(
        physics_updated_specific_humidity=_t(upd["qvapor"]), physics_updated_qliquid=_t(upd["qliquid"]),
        physics_updated_qrain=_t(upd["qrain"]), physics_updated_qsnow=_t(upd["qsnow"]),
        physics_updated_qice=_t(upd["qice"]), physics_updated_qgraupel=_t(upd["qgraupel"]),
        physics_updated_pt=_t(upd_pt), physics_updated_ua=_t(upd_ua), physics_updated_va=_t(upd_va),
        ua=_t(s["ua"]), va=_t(s["va"]), pt=_t(s["pt"]), prsi=_t(prsi))
    halo_update = None
    if grid_type <= 3:
        te = {k: _t(v) for k, v in edge.items()}

        def halo_update(u, v):
            for c, x in enumerate((u, v)):
                x[0, 1:-1], x[-1, 1:-1] = te["w"][c], te["e"][c]
                x[1:-1, 0], x[1:-1, -1] = te["s"][c], te["n"][c]
    grid_info = types.SimpleNamespace(grid_type=grid_type, **{n: _t(x) for n, x in _flat_grid(g).items()})
    nml = types.SimpleNamespace(dt_atmos=dt, npx=ni + 1, npy=nj + 1)
    op = UpdateAtmosphereState(namelist=nml, grid_info=grid_info, apply_tendencies=True, halo_update=halo_update)
    u_dt, v_dt, pt_dt = _t(u_dt0), _t(v_dt0), _t(pt_dt0)
    op(dyn, phy, u_dt, v_dt, pt_dt, dt)
    torch.cuda.synchronize()
    got = {n: getattr(dyn, n) for n in ("pt", "pe", "peln", "pk", "ps", "u", "v")}
    got.update(t_dt=pt_dt, u_srf=op._apply_physics_to_dycore.u_srf, v_srf=op._apply_physics_to_dycore.v_srf)
    want = {n: want[n] for n in got}
    _check(got, want, f"chain grid_type {grid_type}")
    for n, x in q2.items():
        np.testing.assert_allclose(getattr(dyn, n).cpu().numpy(), x, rtol=1e-13, atol=1e-14 * np.abs(x).max(),
                                   err_msg=n)
    if grid_type <= 3:
        assert not u_dt.any() and not v_dt.any()   # set_winds_zero
    else:
        np.testing.assert_allclose(u_dt.cpu().numpy(), u_dt1, rtol=1e-13, atol=1e-14 * np.abs(u_dt1).max())
