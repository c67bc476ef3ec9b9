"""The state-update kernels against the NumPy restatement
(oracle/update_numpy.py): prepare_microphysics,
update_physics_state_with_tendencies, fill_gfs_delp and
prepare_tendencies_and_update_tracers.  Point-wise stencils must match to
1e-15 relative (they are single FP64 expressions); fill_gfs_delp bit for bit
on integer-free inputs up to 1 ulp."""

import numpy as np
import pytest

from oracle import update_numpy as U

pytestmark = pytest.mark.gpu

NI, NJ, NK = 5, 7, 79


def _t(x):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda:0").clone()


def _halo_view(x, h=2):
    """x placed inside a larger NaN-filled buffer, returned as a strided view."""
    import torch
    ni, nj, nk = x.shape
    big = torch.full((ni + 2 * h, nj + 2 * h, nk + 3), float("nan"), dtype=torch.float64, device="cuda:0")
    big[h:h + ni, h:h + nj, :nk] = torch.as_tensor(x, device="cuda:0")
    return big[h:h + ni, h:h + nj, :nk]


@pytest.mark.parametrize("strided", [False, True])
def test_prepare_microphysics(strided):
    import torch
    from pyshield_amd.update import prepare_microphysics
    rng = np.random.default_rng(1)
    phii = np.cumsum(rng.uniform(100, 3000, (NI, NJ, NK + 1))[..., ::-1], axis=-1)[..., ::-1] * 9.8
    omga, qv = rng.normal(0, 1, (NI, NJ, NK)), rng.uniform(0, 0.02, (NI, NJ, NK))
    pt, delp = rng.uniform(200, 300, (NI, NJ, NK)), rng.uniform(100, 2000, (NI, NJ, NK))
    mk = _halo_view if strided else _t
    dz, wmp = mk(np.zeros((NI, NJ, NK))), mk(np.zeros((NI, NJ, NK)))
    tend = [mk(np.full((NI, NJ, NK), 7.0)) for _ in range(10)]
    prepare_microphysics(dz, mk(phii), wmp, mk(omga), mk(qv), mk(pt), mk(delp), *tend)
    torch.cuda.synchronize()
    rdz, rw = U.prepare_microphysics(phii, omga, qv, pt, delp)
    np.testing.assert_allclose(dz.cpu().numpy(), rdz, rtol=1e-15, atol=0)
    np.testing.assert_allclose(wmp.cpu().numpy(), rw, rtol=1e-15, atol=0)
    for t in tend:
        assert torch.all(t == 0.0)


def test_update_physics_state_with_tendencies():
    import torch
    from pyshield_amd.update import update_physics_state_with_tendencies
    rng = np.random.default_rng(2)
    ins = [rng.uniform(0, 1, (NI, NJ, NK)) for _ in range(10)]
    tds = [rng.normal(0, 1e-3, (NI, NJ, NK)) for _ in range(10)]
    outs = [_t(np.zeros((NI, NJ, NK))) for _ in range(10)]
    update_physics_state_with_tendencies(*[_t(x) for x in ins], *[_t(x) for x in tds], *outs, 225.0)
    torch.cuda.synchronize()
    ref = U.update_physics_state_with_tendencies(ins, tds, 225.0)
    # the GPU contracts q + t*dt into one FMA: agreement to an ulp of the inputs
    for o, r in zip(outs, ref):
        np.testing.assert_allclose(o.cpu().numpy(), r, rtol=1e-15, atol=1e-16)


# the default shape; ragged blocks of 12 columns (200 levels); halo'd storage
@pytest.mark.parametrize("shape,strided", [((NI, NJ, NK), False), ((9, 11, 200), False), ((NI, NJ, NK), True)])
def test_fill_gfs_delp(shape, strided):
    import torch
    from pyshield_amd.update import fill_gfs_delp
    rng = np.random.default_rng(3)
    q = rng.uniform(-2e-9, 5e-9, shape)
    q[0, 0, :] = -1e-8          # a column that stays negative down to the surface
    q[1, 1, 10:20] = 1e-3
    delp = rng.uniform(100, 2000, shape)
    mk = _halo_view if strided else _t
    tq = mk(q)
    fill_gfs_delp(mk(delp), tq, 1.0e-9)
    torch.cuda.synchronize()
    ref = U.fill_gfs_delp(delp, q, 1.0e-9)
    np.testing.assert_allclose(tq.cpu().numpy(), ref, rtol=1e-15, atol=1e-30)
    assert np.all(tq.cpu().numpy() >= 0.0)


@pytest.mark.parametrize("strided", [False, True])
def test_prepare_tendencies_and_update_tracers(strided):
    import torch
    from pyshield_amd.update import prepare_tendencies_and_update_tracers
    rng = np.random.default_rng(4)
    sh = (NI, NJ, NK)
    f = {n: rng.uniform(0, 1e-3, sh) for n in ("qv_t1", "ql_t1", "qr_t1", "qs_t1", "qi_t1", "qg_t1",
                                              "qv_t0", "ql_t0", "qr_t0", "qs_t0", "qi_t0", "qg_t0")}
    for n in ("u_dt", "v_dt", "pt_dt", "u_t1", "v_t1", "u_t0", "v_t0"):
        f[n] = rng.normal(0, 5, sh)
    f["pt_t1"], f["pt_t0"] = rng.uniform(200, 300, sh), rng.uniform(200, 300, sh)
    prsi = np.cumsum(np.concatenate([np.full((NI, NJ, 1), 300.0), rng.uniform(100, 2000, sh)], -1), -1)
    delp = prsi[..., 1:] - prsi[..., :-1] + rng.normal(0, 1, sh)
    mk = _halo_view if strided else _t
    t = {n: mk(v) for n, v in f.items()}
    tdelp, tprsi = mk(delp), mk(prsi)
    order = ("u_dt", "v_dt", "pt_dt", "u_t1", "v_t1", "pt_t1", "qv_t1", "ql_t1", "qr_t1", "qs_t1", "qi_t1",
             "qg_t1", "u_t0", "v_t0", "pt_t0", "qv_t0", "ql_t0", "qr_t0", "qs_t0", "qi_t0", "qg_t0")
    prepare_tendencies_and_update_tracers(*[t[n] for n in order], tprsi, tdelp, 1.0 / 225.0)
    torch.cuda.synchronize()
    ref = U.prepare_tendencies_and_update_tracers(*[f[n] for n in order], prsi, delp, 1.0 / 225.0)
    got = [t[n] for n in ("u_dt", "v_dt", "pt_dt", "qv_t0", "ql_t0", "qr_t0", "qs_t0", "qi_t0", "qg_t0")] + [tdelp]
    for g, r in zip(got, ref):
        np.testing.assert_allclose(g.cpu().numpy(), r, rtol=1e-14, atol=1e-300)
