"""The hand-written adjoint (shield_mp_adjoint) against:

* torch autograd through the FP64 torch transcription of the reference column
  code (oracle/mp_numpy.py under TorchOps): the vector-Jacobian product for
  random output cotangents, for the twelve input fields, qaz and the six
  process parameters, at one and two sub-steps;
* the tangent-linear / adjoint dot-product identity <J v, w> = <v, J^T w>,
  with J v from forward-mode AD of the same FP64 transcription (to 1e-10) and
  from central finite differences of the FP64 GPU forward;
* the torch.autograd.Function wrapper.
"""

import numpy as np
import pytest

from oracle import mp_numpy as M
from pyshield_amd._config import PhysicsConfig

from helpers import norm_err, small_state

pytestmark = pytest.mark.gpu

STATE = ("pt", "qv", "ql", "qr", "qi", "qs", "qg", "qa", "ua", "va", "delp", "delz")
OUTS = ("pt_dt", "qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt", "qa_dt", "udt", "vdt",
        "rain", "snow", "ice", "graupel")
PARAMS = ("fac_rc", "vr_fac", "vs_fac", "vg_fac", "vi_fac", "qi_gen")
# normalised max error of each input cotangent against autograd
VJP_TOL = 1e-7


def _cotangents(st, seed):
    rng = np.random.default_rng(seed)
    ni, nj, nz = st["pt"].shape
    return {n: rng.normal(size=(ni, nj) if n in ("rain", "snow", "ice", "graupel") else (ni, nj, nz))
            for n in OUTS}


def _oracle_vjp(st, qaz, cfg, dt, cot):
    """Input cotangents and parameter gradient by torch autograd of the oracle."""
    import torch
    flat = {k: (v.reshape(-1, v.shape[-1]) if v.ndim == 3 else v.reshape(-1)) for k, v in st.items()}
    nml = cfg.mp_namelist()
    ts = {k: torch.tensor(v, dtype=torch.float64, requires_grad=k in STATE) for k, v in flat.items()}
    tq = torch.tensor(qaz.reshape(flat["pt"].shape), dtype=torch.float64, requires_grad=True)
    pn = {k: torch.tensor(float(nml[k]), dtype=torch.float64, requires_grad=True) for k in PARAMS[1:]}
    fac_rc = torch.tensor(M.derived(nml, dt)["fac_rc"], dtype=torch.float64, requires_grad=True)
    out = M.run(dict(nml, **pn), dt, ts, qaz=tq, xp=M.TorchOps(), derived_overrides={"fac_rc": fac_rc})
    loss = 0.0
    for n in OUTS:
        loss = loss + (out[n] * torch.tensor(cot[n].reshape(out[n].shape))).sum()
    loss.backward()
    shp = st["pt"].shape

    def grad(t):
        # an input the switches take out of the call gets no gradient
        return np.zeros(shp) if t.grad is None else t.grad.numpy().reshape(shp)

    g = {n: grad(ts[n]) for n in STATE}
    g["qaz"] = grad(tq)
    g["params"] = np.array([0.0 if p.grad is None else float(p.grad) for p in [fac_rc] + [pn[k] for k in PARAMS[1:]]])
    return g


def oracle_jvp_dot(st, qaz, cfg, dt, v, vq, vp, cot):
    """<J v, w>: the tangent-linear of the oracle's torch transcription by
    forward-mode AD along (v, vq, vp) -- the twelve input fields, qaz and the
    six process parameters -- dotted with the output cotangents w = cot."""
    import torch
    import torch.autograd.forward_ad as fwAD
    flat = {k: (x.reshape(-1, x.shape[-1]) if x.ndim == 3 else x.reshape(-1)) for k, x in st.items()}
    nml = cfg.mp_namelist()
    t64 = lambda x: torch.tensor(x, dtype=torch.float64)  # noqa: E731
    with fwAD.dual_level():
        ts = {k: fwAD.make_dual(t64(x), t64(v[k].reshape(x.shape))) if k in STATE else t64(x) for k, x in flat.items()}
        shp = flat["pt"].shape
        tq = fwAD.make_dual(t64(qaz.reshape(shp)), t64(vq.reshape(shp)))
        pn = {k: fwAD.make_dual(t64(float(nml[k])), t64(vp[i + 1])) for i, k in enumerate(PARAMS[1:])}
        fac_rc = fwAD.make_dual(t64(M.derived(nml, dt)["fac_rc"]), t64(vp[0]))
        out = M.run(dict(nml, **pn), dt, ts, qaz=tq, xp=M.TorchOps(), derived_overrides={"fac_rc": fac_rc})
        total = 0.0
        for n in OUTS:
            tan = fwAD.unpack_dual(out[n]).tangent
            if tan is not None:
                total += float((tan.numpy() * cot[n].reshape(tan.shape)).sum())
    return total


def _directions(st, seed):
    rng = np.random.default_rng(seed)
    v = {n: rng.normal(size=st[n].shape) * (np.abs(st[n]) + 1e-12) for n in STATE}
    return v, rng.normal(size=st["pt"].shape) * 1e-3, rng.normal(size=6) * 1e-2


def _gpu_fields(st, qaz):
    import torch
    dev = torch.device("cuda:0")

    def t(x):
        return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev).clone()

    f = {n: t(st[n]) for n in STATE + ("land", "area")}
    f["qaz"] = t(qaz)
    f["w"] = torch.zeros_like(f["pt"])
    return f


def _gpu_vjp(st, qaz, cfg, dt, cot):
    import torch
    from pyshield_amd import _lib
    from pyshield_amd.microphysics import adjoint
    f = _gpu_fields(st, qaz)
    params = _lib.setup(cfg.mp_namelist(), dt)
    params.accumulate = 0
    c = {n: torch.as_tensor(np.ascontiguousarray(v), dtype=torch.float64, device="cuda:0") for n, v in cot.items()}
    g = {n: torch.full_like(f["pt"], np.nan) for n in STATE + ("qaz",)}
    g["params"] = torch.zeros(6, dtype=torch.float64, device="cuda:0")
    adjoint(f, c, g, params, waves=64)
    torch.cuda.synchronize()
    return {n: v.cpu().numpy() for n, v in g.items()}


# (dt, seed, ni, nj, nz): 79 levels run 2 columns per wavefront in 3 slots (9
# columns leave a last wavefront with one), 40 levels 3 columns in 2 slots
@pytest.mark.parametrize("dt,seed,ni,nj,nz", [(150.0, 1000, 2, 3, 79), (225.0, 2001, 2, 3, 79),
                                              (150.0, 3003, 3, 3, 79), (150.0, 4004, 5, 1, 40)])
def test_adjoint_matches_autograd_of_oracle(dt, seed, ni, nj, nz):
    print(f"case dt={dt} seed={seed} shape=({ni}, {nj}, {nz})", flush=True)
    st = small_state(ni, nj, nz, seed=seed)
    qaz = np.random.default_rng(seed).uniform(0.0, 1.0, st["pt"].shape)
    cfg = PhysicsConfig(dt_atmos=int(dt), npz=nz)
    cot = _cotangents(st, seed + 1)
    ref = _oracle_vjp(st, qaz, cfg, dt, cot)
    got = _gpu_vjp(st, qaz, cfg, dt, cot)
    errs = {n: norm_err(got[n], ref[n], floor=1e-300) for n in STATE + ("qaz",)}
    perr = np.abs(got["params"] - ref["params"]) / np.maximum(np.abs(ref["params"]), 1e-300)
    print({k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    print("params", dict(zip(PARAMS, [f"{x:.2e}/{y:.2e}" for x, y in zip(got["params"], ref["params"])])), flush=True)
    assert all(np.isfinite(got[n]).all() for n in STATE + ("qaz",))
    assert all(v <= VJP_TOL for v in errs.values()), errs
    assert np.all(perr <= VJP_TOL), perr


# every namelist switch the adjoint supports, away from its default
SWITCHES = [
    {"do_qa": True},
    {"const_vi": True, "const_vs": True, "const_vg": True, "const_vr": True},
    {"irain_f": 1},
    {"z_slope_liq": False},
    {"do_sedi_heat": False, "sedi_transport": False},
    {"fast_sat_adj": True, "rad_snow": False, "rad_rain": False},
    {"reset_qaz": True},
    {"z_slope_ice": True},
    {"z_slope_ice": True, "fix_negative": True},
]


@pytest.mark.parametrize("switches", SWITCHES, ids=lambda d: "+".join(sorted(d)))
def test_adjoint_switches_match_autograd_of_oracle(switches):
    dt, seed = 150.0, 6006
    print(f"case switches={switches}", flush=True)
    st = small_state(2, 3, 79, seed=seed)
    qaz = np.random.default_rng(seed).uniform(0.0, 1.0, st["pt"].shape)
    cfg = PhysicsConfig(dt_atmos=int(dt), npz=79, **switches)
    cot = _cotangents(st, seed + 1)
    ref = _oracle_vjp(st, qaz, cfg, dt, cot)
    got = _gpu_vjp(st, qaz, cfg, dt, cot)
    errs = {n: norm_err(got[n], ref[n], floor=1e-300) for n in STATE + ("qaz",)}
    perr = np.abs(got["params"] - ref["params"]) / np.maximum(np.abs(ref["params"]), 1e-300)
    print({k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert all(np.isfinite(got[n]).all() for n in STATE + ("qaz",))
    assert all(v <= VJP_TOL for v in errs.values()), errs
    # a parameter the switches take out of the call has no gradient
    live = np.abs(ref["params"]) > 0.0
    assert np.all(perr[live] <= VJP_TOL), perr
    assert np.all(np.abs(got["params"][~live]) == 0.0), got["params"]


def negative_species_state(seed=8008):
    """small_state with negative species in some cells: every link of
    fix_negative's borrowing chain (ice from snow, snow from graupel, graupel
    from rain, rain from cloud water, cloud water from vapour), negative
    vapour that borrows down a column, and a negative surface layer that
    borrows from the layer above."""
    st = small_state(2, 3, 79, seed=seed)
    rng = np.random.default_rng(seed)
    shp = st["pt"].shape
    for q, frac in (("qi", 0.08), ("qs", 0.06), ("qg", 0.06), ("qr", 0.06), ("ql", 0.06)):
        m = rng.random(shp) < frac
        st[q] = np.where(m, -rng.uniform(1e-7, 5e-5, shp), st[q])
    # vapour: a negative run in column (0, 0) above level 40, negative surface layers
    st["qv"][0, 0, 38:42] = -rng.uniform(1e-6, 1e-5, 4)
    st["qv"][1, 2, 70] = -2e-6
    st["qv"][0, 1, -1] = -3e-6
    st["qv"][1, 1, -1] = -1e-3  # more than the layer above can lend
    return st


def test_adjoint_fix_negative_matches_autograd_of_oracle():
    dt, seed = 150.0, 8008
    st = negative_species_state(seed)
    qaz = np.random.default_rng(seed).uniform(0.0, 1.0, st["pt"].shape)
    cfg = PhysicsConfig(dt_atmos=int(dt), npz=79, fix_negative=True)
    cot = _cotangents(st, seed + 1)
    ref = _oracle_vjp(st, qaz, cfg, dt, cot)
    got = _gpu_vjp(st, qaz, cfg, dt, cot)
    errs = {n: norm_err(got[n], ref[n], floor=1e-300) for n in STATE + ("qaz",)}
    perr = np.abs(got["params"] - ref["params"]) / np.maximum(np.abs(ref["params"]), 1e-300)
    print({k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert all(np.isfinite(got[n]).all() for n in STATE + ("qaz",))
    assert all(v <= VJP_TOL for v in errs.values()), errs
    assert np.all(perr <= VJP_TOL), perr


def test_adjoint_sedi_w_matches_autograd_of_oracle():
    """do_sedi_w (non-hydrostatic): the vertical velocity is carried with the
    falling mass of every fall; w's cotangent after the call is an input of
    the adjoint and w's cotangent before it an output."""
    import torch
    from pyshield_amd import _lib
    from pyshield_amd.microphysics import adjoint
    dt, seed = 150.0, 9009
    st = small_state(2, 3, 79, seed=seed)
    rng = np.random.default_rng(seed)
    st["w"] = rng.normal(0.0, 0.5, st["pt"].shape)
    qaz = rng.uniform(0.0, 1.0, st["pt"].shape)
    cfg = PhysicsConfig(dt_atmos=int(dt), npz=79, hydrostatic=False, do_sedi_w=True)
    cot = _cotangents(st, seed + 1)
    cw = rng.normal(size=st["pt"].shape)
    # oracle: autograd through the torch transcription, w an input and an output
    flat = {k: (v.reshape(-1, v.shape[-1]) if v.ndim == 3 else v.reshape(-1)) for k, v in st.items()}
    nml = cfg.mp_namelist()
    ts = {k: torch.tensor(v, dtype=torch.float64, requires_grad=k in STATE + ("w",)) for k, v in flat.items()}
    tq = torch.tensor(qaz.reshape(flat["pt"].shape), dtype=torch.float64, requires_grad=True)
    out = M.run(nml, dt, ts, qaz=tq, xp=M.TorchOps())
    loss = (out["w"] * torch.tensor(cw.reshape(out["w"].shape))).sum()
    for n in OUTS:
        loss = loss + (out[n] * torch.tensor(cot[n].reshape(out[n].shape))).sum()
    loss.backward()
    ref = {n: ts[n].grad.numpy().reshape(st["pt"].shape) for n in STATE + ("w",)}
    # GPU adjoint
    f = _gpu_fields(st, qaz)
    f["w"] = torch.as_tensor(st["w"], dtype=torch.float64, device="cuda:0").clone()
    params = _lib.setup(nml, dt)
    params.accumulate = 0
    assert params.do_sedi_w == 1
    c = {n: torch.as_tensor(np.ascontiguousarray(v), dtype=torch.float64, device="cuda:0") for n, v in cot.items()}
    c["w"] = torch.as_tensor(cw, dtype=torch.float64, device="cuda:0")
    g = {n: torch.full_like(f["pt"], np.nan) for n in STATE + ("qaz", "w")}
    g["params"] = torch.zeros(6, dtype=torch.float64, device="cuda:0")
    w_in = f["w"].clone()
    adjoint(f, c, g, params, waves=64)
    torch.cuda.synchronize()
    assert torch.equal(f["w"], w_in)  # the adjoint leaves w alone
    got = {n: g[n].cpu().numpy() for n in STATE + ("w",)}
    errs = {n: norm_err(got[n], ref[n], floor=1e-300) for n in STATE + ("w",)}
    print({k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert all(np.isfinite(got[n]).all() for n in got)
    assert all(v <= VJP_TOL for v in errs.values()), errs
    # the w path is live: the transport moves w's gradient away from identity
    assert np.abs(ref["w"] - cw).max() > 1e-5 * np.abs(cw).max()


def test_adjoint_dot_product_with_finite_differences():
    """<J v, w> from central differences of the GPU FP64 forward equals
    <v, J^T w> from the adjoint, along smooth random directions."""
    import torch
    from pyshield_amd import _lib
    from pyshield_amd.microphysics import forward
    dt = 150.0
    st = small_state(2, 3, 79, seed=77)
    qaz = np.zeros(st["pt"].shape)
    cfg = PhysicsConfig(dt_atmos=int(dt), npz=79)
    cot = _cotangents(st, 5)
    got = _gpu_vjp(st, qaz, cfg, dt, cot)
    params = _lib.setup(cfg.mp_namelist(), dt)
    params.accumulate = 0
    rng = np.random.default_rng(11)
    results = []
    for trial in range(3):
        # a direction scaled to each field (relative perturbation of every value)
        v = {n: rng.normal(size=st[n].shape) * (np.abs(st[n]) + 1e-12) for n in STATE}

        def loss(h):
            s2 = dict(st)
            for n in STATE:
                s2[n] = st[n] + h * v[n]
            f = _gpu_fields(s2, qaz)
            ni, nj, nz = st["pt"].shape
            for n in OUTS:
                f[n] = torch.zeros((ni, nj) if n in ("rain", "snow", "ice", "graupel") else (ni, nj, nz),
                                   dtype=torch.float64, device="cuda:0")
            forward(f, params, "fp64")
            torch.cuda.synchronize()
            return sum(float((f[n].cpu().numpy() * cot[n]).sum()) for n in OUTS)

        h = 1e-7
        fd = (loss(h) - loss(-h)) / (2 * h)
        ad = sum(float((got[n] * v[n]).sum()) for n in STATE)
        results.append((fd, ad))
        print(f"direction {trial}: finite difference {fd:.10e} adjoint {ad:.10e} rel {abs(fd - ad) / abs(ad):.1e}",
              flush=True)
    # central differences at h = 1e-7 carry ~1e-6 relative truncation and
    # round-off; a process threshold crossed inside +-h shows as a larger gap
    assert sum(abs(fd - ad) / abs(ad) <= 1e-4 for fd, ad in results) >= 2, results


# the tangent-linear / adjoint identity in FP64: forward-mode AD of the oracle
# against the hand-written adjoint on the GPU, bar 1e-10 relative
DOT_TOL = 1e-10


@pytest.mark.parametrize("dt,seed", [(150.0, 77), (225.0, 78)])
def test_adjoint_dot_product_with_tangent_linear(dt, seed):
    """<J v, w> from forward-mode AD of the FP64 oracle equals <v, J^T w>
    from the adjoint, for random directions v over the input fields, qaz and
    the six parameters, and random output cotangents w."""
    st = small_state(2, 3, 79, seed=seed)
    qaz = np.random.default_rng(seed).uniform(0.0, 1.0, st["pt"].shape)
    cfg = PhysicsConfig(dt_atmos=int(dt), npz=79)
    cot = _cotangents(st, seed + 1)
    got = _gpu_vjp(st, qaz, cfg, dt, cot)
    for trial in range(2):
        v, vq, vp = _directions(st, 100 * seed + trial)
        jv_w = oracle_jvp_dot(st, qaz, cfg, dt, v, vq, vp, cot)
        v_jtw = (sum(float((got[n] * v[n]).sum()) for n in STATE) + float((got["qaz"] * vq).sum())
                 + float((got["params"] * vp).sum()))
        rel = abs(jv_w - v_jtw) / abs(jv_w)
        print(f"direction {trial}: <Jv,w> {jv_w:.15e} <v,J^T w> {v_jtw:.15e} rel {rel:.1e}", flush=True)
        assert rel <= DOT_TOL, (jv_w, v_jtw)


def test_autograd_function_matches_adjoint():
    import torch
    from pyshield_amd import _lib
    from pyshield_amd.microphysics import AD_OUT_NAMES, AD_STATE_NAMES, microphysics_ad
    dt = 150.0
    st = small_state(2, 2, 79, seed=31)
    qaz = np.zeros(st["pt"].shape)
    cfg = PhysicsConfig(dt_atmos=int(dt), npz=79)
    params = _lib.setup(cfg.mp_namelist(), dt)
    f = _gpu_fields(st, qaz)
    for n in AD_STATE_NAMES:
        f[n].requires_grad_(True)
    adp = torch.tensor([params.fac_rc, params.nml.vr_fac, params.nml.vs_fac, params.nml.vg_fac, params.nml.vi_fac,
                        params.nml.qi_gen], dtype=torch.float64, requires_grad=True)
    out = microphysics_ad(f, params, adp)
    cot = _cotangents(st, 3)
    loss = sum((out[n] * torch.as_tensor(cot[n], device="cuda:0")).sum() for n in AD_OUT_NAMES)
    loss.backward()
    ref = _gpu_vjp(st, qaz, cfg, dt, cot)
    for n in AD_STATE_NAMES:
        assert norm_err(f[n].grad.cpu().numpy(), ref[n], floor=1e-300) <= 1e-12, n
    np.testing.assert_allclose(adp.grad.numpy(), ref["params"], rtol=1e-12)


def test_autograd_function_on_halo_views():
    """The autograd path on fields given as interior views of halo-padded
    storages (the reference's Quantity layout): the same outputs and
    gradients as on dense tensors, bit for bit, and a second backward reuses
    the workspace."""
    import torch
    from pyshield_amd import _lib
    from pyshield_amd.microphysics import AD_OUT_NAMES, AD_STATE_NAMES, microphysics_ad
    dt = 150.0
    st = small_state(3, 4, 79, seed=37)
    cfg = PhysicsConfig(dt_atmos=int(dt), npz=79)
    params = _lib.setup(cfg.mp_namelist(), dt)
    ni, nj, nz = st["pt"].shape
    h, dev = 3, torch.device("cuda:0")

    def padded(x):
        shape = (ni + 2 * h + 1, nj + 2 * h + 1) + ((nz + 1,) if x.ndim == 3 else ())
        big = torch.full(shape, float("nan"), dtype=torch.float64, device=dev)
        view = big[h:h + ni, h:h + nj, :nz] if x.ndim == 3 else big[h:h + ni, h:h + nj]
        view.copy_(torch.as_tensor(x, device=dev))
        return view

    cot = _cotangents(st, 5)
    results = []
    for make in (lambda x: torch.as_tensor(np.ascontiguousarray(x), device=dev).clone(), padded):
        f = {n: make(st[n]) for n in AD_STATE_NAMES + ("land", "area")}
        f["qaz"] = make(np.zeros(st["pt"].shape))
        for n in AD_STATE_NAMES:
            f[n].requires_grad_(True)
        grads = []
        for _ in range(2):
            out = microphysics_ad(f, params)
            loss = sum((out[n] * torch.as_tensor(cot[n], device=dev)).sum() for n in AD_OUT_NAMES)
            g = torch.autograd.grad(loss, [f[n] for n in AD_STATE_NAMES])
            grads.append([x.detach().cpu().numpy() for x in g])
        assert all(np.array_equal(a, b) for a, b in zip(*grads))
        results.append(({n: out[n].detach().cpu().numpy() for n in AD_OUT_NAMES}, grads[0]))
    (o_dense, g_dense), (o_halo, g_halo) = results
    for n in AD_OUT_NAMES:
        assert np.array_equal(o_dense[n], o_halo[n]), n
    for n, a, b in zip(AD_STATE_NAMES, g_dense, g_halo):
        assert np.array_equal(a, b), n
