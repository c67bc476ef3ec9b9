"""The benchmark workloads at their full sizes, beyond one forward call.

* C384 (6 x 384 x 384 columns x 79 levels, dt 150 s): ten consecutive calls,
  the state advanced between calls by update_physics_state_with_tendencies
  (physics.py:146-194) and the cloud-fraction accumulator carried.  Every call
  is checked against the oracle on sampled columns of the state the GPU has
  reached, the mixed-precision call against the FP64 call on the same state
  over every cell (1e-3 K and 1e-5 of the vapour scale, except cells the
  oracle's own FP32 envelope shows to sit on a threshold), and water
  conservation over every column; the GPU chain
  and an oracle chain started from the same sampled columns are compared
  after ten calls.
* C48 (6 x 48 x 48 x 79): the FP64 adjoint over the whole domain, against
  autograd of the oracle's torch transcription on sampled columns (for the
  gradients of the column-mean updated T, qv and condensates and of surface
  precipitation, and for random output cotangents), the dot-product test
  <J v, w> = <v, J^T w> over the whole domain with J v from forward-mode AD
  of the same transcription, and central finite differences of the FP64
  forward (step 1e-6 relative) at points at least 1e-3 (relative) from every
  branch threshold, picked from the oracle's branch records.  For the mixed
  forward: finite differences at 1e-3 against the FP64 adjoint.
* 768 x 768 x 127 (dt 90 s): the adjoint over the whole domain against
  autograd on sampled columns, with its workspace and its time relative to
  the FP64 forward reported.
"""

import numpy as np
import pytest

from pyshield_amd._config import PhysicsConfig

import branch_margins as BM
from helpers import PRECIP, TEND, WATER, norm_err, run_gpu, run_oracle
from test_gpu_adjoint import OUTS, STATE, _oracle_vjp, oracle_jvp_dot
from test_gpu_configs import _column_water_imbalance, _load, explain_cells

pytestmark = pytest.mark.gpu

UPDATED = ("qv", "ql", "qr", "qi", "qs", "qg", "qa", "pt", "ua", "va")
UPDATED_TEND = ("qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt", "qa_dt", "pt_dt", "udt", "vdt")
FP64_TOL = 1e-9           # normalised max error of a call's tendencies, FP64 vs oracle
VJP_TOL = 1e-7            # normalised max error of input cotangents vs autograd
DOT_TOL = 1e-10           # dot-product identity, relative


def _take(f, cols):
    """Host copy of the sampled columns of every field, [n, 1, nz] / [n, 1]."""
    import torch
    idx = torch.as_tensor(cols, device=next(iter(f.values())).device)
    out = {}
    for n, t in f.items():
        flat = t.reshape((-1,) + tuple(t.shape[2:]))
        out[n] = flat[idx].cpu().numpy()[:, None]
    return out


def test_c384_ten_consecutive_calls():
    import torch
    from pyshield_amd import _lib
    from pyshield_amd.microphysics import forward
    from pyshield_amd.update import update_physics_state_with_tendencies
    print("c384: loading", flush=True)
    w, f, cols, sample = _load("c384", nsample=192, seed=11)
    dt = w.timestep
    cfg = PhysicsConfig(dt_atmos=int(dt), npz=w.nz)
    params = _lib.setup(cfg.mp_namelist(), dt)
    params.accumulate = 0
    # the mixed call writes its own tendencies, accumulator and precipitation
    fm = dict(f)
    for n in TEND + ("qaz",):
        fm[n] = torch.empty_like(f[n])
    for n in PRECIP:
        fm[n] = torch.empty_like(f[n])
    chain = {n: sample[n].copy() for n in sample}
    chain_qaz = np.zeros_like(sample["pt"])
    worst = {"fp64": 0.0, "mixed_T": 0.0}
    for call in range(10):
        state = _take({n: f[n] for n in STATE + ("land", "area", "qaz")}, cols)
        qaz0 = f["qaz"].clone()
        fm["qaz"].copy_(f["qaz"])
        forward(fm, params, "mixed")
        forward(f, params, "fp64")
        torch.cuda.synchronize()
        for mode, g in (("fp64", f), ("mixed", fm)):
            imb = _column_water_imbalance(g, dt).max().item()
            assert imb <= 1e-12, (call, mode, imb)
        dT3 = ((fm["pt_dt"] - f["pt_dt"]) * dt).abs()
        dq3 = ((fm["qv_dt"] - f["qv_dt"]) * dt).abs()
        qscale = (f["qv_dt"] * dt).abs().max().item()
        beyond = (dT3 > 1e-3) | (dq3 > 1e-5 * qscale)
        ncols, bad = explain_cells(f, qaz0, beyond, {"pt": fm["pt_dt"], "qv": fm["qv_dt"]},
                                   {"pt": f["pt_dt"], "qv": f["qv_dt"]}, dt, cfg.mp_namelist())
        dT, dq = dT3.max().item(), dq3.max().item()
        nbeyond = int(beyond.sum().item())
        del dT3, dq3, beyond, qaz0
        # the GPU call against the oracle on the state the GPU chain reached
        ora = run_oracle({n: state[n] for n in STATE + ("land", "area")}, cfg, dt, qaz=state["qaz"])
        got = _take({n: f[n] for n in TEND + PRECIP}, cols)
        errs = {n: norm_err(got[n], ora[n], floor=1e-30) for n in TEND}
        worst["fp64"] = max(worst["fp64"], max(errs.values()))
        worst["mixed_T"] = max(worst["mixed_T"], dT)
        print(f"call {call + 1}: fp64 vs oracle {max(errs.values()):.1e}; mixed vs fp64 max |dT| {dT:.2e} K, "
              f"qv {dq / qscale:.1e} of scale; {nbeyond} cells beyond 1e-3 K / 1e-5 in {ncols} columns, "
              f"{len(bad)} not accounted for by the FP32 envelope", flush=True)
        assert all(v <= FP64_TOL for v in errs.values()), (call, errs)
        assert not bad, (call, bad[:10])
        assert ncols <= 1e-5 * w.ncol, (call, ncols)
        # advance both chains by their own tendencies
        st = [f[n] for n in UPDATED]
        update_physics_state_with_tendencies(*st, *[f[n] for n in UPDATED_TEND], *st, dt)
        oc = run_oracle({n: chain[n] for n in STATE + ("land", "area")}, cfg, dt, qaz=chain_qaz)
        for x, xt in zip(UPDATED, UPDATED_TEND):
            chain[x] = chain[x] + oc[xt] * dt
        chain_qaz = oc["qaz"]
    torch.cuda.synchronize()
    final = _take({n: f[n] for n in UPDATED}, cols)
    dev = {n: norm_err(final[n], chain[n], floor=1e-30) for n in UPDATED}
    print("after 10 calls, GPU chain vs oracle chain (normalised max difference):",
          {k: f"{v:.1e}" for k, v in dev.items()}, flush=True)
    assert float(np.max(np.abs(final["pt"] - chain["pt"]))) <= 1e-3
    for q in WATER:
        assert dev[q] <= 1e-5, (q, dev[q])


def _structured_cotangents(nz, dt, n):
    """Cotangents of the column-mean updated T, qv and condensates and of the
    four surface precipitation rates: mean_k (X + X_dt dt) is linear in X_dt
    with weight dt / nz."""
    rng = np.random.default_rng(5)
    cot = {}
    for name in OUTS:
        shape = (n, 1) if name in PRECIP else (n, 1, nz)
        if name in ("pt_dt", "qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt"):
            a = rng.normal(size=(n, 1, 1)) * dt / nz
            cot[name] = np.broadcast_to(a, shape).copy()
        elif name in PRECIP:
            cot[name] = rng.normal(size=shape)
        else:
            cot[name] = np.zeros(shape)
    return cot


def _gpu_adjoint(f, cot_dev, params, waves=2048):
    import torch
    from pyshield_amd.microphysics import adjoint
    g = {n: torch.full_like(f["pt"], np.nan) for n in STATE + ("qaz",)}
    g["params"] = torch.zeros(6, dtype=torch.float64, device=f["pt"].device)
    ws = adjoint(f, cot_dev, g, params, waves=waves)
    return g, ws


def _scatter(cot_cols, cols, f):
    """Whole-domain cotangents that are cot_cols on the sampled columns and 0
    elsewhere."""
    import torch
    out = {}
    idx = torch.as_tensor(cols, device=f["pt"].device)
    for n, v in cot_cols.items():
        ref = f["land"] if n in PRECIP else f["pt"]
        t = torch.zeros_like(ref)
        flat = t.reshape((-1,) + tuple(ref.shape[2:]))
        flat[idx] = torch.as_tensor(v.reshape((len(cols),) + tuple(ref.shape[2:])), device=ref.device)
        out[n] = t
    return out


def _check_vjp(g, cols, sample, qaz_s, cfg, dt, cot_cols, label):
    ora = _oracle_vjp(sample, qaz_s, cfg, dt, cot_cols)
    got = _take({n: g[n] for n in STATE + ("qaz",)}, cols)
    errs = {n: norm_err(got[n], ora[n], floor=1e-300) for n in STATE + ("qaz",)}
    gp = g["params"].cpu().numpy()
    perr = np.abs(gp - ora["params"]) / np.maximum(np.abs(ora["params"]), 1e-300)
    print(label, "adjoint vs autograd:", {k: f"{v:.1e}" for k, v in errs.items()},
          "params", [f"{x:.1e}" for x in perr], flush=True)
    assert all(v <= VJP_TOL for v in errs.values()), errs
    assert np.all(perr <= VJP_TOL), perr
    # columns without an output cotangent get exactly zero input cotangents
    import torch
    mask = torch.ones(g["pt"].shape[0] * g["pt"].shape[1], dtype=torch.bool, device=g["pt"].device)
    mask[torch.as_tensor(cols, device=mask.device)] = False
    for n in STATE:
        assert g[n].reshape(mask.numel(), -1)[mask].abs().max().item() == 0.0, n


def _fd_levels(sample, rng):
    """One level per sampled column, inside a cloud where the column has one."""
    cond = sum(sample[q][:, 0] for q in ("ql", "qr", "qi", "qs", "qg"))
    levels = []
    for c in range(cond.shape[0]):
        wet = np.flatnonzero(cond[c] > 1e-6)
        levels.append(int(rng.choice(wet)) if wet.size else int(rng.integers(cond.shape[1])))
    return levels


def _fd_candidates(sample, levels, rng, per_col=3):
    """(column, field, level) at each column's level, for fields with a
    nonzero value there."""
    out = []
    for c, k in enumerate(levels):
        fields = [n for n in STATE if abs(sample[n][c, 0, k]) > 1e-10]
        for field in rng.choice(fields, size=min(per_col, len(fields)), replace=False):
            out.append((c, str(field), k))
    return out


def _local_cotangents(sample, levels, rng):
    """Cotangents of the ten tendencies at one level per column (zero
    elsewhere, and on the precipitation): <out, w> then depends on the
    perturbed cell and the cells above it, and its rounding stays small."""
    cot = {n: np.zeros(sample["land" if n in PRECIP else "pt"].shape) for n in OUTS}
    for c, k in enumerate(levels):
        for n in TEND:
            cot[n][c, 0, k] = rng.normal()
    return cot


# the state field whose magnitude sets the rounding of each tendency
_STATE_OF = {"pt_dt": "pt", "qv_dt": "qv", "ql_dt": "ql", "qr_dt": "qr", "qi_dt": "qi", "qs_dt": "qs",
             "qg_dt": "qg", "qa_dt": "qa", "udt": "ua", "vdt": "va"}


def _loss(out, cot_cols, c, column, dt):
    """<outputs, w of sampled column c> for each member of an FD pair, the
    sum of |w out| and the sum of |w| |X| / dt over the tendencies (a
    tendency is a difference of states over dt, so its rounding scales with
    the state X)."""
    tot, mag_out, mag_state = 0.0, 0.0, 0.0
    for n in OUTS:
        wn = cot_cols[n][c][None]
        term = out[n] * wn
        tot = tot + term.reshape(term.shape[0], -1).sum(-1)
        mag_out = mag_out + np.abs(term).reshape(term.shape[0], -1).sum(-1).max()
        if n in _STATE_OF:
            mag_state = mag_state + float((np.abs(wn[0]) * np.abs(column[_STATE_OF[n]]) / dt).sum())
    return tot, mag_out, mag_state


def _finite_differences(sample, qaz_s, cfg, dt, cot_cols, cands, rel_step, mode):
    """Central differences of <out, w> along each candidate's input, from one
    GPU forward over 2 columns per candidate."""
    n = len(cands)
    batch = {k: np.concatenate([sample[k][[c]] for c, _, _ in cands for _ in (0, 1)]) for k in sample}
    qz = np.concatenate([qaz_s[[c]] for c, _, _ in cands for _ in (0, 1)])
    h = np.empty(n)
    for i, (c, field, k) in enumerate(cands):
        x = sample[field][c, 0, k]
        h[i] = rel_step * abs(x)
        batch[field][2 * i, 0, k] = x + h[i]
        batch[field][2 * i + 1, 0, k] = x - h[i]
    out = run_gpu(batch, cfg, dt, mode, qaz=qz)
    fd, n_out, n_state = np.empty(n), np.empty(n), np.empty(n)
    for i, (c, _, _) in enumerate(cands):
        pair = {k: out[k][2 * i:2 * i + 2] for k in OUTS}
        tot, mag_out, mag_state = _loss(pair, cot_cols, c, {k: sample[k][c] for k in sample}, dt)
        fd[i] = (tot[0] - tot[1]) / (2 * h[i])
        n_out[i], n_state[i] = mag_out / (2 * h[i]), mag_state / (2 * h[i])
    return fd, n_out, n_state


def test_c48_adjoint_whole_domain():
    import torch
    from pyshield_amd import _lib
    from pyshield_amd.synthetic import make_workload
    print("c48: loading", flush=True)
    w, f, cols, sample = _load("c48", nsample=24, seed=13)
    dt = w.timestep
    cfg = PhysicsConfig(dt_atmos=int(dt), npz=w.nz)
    params = _lib.setup(cfg.mp_namelist(), dt)
    params.accumulate = 0
    rng = np.random.default_rng(17)
    qaz_all = rng.uniform(0.0, 1.0, f["pt"].shape)
    f["qaz"].copy_(torch.as_tensor(qaz_all))
    qaz_s = qaz_all.reshape(-1, w.nz)[cols][:, None]
    # 1. sampled columns against autograd of the oracle: the column means and
    # precipitation, then random cotangents
    for label, cot_cols in (("column means + precipitation", _structured_cotangents(w.nz, dt, len(cols))),
                            ("random", {n: rng.normal(size=sample["land" if n in PRECIP else "pt"].shape)
                                        for n in OUTS})):
        g, _ = _gpu_adjoint(f, _scatter(cot_cols, cols, f), params)
        torch.cuda.synchronize()
        _check_vjp(g, cols, sample, qaz_s, cfg, dt, cot_cols, label)
    # 2. the dot-product identity over the whole domain
    full = make_workload("c48")
    wcot = {n: rng.normal(size=full["land" if n in PRECIP else "pt"].shape) for n in OUTS}
    v = {n: rng.normal(size=full[n].shape) * (np.abs(full[n]) + 1e-12) for n in STATE}
    vq = rng.normal(size=full["pt"].shape) * 1e-3
    vp = rng.normal(size=6) * 1e-2
    dev = {n: torch.as_tensor(x, device="cuda:0") for n, x in wcot.items()}
    g, _ = _gpu_adjoint(f, dev, params)
    torch.cuda.synchronize()
    gh = {n: g[n].cpu().numpy() for n in STATE + ("qaz",)}
    v_jtw = (sum(float((gh[n] * v[n]).sum()) for n in STATE) + float((gh["qaz"] * vq).sum())
             + float((g["params"].cpu().numpy() * vp).sum()))
    jv_w = 0.0
    for t in range(w.ntiles):          # the tangent-linear tile by tile (columns are independent)
        sl = slice(t * w.ni, (t + 1) * w.ni)
        part = {n: full[n][sl] for n in full}
        jv_w += oracle_jvp_dot(part, qaz_all[sl], cfg, dt, {n: v[n][sl] for n in STATE}, vq[sl], vp,
                               {n: wcot[n][sl] for n in OUTS})
    rel = abs(jv_w - v_jtw) / abs(jv_w)
    print(f"whole-domain dot product: <Jv,w> {jv_w:.15e} <v,J^T w> {v_jtw:.15e} rel {rel:.1e}", flush=True)
    assert rel <= DOT_TOL
    # 3. finite differences at points 1e-3 (relative) from every threshold
    levels = _fd_levels(sample, rng)
    cot_cols = _local_cotangents(sample, levels, rng)
    g, _ = _gpu_adjoint(f, _scatter(cot_cols, cols, f), params)
    torch.cuda.synchronize()
    gh = {n: g[n].reshape(-1, w.nz)[torch.as_tensor(cols, device="cuda:0")].cpu().numpy() for n in STATE}
    nml = cfg.mp_namelist()
    flat = {k: sample[k][:, 0] for k in sample if k in STATE + ("land", "area")}
    for mode, step, margin, tol in (("fp64", 1e-6, 1e-3, 1e-6), ("mixed", 1e-3, 2e-3, 2e-2)):
        cands = _fd_candidates(sample, levels, np.random.default_rng(23 if mode == "fp64" else 29))
        safe = BM.safe_points(nml, dt, flat, cands, margin)
        pts = [c for c, s in zip(cands, safe) if s]
        print(f"{mode}: {len(pts)} of {len(cands)} candidate points are {margin:g} from every threshold",
              flush=True)
        assert len(pts) >= 24
        fd, n_out, n_state = _finite_differences(sample, qaz_s, cfg, dt, cot_cols, pts, step, mode)
        adj = np.array([gh[field][c, k] for c, field, k in pts])
        # rounding of <out, w> over the step 2h: FP64 rounding of the states
        # the tendencies are differences of, and of the outputs (in mixed
        # precision the rate laws' FP32 rounding, a few ulp of the outputs)
        noise = 1e-14 * n_state + (1e-14 if mode == "fp64" else 3e-7) * n_out
        bound = tol * np.abs(adj) + noise
        err = np.abs(fd - adj)
        # the points whose derivative stands well above the rounding of the
        # difference quotient: there the relative error is what is measured
        sharp = np.abs(adj) > 1e3 * noise
        rel_err = err[sharp] / np.abs(adj[sharp])
        print(f"{mode}: finite differences (step {step:g}) vs FP64 adjoint at {len(pts)} points: "
              f"{int(sharp.sum())} with |dL/dx| > 1e3 x rounding, worst relative error there "
              f"{rel_err.max():.1e}, median {np.median(rel_err):.1e}; worst {np.max(err / bound):.2f} x bound overall",
              flush=True)
        assert sharp.sum() >= 12
        bad = [(pts[i], fd[i], adj[i], bound[i]) for i in np.flatnonzero(err > bound)]
        assert not bad, bad[:5]


def test_deep768_adjoint_whole_domain():
    import torch
    from pyshield_amd import _lib
    from pyshield_amd.microphysics import forward
    print("deep768: loading", flush=True)
    w, f, cols, sample = _load("deep768", nsample=8, seed=19)
    dt = w.timestep
    cfg = PhysicsConfig(dt_atmos=int(dt), npz=w.nz)
    params = _lib.setup(cfg.mp_namelist(), dt)
    params.accumulate = 0
    rng = np.random.default_rng(31)
    qaz_s = np.zeros_like(sample["pt"])
    cot_cols = {n: rng.normal(size=sample["land" if n in PRECIP else "pt"].shape) for n in OUTS}
    cot = _scatter(cot_cols, cols, f)
    g, ws = _gpu_adjoint(f, cot, params)
    torch.cuda.synchronize()
    _check_vjp(g, cols, sample, qaz_s, cfg, dt, cot_cols, "deep768")
    # cost: the adjoint against the FP64 forward on the same state
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record()
    forward(f, params, "fp64")
    ev[1].record()
    from pyshield_amd.microphysics import adjoint
    ev[2].record()
    adjoint(f, cot, g, params, workspace=ws, waves=2048)
    ev[3].record()
    torch.cuda.synchronize()
    fwd_ms, adj_ms = ev[0].elapsed_time(ev[1]), ev[2].elapsed_time(ev[3])
    fields = 8 * f["pt"].numel()
    print(f"deep768: adjoint workspace {ws.numel() * 8 / 2 ** 20:.1f} MiB for 2048 wavefronts "
          f"(one 127-level column each, 77 recorded values per cell and sub-step); cotangents "
          f"{(10 + 13) * fields / 2 ** 30:.1f} GiB; FP64 forward {fwd_ms:.2f} ms, adjoint {adj_ms:.2f} ms "
          f"({adj_ms / fwd_ms:.2f}x the forward, which it replays)", flush=True)
    assert ws.numel() * 8 <= 2048 * (77 * 128 + 6) * 8
