"""Throughput of the GFDL microphysics forward on one MI355X.

A step is one Microphysics call (pySHiELD/stencils/microphysics.py:1981-2263)
over the whole workload, launched through the C ABI (shield_mp_forward), with
the state resident in HBM.  Calls are consecutive: between two calls the
state is advanced by the first call's tendencies on the device
(update_physics_state_with_tendencies, physics.py:146-194, outside the timed
kernel events) and the cloud-fraction accumulator carries over, so the
branch mix evolves as in a run.  Tendencies are written, not accumulated, which is the
Physics driver's use (prepare_microphysics zeroes them before the call,
physics.py:133-143); the zeroing is fused into the kernel.

Default workload: C384 (6 x 384 x 384 columns x 79 levels, dt 150 s),
mixed precision.  The JSON line also carries the plain FP64 kernel's rate on
the same state, the dominant kernel's HBM roofline fraction (algorithmic
bytes / HIP-event kernel time / 8 TB/s), the FP64 adjoint's time, the fused
microphysics + state update against its separate kernels, the
ApplyPhysicsToDycore kernels' and the physics prologue's times on one tile
(384 x 384 x 79 at C384) and the CPU baseline (the C oracle on the host cores, on a bounded sample of the
same workload).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_CELL = 8 * (13 + 11)   # FP64 fields read (12 state + qaz) and written (10 tendencies + qaz)
B_COL = 8 * (2 + 4)      # land, area in; 4 surface precipitation fields out


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c384")
    ap.add_argument("--mode", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-compare", action="store_true", help="skip the second precision")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the FP64 adjoint timing")
    ap.add_argument("--adjoint-steps", type=int, default=2)
    ap.add_argument("--no-dycore", action="store_true",
                    help="skip the ApplyPhysicsToDycore and physics-prologue timings")
    ap.add_argument("--only-kernel", action="store_true",
                    help="launch only the timed kernel (for profiler runs)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="instead of the benchmark, copy one buffer of this many bytes "
                         "(calibrates the HBM byte counters of a profiler run)")
    return ap.parse_args()


def copy_calibration(nbytes):
    import torch
    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 8, "note": "one device-to-device copy; reads and writes n bytes"}))


def load_state(name, device):
    import torch
    from pyshield_amd.synthetic import WORKLOADS, make_state, WORKLOAD_IDS
    w = WORKLOADS[name]
    ni, nj, nz = w.ni * w.ntiles, w.nj, w.nz
    names3 = ("pt", "qv", "ql", "qr", "qi", "qs", "qg", "qa", "ua", "va", "delp", "delz", "w")
    f = {n: torch.empty((ni, nj, nz), dtype=torch.float64, device=device) for n in names3}
    for n in ("land", "area"):
        f[n] = torch.empty((ni, nj), dtype=torch.float64, device=device)
    for t in range(w.ntiles):
        off = (-10.0 + 4.0 * t) if w.tile_offsets else 0.0
        st = make_state(w.ni, w.nj, w.nz, 1000 * WORKLOAD_IDS[name] + t, ptop=w.ptop, t_offset=off,
                        cloud_top=w.cloud_top, median=w.median, qg_median=w.qg_median)
        sl = slice(t * w.ni, (t + 1) * w.ni)
        for n in names3 + ("land", "area"):
            f[n][sl].copy_(torch.from_numpy(st[n]))
        del st
    for n in ("pt_dt", "qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt", "qa_dt", "udt", "vdt", "qaz"):
        f[n] = torch.zeros((ni, nj, nz), dtype=torch.float64, device=device)
    for n in ("rain", "snow", "ice", "graupel"):
        f[n] = torch.zeros((ni, nj), dtype=torch.float64, device=device)
    return w, f


STATE_UPDATED = ("qv", "ql", "qr", "qi", "qs", "qg", "qa", "pt", "ua", "va")
STATE_TENDS = ("qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt", "qa_dt", "pt_dt", "udt", "vdt")


def carry_state(f, dt):
    """Advance the state by the call's tendencies in place, X += X_dt dt
    (update_physics_state_with_tendencies, physics.py:146-194), so the next
    call sees the state this one left."""
    from pyshield_amd.update import update_physics_state_with_tendencies
    st = [f[n] for n in STATE_UPDATED]
    update_physics_state_with_tendencies(*st, *[f[n] for n in STATE_TENDS], *st, dt)


def time_mode(f, params, mode, steps, warmup, carry=True):
    """`steps` consecutive microphysics calls after `warmup` untimed ones.
    Each call's kernel is timed with HIP events on its launch stream; with
    `carry` the state is then advanced by the call's tendencies (outside the
    events) so the branch mix evolves from call to call.  Returns the wall
    time of the timed loop (updates included), the per-call kernel times in
    ms and the per-call update times in ms."""
    import torch
    from pyshield_amd.microphysics import forward
    dt = params.timestep
    for _ in range(warmup):
        forward(f, params, mode)
        if carry:
            carry_state(f, dt)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b, c in ev:
        a.record()
        forward(f, params, mode)
        b.record()
        if carry:
            carry_state(f, dt)
        c.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b, _ in ev]
    update_ms = [b.elapsed_time(c) for _, b, c in ev]
    return wall, kernel_ms, update_ms


def time_adjoint(f, params, steps):
    """The FP64 adjoint (which replays the forward of every column) for unit
    output cotangents; returns the kernel times in ms (HIP events) and the
    allocated bytes."""
    import torch
    from pyshield_amd.microphysics import AD_OUT_NAMES, AD_STATE_NAMES, adjoint
    ref3, ref2 = f["pt"], f["land"]
    cot = {n: torch.ones_like(ref2 if n in ("rain", "snow", "ice", "graupel") else ref3) for n in AD_OUT_NAMES}
    grads = {n: torch.empty_like(ref3) for n in AD_STATE_NAMES + ("qaz",)}
    grads["params"] = torch.zeros(6, dtype=torch.float64, device=ref3.device)
    ws = adjoint(f, cot, grads, params, waves=2048)  # warm-up
    torch.cuda.synchronize()
    ms = []
    for _ in range(steps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        adjoint(f, cot, grads, params, workspace=ws, waves=2048)
        b.record()
        torch.cuda.synchronize()
        ms.append(a.elapsed_time(b))
    return ms


def cpu_baseline(name, seconds):
    """The C oracle built -O3 -march=native on this host, OpenMP over columns
    with the threads bound (OMP_PROC_BIND=close, OMP_PLACES=cores), on a
    bounded sample of tile 0 of the workload (oracle/cpu_baseline.py, run as
    a child process so the binding takes effect before OpenMP starts)."""
    import subprocess
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    env = dict(os.environ, OMP_NUM_THREADS=str(nthreads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", "--workload", name, "--seconds", str(seconds)],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed:\n{r.stdout}\n{r.stderr}")
    return json.loads(r.stdout.strip().splitlines()[-1])


# ApplyPhysicsToDycore per cell: the column step reads 9 layer fields and
# writes pt, t_dt and the 3 interface fields; the wind step reads and writes
# u, v, reads u_dt, v_dt and clears them (FP64)
B_DYCORE_CELL = 8 * (9 + 2 + 3 + 4 + 2 + 2)


def time_dycore(n, nk, steps):
    """shield_apply_physics_to_dycore on one n x n tile of nk levels
    (synthetic fields and grid metrics; its three kernels, HIP events on the
    launch stream); returns the times in ms."""
    import torch
    from pyshield_amd.dycore import GRID_NAMES, apply_physics_to_dycore
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(3)

    def u(shape, lo, hi):
        return torch.rand(shape, generator=g, device=dev, dtype=torch.float64) * (hi - lo) + lo

    f = {q: u((n, n, nk), 0.0, 2e-3) for q in ("qvapor", "qliquid", "qrain", "qsnow", "qice", "qgraupel")}
    f.update(pt=u((n, n, nk), 200, 300), t_dt=u((n, n, nk), -1e-4, 1e-4), delp=u((n, n, nk), 50, 2000),
             ua=u((n, n, nk), -10, 10), va=u((n, n, nk), -10, 10), u=u((n, n + 1, nk), -10, 10),
             v=u((n + 1, n, nk), -10, 10), u_dt=u((n + 2, n + 2, nk), -1e-3, 1e-3),
             v_dt=u((n + 2, n + 2, nk), -1e-3, 1e-3))
    for q in ("pe", "peln", "pk"):
        f[q] = torch.full((n, n, nk + 1), 300.0, dtype=torch.float64, device=dev)
    for q in ("ps", "u_srf", "v_srf"):
        f[q] = torch.zeros((n, n), dtype=torch.float64, device=dev)
    shapes = {**{q: (n + 2, n + 2) for q in GRID_NAMES[:6]}, **{q: (n, n + 1) for q in GRID_NAMES[6:9]},
              **{q: (n + 1, n) for q in GRID_NAMES[9:12]}, **{q: (n,) for q in GRID_NAMES[12:]}}
    grid = {q: u(shp, -1, 1) if not q.startswith("edge") else u(shp, 0, 1) for q, shp in shapes.items()}
    apply_physics_to_dycore(f, grid, 150.0, n + 1, n + 1)
    torch.cuda.synchronize()
    ms = []
    for _ in range(steps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        apply_physics_to_dycore(f, grid, 150.0, n + 1, n + 1)
        b.record()
        torch.cuda.synchronize()
        ms.append(a.elapsed_time(b))
    return ms


# physics prologue per cell: reads delz, delp, 7 tracers, qsgs_tke, pt, omga
# (12) and writes delp, 7 tracers, qsgs_tke, dm, delprsi, del_gz, phil, dz,
# wmp and 10 zeroed tendencies (24), plus 3 interface fields (FP64)
B_PROLOGUE_CELL = 8 * (12 + 24 + 3)


def time_prologue(n, nk, steps):
    """shield_physics_prologue on one n x n tile of nk levels (synthetic
    state, restored before each call outside the timed region); returns the
    times in ms."""
    import torch
    from pyshield_amd.physics import SPECIES, physics_prologue
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(5)

    def u(lo, hi, shape=(n, n, nk)):
        return torch.rand(shape, generator=g, device=dev, dtype=torch.float64) * (hi - lo) + lo

    src = {"delz": u(-600.0, -20.0), "delp": u(50.0, 2000.0), "qsgs_tke": u(0.0, 1.0), "pt": u(200.0, 300.0),
           "omga": u(-1.0, 1.0)}
    for sp in SPECIES:
        src[sp] = u(1e-6, 2e-2) if sp == SPECIES[0] else u(0.0, 1e-3)
    work = {k: v.clone() for k, v in src.items()}
    out = {k: torch.zeros((n, n, nk), dtype=torch.float64, device=dev)
           for k in ("dm", "delprsi", "del_gz", "phil", "dz", "wmp")}
    iface = {k: torch.zeros((n, n, nk + 1), dtype=torch.float64, device=dev) for k in ("phii", "prsi", "prsik")}
    tend = [torch.zeros((n, n, nk), dtype=torch.float64, device=dev) for _ in range(10)]

    def call():
        physics_prologue(work["delz"], work["delp"], {sp: work[sp] for sp in SPECIES}, work["qsgs_tke"],
                         work["pt"], iface["phii"], iface["prsi"], iface["prsik"], out["dm"], out["delprsi"],
                         out["del_gz"], out["phil"], 300.0, omga=work["omga"], dz=out["dz"], wmp=out["wmp"],
                         tend=tend)

    call()
    ms = []
    for _ in range(steps):
        for k, v in src.items():
            work[k].copy_(v)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        call()
        b.record()
        torch.cuda.synchronize()
        ms.append(a.elapsed_time(b))
    return ms


def time_fused(f, params, mode, dt, steps):
    """The microphysics call followed by the Physics / UpdateAtmosphereState
    state update (forward Euler, fill_gfs_delp, the tracer conversion): the
    fused kernel (shield_mp_forward_update) and the four separate launches,
    on the bench state (synthetic update inputs); returns (fused ms,
    separate ms), HIP events around each sequence."""
    import torch
    from pyshield_amd import _lib
    from pyshield_amd.microphysics import forward, forward_update
    from pyshield_amd.update import (fill_gfs_delp, prepare_tendencies_and_update_tracers,
                                     update_physics_state_with_tendencies)
    ref = f["pt"]
    ni, nj, nk = ref.shape
    z = lambda: torch.zeros_like(ref)  # noqa: E731
    prsi = torch.empty((ni, nj, nk + 1), dtype=torch.float64, device=ref.device)
    prsi[..., 0] = 300.0
    prsi[..., 1:] = 300.0 + torch.cumsum(f["delp"], dim=2)
    upd = {"updated": {n: z() for n in _lib.UPDATED_NAMES}, "u_dt": z(), "v_dt": z(), "pt_dt": z(),
           "tracers": {n: f[src].clone() for n, src in zip(_lib.TRACER_NAMES, ("qv", "ql", "qr", "qs", "qi", "qg"))},
           "delp": f["delp"].clone(), "prsi": prsi}
    src = ("qv", "ql", "qr", "qi", "qs", "qg", "qa", "pt", "ua", "va")
    tend = ("qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt", "qa_dt", "pt_dt", "udt", "vdt")

    def fused():
        forward_update(f, upd, params, dt=dt, rdt=1.0 / dt, mode=mode)

    def separate():
        forward(f, params, mode)
        u = upd["updated"]
        update_physics_state_with_tendencies(*[f[n] for n in src], *[f[n] for n in tend],
                                             *[u[n] for n in _lib.UPDATED_NAMES], dt)
        fill_gfs_delp(upd["delp"], u["qvapor"], 1.0e-9)
        tr = upd["tracers"]
        prepare_tendencies_and_update_tracers(
            upd["u_dt"], upd["v_dt"], upd["pt_dt"], u["ua"], u["va"], u["pt"], u["qvapor"], u["qliquid"],
            u["qrain"], u["qsnow"], u["qice"], u["qgraupel"], f["ua"], f["va"], f["pt"],
            *[tr[n] for n in _lib.TRACER_NAMES], upd["prsi"], upd["delp"], 1.0 / dt)

    out = []
    for run in (fused, separate):
        run()
        torch.cuda.synchronize()
        ms = []
        for _ in range(steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            run()
            b.record()
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        out.append(float(np.mean(ms)))
    return out


def time_fill(n, nk, steps):
    """fill_gfs_delp on one n x n tile of nk levels (vapour near q_min so the
    fills run; restored before each call); returns the times in ms."""
    import torch
    from pyshield_amd.update import fill_gfs_delp
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(7)
    q0 = torch.rand((n, n, nk), generator=g, device=dev, dtype=torch.float64) * 3e-9 - 1e-9
    delp = torch.rand((n, n, nk), generator=g, device=dev, dtype=torch.float64) * 1900.0 + 100.0
    q = q0.clone()
    fill_gfs_delp(delp, q, 1.0e-9)
    ms = []
    for _ in range(steps):
        q.copy_(q0)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fill_gfs_delp(delp, q, 1.0e-9)
        b.record()
        torch.cuda.synchronize()
        ms.append(a.elapsed_time(b))
    return ms


def main():
    args = parse()
    if args.copy_bytes:
        copy_calibration(args.copy_bytes)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the microphysics shards by columns with no "
              f"exchange, so --gpus {args.gpus} is not run here", file=sys.stderr)
        sys.exit(2)
    import torch
    from pyshield_amd import _lib
    from pyshield_amd._config import PhysicsConfig

    dev = torch.device("cuda:0")
    w, f = load_state(args.workload, dev)
    ncol = w.ni * w.nj * w.ntiles
    cfg = PhysicsConfig(dt_atmos=int(w.timestep), npz=w.nz)
    params = _lib.setup(cfg.mp_namelist(), w.timestep)
    params.accumulate = 0
    print(f"# workload {w.name}: {ncol} columns x {w.nz} levels, dt {w.timestep} s, "
          f"layout {_lib.layout(w.nz, args.mode, params.ntimes)}", file=sys.stderr, flush=True)

    wall, kms, ums = time_mode(f, params, args.mode, args.steps, args.warmup)
    # a step is one microphysics call on the state the previous call left;
    # the forward-Euler update between calls is reported beside it
    k_ms = float(np.mean(kms))
    ms_per_step = k_ms
    value = ncol / (k_ms * 1e-3)
    alg_bytes = ncol * (w.nz * B_CELL + B_COL)
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9
    print(f"# {args.mode}: {args.steps} consecutive calls, kernel {k_ms:.3f} ms "
          f"(min {min(kms):.3f}, max {max(kms):.3f}), {achieved:.0f} GB/s; state update {np.mean(ums):.3f} ms; "
          f"wall {1e3 * wall / args.steps:.3f} ms/step", file=sys.stderr, flush=True)

    extra = {}
    if not args.no_compare and not args.only_kernel:
        other = "fp64" if args.mode == "mixed" else "mixed"
        wall2, kms2, _ = time_mode(f, params, other, max(2, args.steps // 2), 1, carry=False)
        k2 = float(np.mean(kms2))
        extra = {f"{other}_columns_per_s": ncol / (k2 * 1e-3),
                 f"{other}_kernel_ms": k2,
                 "mixed_over_fp64": (k2 / k_ms) if args.mode == "mixed" else (k_ms / k2)}
        print(f"# {other}: kernel {k2:.3f} ms", file=sys.stderr, flush=True)

    if not args.no_adjoint and not args.only_kernel:
        p64 = _lib.setup(cfg.mp_namelist(), w.timestep)
        p64.accumulate = 0
        ams = float(np.mean(time_adjoint(f, p64, args.adjoint_steps)))
        # forward inputs re-read (13 F), output cotangents read (10 F), input
        # cotangents written (13 F); per column land, area and 4 precipitation cotangents
        adj_bytes = ncol * (w.nz * 8 * (13 + 10 + 13) + 8 * 6)
        extra.update({"adjoint_kernel_ms": ams, "adjoint_columns_per_s_fp64": ncol / (ams * 1e-3),
                      "fwd_plus_adjoint_columns_per_s": ncol / ((k_ms + ams) * 1e-3),
                      "adjoint_algorithmic_bytes": adj_bytes,
                      "adjoint_roofline_frac": adj_bytes / (ams * 1e-3) / 1e9 / HBM_PEAK_GBS})
        fp64_ms = k_ms if args.mode == "fp64" else extra.get("fp64_kernel_ms")
        if fp64_ms:
            # the adjoint's cost in units of the FP64 forward of the same call
            extra["adjoint_over_fp64_forward"] = ams / fp64_ms
        print(f"# adjoint (fp64): kernel {ams:.3f} ms", file=sys.stderr, flush=True)

    if not args.no_dycore and not args.only_kernel:
        dms = float(np.mean(time_dycore(w.ni, w.nz, 5)))
        dbytes = w.ni * w.nj * w.nz * B_DYCORE_CELL
        extra.update({"dycore_tile_ms": dms, "dycore_algorithmic_bytes": dbytes,
                      "dycore_roofline_frac": dbytes / (dms * 1e-3) / 1e9 / HBM_PEAK_GBS})
        print(f"# apply_physics_to_dycore ({w.ni}x{w.nj}x{w.nz} tile): {dms:.3f} ms", file=sys.stderr, flush=True)

    if not args.no_dycore and not args.only_kernel:
        fu_ms, sep_ms = time_fused(f, params, args.mode, w.timestep, 3)
        extra.update({"fused_update_ms": fu_ms, "separate_update_ms": sep_ms})
        print(f"# microphysics + state update: fused {fu_ms:.3f} ms, separate kernels {sep_ms:.3f} ms",
              file=sys.stderr, flush=True)
        fms = float(np.mean(time_fill(w.ni, w.nz, 5)))
        extra.update({"fill_gfs_delp_tile_ms": fms})
        print(f"# fill_gfs_delp ({w.ni}x{w.nj}x{w.nz} tile): {fms:.3f} ms", file=sys.stderr, flush=True)
        pms = float(np.mean(time_prologue(w.ni, w.nz, 5)))
        pbytes = w.ni * w.nj * w.nz * B_PROLOGUE_CELL
        extra.update({"prologue_tile_ms": pms, "prologue_algorithmic_bytes": pbytes,
                      "prologue_roofline_frac": pbytes / (pms * 1e-3) / 1e9 / HBM_PEAK_GBS})
        print(f"# physics_prologue ({w.ni}x{w.nj}x{w.nz} tile): {pms:.3f} ms", file=sys.stderr, flush=True)

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"mp_forward_{w.name}_{args.mode}_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline and not args.only_kernel:
        del f
        torch.cuda.empty_cache()
        cpu = cpu_baseline(args.workload, args.cpu_seconds)

    line = {
        "metric": "column-level updates/s, fwd+adjoint, mixed vs FP64 on 1 MI355X at C384×79",
        "value": value,
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.mode == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: {w.ntiles}x{w.ni}x{w.nj} columns x {w.nz} levels, "
                               f"dt {w.timestep:g} s, {args.mode} forward",
                   "columns": ncol, "levels": w.nz, "timestep": w.timestep, "mode": args.mode},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "kernel_ms": k_ms,
        "kernel_ms_per_call": kms,
        "state_update_ms": float(np.mean(ums)),
        "wall_ms_per_step_with_update": 1e3 * wall / args.steps,
        "algorithmic_bytes_per_launch": alg_bytes,
        "adjoint_columns_per_s": extra.get("adjoint_columns_per_s_fp64"),
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
