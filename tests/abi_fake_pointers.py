"""Argument checks of the C ABI that pass made-up device addresses.

Every entry point validates its arguments before it touches a device; these
checks call each one with fake pointers (0x1000, 0x200000, ...) and expect the
documented error code.  A call whose arguments all pass validation would
launch a kernel on those addresses, so this module runs only in a process
that sees no HIP device: `main` refuses to run the checks otherwise.
tests/test_abi.py starts it as a child process with the devices hidden.
Not collected by pytest (no test_ prefix).
"""

import ctypes
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from pyshield_amd import _lib  # noqa: E402
from pyshield_amd._config import PhysicsConfig  # noqa: E402

NO_DEVICE_EXIT = 5


def check_prologue_rejects_bad_arguments():
    L = _lib.lib()
    shape = _lib.ShieldMpShape(3, 4, 79, 79 * 4, 79, 0, 0)
    f = _lib.ShieldPrologueFields()
    names = [n for n, _ in f._fields_ if n not in ("q", "tend", "omga", "dz", "wmp", "si_i", "sj_i", "ptop")]
    for i, n in enumerate(names):
        setattr(f, n, 0x1000 * (i + 1))
    for i in range(7):
        f.q[i] = 0x100000 + 0x1000 * i
    f.si_i, f.sj_i, f.ptop = 80 * 4, 80, 300.0
    assert L.shield_physics_prologue(ctypes.byref(shape), None, None) == -1
    f.q[4] = None
    assert L.shield_physics_prologue(ctypes.byref(shape), ctypes.byref(f), None) == -1
    f.q[4] = 0x200000
    f.omga = 0x300000          # prepare_microphysics requested without its outputs
    assert L.shield_physics_prologue(ctypes.byref(shape), ctypes.byref(f), None) == -1
    f.omga = None
    f.ptop = 0.0
    assert L.shield_physics_prologue(ctypes.byref(shape), ctypes.byref(f), None) == -2
    f.ptop = 300.0
    assert L.shield_physics_prologue(ctypes.byref(shape), ctypes.byref(f), None) == -8


def _fake_update(base=0x100000, alias=None):
    u = _lib.ShieldUpdateFields()
    ptrs = iter(range(base, base + 0x1000 * 30, 0x1000))
    for i in range(10):
        u.updated[i] = next(ptrs)
    for i in range(6):
        u.tracers[i] = next(ptrs)
    u.u_dt, u.v_dt, u.pt_dt, u.delp, u.prsi = (next(ptrs) for _ in range(5))
    u.si_i, u.sj_i, u.dt, u.rdt, u.q_min = 80, 80 * 4, 150.0, 1.0 / 150.0, 1.0e-9
    if alias:
        alias(u)
    return u


def check_forward_update_rejects_bad_arguments():
    L = _lib.lib()
    p = _lib.setup(PhysicsConfig().mp_namelist(), 150.0)
    shape = _lib.ShieldMpShape(3, 4, 79, 79 * 4, 79, 4, 1)
    fwd = L.shield_mp_forward_update
    f = _fake_fields()
    assert fwd(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(f), None, 0, None) == -1
    missing = _fake_update()
    missing.tracers[3] = None
    assert fwd(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(f), ctypes.byref(missing), 0, None) == -1
    # an updated field written over a microphysics input
    def over_input(u):
        u.updated[2] = f.qr
    bad = _fake_update(alias=over_input)
    assert fwd(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(f), ctypes.byref(bad), 0, None) == -5
    # two update outputs sharing memory
    def twice(u):
        u.tracers[0] = u.delp
    bad = _fake_update(alias=twice)
    assert fwd(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(f), ctypes.byref(bad), 0, None) == -5
    bad = _fake_update()
    bad.dt = 0.0
    assert fwd(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(f), ctypes.byref(bad), 0, None) == -7
    # every check passes: without a GPU the call reports it
    ok = _fake_update()
    assert fwd(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(f), ctypes.byref(ok), 0, None) == -8


def _fake_fields(alias=None):
    f = _lib.ShieldMpFields()
    for i, n in enumerate(_lib.FIELD_NAMES):
        setattr(f, n, 0x1000 * (i + 1))
    if alias:
        setattr(f, alias[0], getattr(f, alias[1]))
    return f


def check_forward_rejects_before_launch():
    """Validation errors are returned before any device call, with the
    documented codes (these calls never launch: the fake pointers are caught)."""
    L = _lib.lib()
    p = _lib.setup(PhysicsConfig().mp_namelist(), 150.0)
    shape = _lib.ShieldMpShape(2, 2, 79, 2 * 79, 79, 2, 1)
    fwd = L.shield_mp_forward
    # unknown mode
    assert fwd(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(_fake_fields()), 7, None) == -4
    # nk out of range
    bad = _lib.ShieldMpShape(2, 2, 1, 2, 1, 2, 1)
    assert fwd(ctypes.byref(bad), ctypes.byref(p), ctypes.byref(_fake_fields()), 0, None) == -2
    bad = _lib.ShieldMpShape(2, 2, 400, 800, 400, 2, 1)
    assert fwd(ctypes.byref(bad), ctypes.byref(p), ctypes.byref(_fake_fields()), 0, None) == -2
    # missing input
    f = _fake_fields()
    f.pt = None
    assert fwd(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(f), 0, None) == -1
    # an output aliasing an input or another output
    assert fwd(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(_fake_fields(("qv_dt", "qv"))), 0, None) == -5
    assert fwd(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(_fake_fields(("udt", "vdt"))), 0, None) == -5
    # unsupported namelist switches
    p2 = _lib.setup(PhysicsConfig().mp_namelist(), 150.0)
    p2.nml.use_ppm = 1
    assert fwd(ctypes.byref(shape), ctypes.byref(p2), ctypes.byref(_fake_fields()), 0, None) == -3
    # empty domain is a no-op
    empty = _lib.ShieldMpShape(0, 4, 79, 0, 79, 0, 1)
    assert fwd(ctypes.byref(empty), ctypes.byref(p), ctypes.byref(_fake_fields()), 0, None) == 0


def _fake_adjoint():
    a = _lib.ShieldMpAdjointFields()
    for i, (n, _) in enumerate(a._fields_):
        setattr(a, n, 0x200000 + 0x1000 * i)
    return a


def check_adjoint_rejects_before_launch():
    """shield_mp_adjoint validates its arguments before any device call
    (these calls never launch: the fake pointers are caught); with valid
    arguments and no GPU it reports SHIELD_ERR_NOGPU."""
    L = _lib.lib()
    adj = L.shield_mp_adjoint
    p = _lib.setup(PhysicsConfig().mp_namelist(), 150.0)
    shape = _lib.ShieldMpShape(2, 2, 79, 2 * 79, 79, 2, 1)
    f, a, ws = _fake_fields(), _fake_adjoint(), ctypes.c_void_p(0x900000)

    def call(shape=shape, p=p, f=f, a=a, ws=ws, nbytes=1 << 30):
        return adj(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(f), ctypes.byref(a), ws, nbytes, None)

    assert call(ws=None) == -1
    a2 = _fake_adjoint()
    a2.d_pt_dt = None
    assert call(a=a2) == -1
    # do_sedi_w needs w
    p2 = _lib.setup(PhysicsConfig().mp_namelist(), 150.0)
    p2.do_sedi_w = 1
    f2 = _fake_fields()
    f2.w = None
    assert call(p=p2, f=f2) == -1
    # de_ice acts only on accumulated tendencies, which the adjoint does not take
    p2 = _lib.setup(PhysicsConfig().mp_namelist(), 150.0)
    p2.nml.de_ice = 1
    p2.accumulate = 1
    assert call(p=p2) == -9
    # nk above the adjoint's two slots
    assert call(shape=_lib.ShieldMpShape(2, 2, 200, 400, 200, 2, 1)) == -2
    # empty domain is a no-op
    assert call(shape=_lib.ShieldMpShape(0, 4, 79, 0, 79, 0, 1)) == 0
    # written cotangents sharing memory with a field read, with each other,
    # or with the workspace
    ref = _fake_adjoint()
    for written, address in (("g_qv", f.qv), ("g_delp", f.qaz), ("g_ql", ref.g_qi), ("params", ws.value),
                             ("g_pt", ref.d_pt_dt), ("g_qaz", ref.d_rain)):
        a2 = _fake_adjoint()
        setattr(a2, written, address)
        assert call(a=a2) == -5, written
    # every check passes: without a GPU the call reports it
    assert call() == -8
    # the variant without a workspace argument makes the same checks
    noalloc = L.shield_mp_adjoint_noalloc
    assert noalloc(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(f), None, None) == -1
    aliased = _fake_adjoint()
    aliased.g_qv = f.qv
    assert noalloc(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(f), ctypes.byref(aliased), None) == -5
    a3 = _fake_adjoint()
    a3.d_qv_dt = None
    assert noalloc(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(f), ctypes.byref(a3), None) == -1
    assert noalloc(ctypes.byref(_lib.ShieldMpShape(2, 2, 200, 400, 200, 2, 1)), ctypes.byref(p), ctypes.byref(f),
                   ctypes.byref(a), None) == -2
    assert noalloc(ctypes.byref(_lib.ShieldMpShape(0, 4, 79, 0, 79, 0, 1)), ctypes.byref(p), ctypes.byref(f),
                   ctypes.byref(a), None) == 0
    assert noalloc(ctypes.byref(shape), ctypes.byref(p), ctypes.byref(f), ctypes.byref(a), None) == -8
    # workspace: one tape of ntimes x 77 recorded fields x 192 cells (2
    # columns of 79 levels in 3 slots) per wavefront, plus six partial sums;
    # bounded by the number of column pairs
    per_wave = (1 * 77 * 192 + 6) * 8
    ws_bytes = L.shield_mp_adjoint_workspace_bytes
    assert ws_bytes(ctypes.byref(shape), ctypes.byref(p), 1) == per_wave
    assert ws_bytes(ctypes.byref(shape), ctypes.byref(p), 3) == 2 * per_wave
    assert ws_bytes(ctypes.byref(shape), ctypes.byref(p), 100) == 2 * per_wave
    # 127 levels: one column in 2 slots
    tall = _lib.ShieldMpShape(2, 2, 127, 2 * 127, 127, 2, 1)
    assert ws_bytes(ctypes.byref(tall), ctypes.byref(p), 100) == 4 * (1 * 77 * 128 + 6) * 8
    assert ws_bytes(ctypes.byref(_lib.ShieldMpShape(2, 2, 200, 400, 200, 2, 1)), ctypes.byref(p), 3) == 0


def _fake_dycore(ni=4, nj=4):
    f = _lib.ShieldDycoreFields()
    for i, n in enumerate(_lib.DYCORE_PTR_NAMES):
        setattr(f, n, 0x1000 * (i + 1))
    f.dt, f.con_cp, f.kappa = 150.0, 1004.6, 287.05 / 1004.6
    f.grid_type, f.npx, f.npy = 0, ni + 1, nj + 1
    return f


def check_dycore_struct_and_rejects():
    # 37 pointers, three doubles, four int32
    assert ctypes.sizeof(_lib.ShieldDycoreFields) == 37 * 8 + 3 * 8 + 4 * 4
    L = _lib.lib()
    call = L.shield_apply_physics_to_dycore
    shape = _lib.ShieldMpShape(4, 4, 10, 40, 10, 4, 1)
    assert call(None, ctypes.byref(_fake_dycore()), None) == -1
    f = _fake_dycore()
    f.delp = None
    assert call(ctypes.byref(shape), ctypes.byref(f), None) == -1
    f = _fake_dycore()
    f.edge_vect_n = None
    assert call(ctypes.byref(shape), ctypes.byref(f), None) == -1
    f = _fake_dycore()
    f.npx = 9                                   # one tile per rank: npx - 1 == ni
    assert call(ctypes.byref(shape), ctypes.byref(f), None) == -2
    assert call(ctypes.byref(_lib.ShieldMpShape(1, 4, 10, 40, 10, 4, 1)), ctypes.byref(_fake_dycore(1, 4)),
                None) == -2
    f = _fake_dycore()
    f.u = f.u_dt                                # the wind written over its tendency
    assert call(ctypes.byref(shape), ctypes.byref(f), None) == -5
    f = _fake_dycore()
    f.u_srf = f.ua
    assert call(ctypes.byref(shape), ctypes.byref(f), None) == -5
    f = _fake_dycore()
    f.dt = 0.0
    assert call(ctypes.byref(shape), ctypes.byref(f), None) == -7
    # an empty tile is a no-op
    assert call(ctypes.byref(_lib.ShieldMpShape(0, 4, 10, 40, 10, 4, 1)), ctypes.byref(_fake_dycore()), None) == 0
    # every check passes: without a GPU the call reports it
    assert call(ctypes.byref(shape), ctypes.byref(_fake_dycore()), None) == -8
    f = _fake_dycore()
    f.grid_type = 4
    for n in _lib.DYCORE_PTR_NAMES[21:]:
        setattr(f, n, None)                     # the doubly periodic form needs no grid
    f.npx = 100
    assert call(ctypes.byref(shape), ctypes.byref(f), None) == -8


def check_level_limits_of_the_staged_kernels():
    # the dycore update, the prologue and fill_gfs_delp stage whole columns in
    # LDS and reject columns taller than their stage
    L = _lib.lib()
    big = _lib.ShieldMpShape(4, 4, 4097, 4097 * 4, 4097, 4, 1)
    assert L.shield_apply_physics_to_dycore(ctypes.byref(big), ctypes.byref(_fake_dycore()), None) == -2
    f = _lib.ShieldPrologueFields()
    for i, n in enumerate(n for n, _ in f._fields_ if n not in ("q", "tend", "si_i", "sj_i", "ptop", "omga",
                                                                "dz", "wmp")):
        setattr(f, n, 0x1000 * (i + 1))
    for i in range(7):
        f.q[i] = 0x100000 + 0x1000 * i
    f.si_i, f.sj_i, f.ptop = 1026 * 4, 1026, 300.0
    tall = _lib.ShieldMpShape(3, 4, 1025, 1025 * 4, 1025, 0, 0)
    assert L.shield_physics_prologue(ctypes.byref(tall), ctypes.byref(f), None) == -2
    tall = _lib.ShieldMpShape(3, 4, 2049, 2049 * 4, 2049, 0, 0)
    assert L.shield_fill_gfs_delp(ctypes.byref(tall), ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000),
                                  ctypes.c_double(1e-9), None) == -2


CHECKS = [n for n in dir() if n.startswith("check_")]


def main():
    """Run every check; print one JSON object {check: error text or ""}."""
    if _lib.device_count() != 0:
        print("a HIP device is visible: the fake-pointer checks are not run", flush=True)
        sys.exit(NO_DEVICE_EXIT)
    import traceback
    out = {}
    for name in CHECKS:
        try:
            globals()[name]()
            out[name] = ""
        except Exception:
            out[name] = traceback.format_exc()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
