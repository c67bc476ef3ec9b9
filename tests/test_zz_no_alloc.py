"""shield_mp_adjoint_noalloc (the adjoint with its tapes in the library's
static device array) against shield_mp_adjoint with a caller workspace, on
two calls in a row, the second on a smaller domain.

Both launch the same kernel with one wavefront per group of columns when the
domain has fewer groups than either buffer holds tapes for, so every output,
the parameter gradient included, is equal bit for bit.  The second call
checks that what the first left in the shared array does not leak into it.
"""

import ctypes

import numpy as np
import pytest

from pyshield_amd._config import PhysicsConfig

from helpers import small_state

pytestmark = pytest.mark.gpu

STATE = ("pt", "qv", "ql", "qr", "qi", "qs", "qg", "qa", "ua", "va", "delp", "delz")
OUTS = ("pt_dt", "qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt", "qa_dt", "udt", "vdt",
        "rain", "snow", "ice", "graupel")


def _setup(ni, nj, seed):
    import torch
    from pyshield_amd import _lib
    dev = torch.device("cuda:0")
    st = small_state(ni, nj, 79, seed=seed)
    rng = np.random.default_rng(seed + 1)

    def t(x):
        return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev).clone()

    f = {n: t(st[n]) for n in STATE + ("land", "area")}
    f["qaz"] = t(rng.uniform(0.0, 1.0, st["pt"].shape))
    f["w"] = torch.zeros_like(f["pt"])
    cot = {n: t(rng.normal(size=(ni, nj) if n in ("rain", "snow", "ice", "graupel") else (ni, nj, 79)))
           for n in OUTS}
    params = _lib.setup(PhysicsConfig(dt_atmos=150, npz=79).mp_namelist(), 150.0)
    params.accumulate = 0
    return f, cot, params


def _grads(f):
    import torch
    g = {n: torch.full_like(f["pt"], np.nan) for n in STATE + ("qaz",)}
    g["params"] = torch.full((6,), np.nan, dtype=torch.float64, device=f["pt"].device)
    return g


def _noalloc(f, cot, g, params):
    import torch
    from pyshield_amd import _lib
    from pyshield_amd.microphysics import _mp_args
    shape, mf = _mp_args(f)
    a = _lib.ShieldMpAdjointFields()
    for n in OUTS:
        setattr(a, "d_" + n, cot[n].data_ptr())
    for n in STATE + ("qaz",):
        setattr(a, "g_" + n, g[n].data_ptr())
    a.params = g["params"].data_ptr()
    stream = torch.cuda.current_stream(f["pt"].device).cuda_stream
    rc = _lib.lib().shield_mp_adjoint_noalloc(ctypes.byref(shape), ctypes.byref(params), ctypes.byref(mf),
                                              ctypes.byref(a), ctypes.c_void_p(stream))
    _lib.check(rc, "shield_mp_adjoint_noalloc")


def test_noalloc_equals_workspace_adjoint_on_two_calls():
    import torch
    from pyshield_amd.microphysics import adjoint
    for ni, nj, seed in ((24, 24, 91), (6, 5, 92)):
        print(f"case {ni}x{nj}x79 seed {seed}", flush=True)
        f, cot, params = _setup(ni, nj, seed)
        g_ws = _grads(f)
        adjoint(f, cot, g_ws, params, waves=4096)
        g_na = _grads(f)
        _noalloc(f, cot, g_na, params)
        torch.cuda.synchronize()
        for n in STATE + ("qaz", "params"):
            a, b = g_ws[n].cpu().numpy(), g_na[n].cpu().numpy()
            assert np.all(np.isfinite(b)), n
            assert np.array_equal(a, b), (n, float(np.max(np.abs(a - b))))


def test_noalloc_rejects_a_tape_larger_than_the_array():
    from pyshield_amd import _lib
    f, cot, params = _setup(2, 2, 93)
    # ntimes sub-steps of 77 x 192 doubles per wavefront: past 128 MiB the
    # tape of one wavefront no longer fits
    _lib.check(_lib.lib().shield_mp_update_timestep(ctypes.byref(params), 150.0 * 1200), "timestep")
    assert params.ntimes == 1200
    with pytest.raises(_lib.ShieldError, match="workspace"):
        _noalloc(f, cot, _grads(f), params)
