"""The hand-written reverse of the icloud step (pyshield_amd/csrc/mp_icloud_adj.h)
on the host, against the dual-number vector-Jacobian product of the same
process code (mp_process.h: icloud_melt_freeze + icloud_main, one dual pass per
input), on random cells that reach every branch: warm and cold, below T_MIN,
the instant-evaporation collapse, p1 below P_MIN, zero and tiny species.

The block-wise forward inside the reverse must reproduce the process code bit
for bit (every branch of the reverse is the forward's); the products must agree
to round-off, normalised by each input's largest cotangent.  The dual-number
path itself is checked against torch autograd of the oracle on the GPU
(test_gpu_adjoint.py).
"""

import ctypes
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from pyshield_amd import _lib
from pyshield_amd._config import PhysicsConfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INPUTS = ["qv", "ql", "qr", "qi", "qs", "qg", "tz", "qa", "den", "denfac", "vtg", "vtr", "vts", "p1", "qi_gen"]


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    out = str(tmp_path_factory.mktemp("icl") / "libicl_check.so")
    cmd = [hipcc, "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-x", "hip", "-fPIC", "-shared", os.path.join(HERE, "host", "icloud_vjp_check.hip"),
           "-I", os.path.join(ROOT, "include"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lib = ctypes.CDLL(out)
    lib.icl_check.restype = None
    return lib


def _cells(n, seed):
    rng = np.random.default_rng(seed)

    def lq(lo, hi, pz=0.15):
        v = 10 ** rng.uniform(lo, hi, n)
        v[rng.random(n) < pz] = 0.0
        return v

    cells = np.stack([10 ** rng.uniform(-6, -1.7, n), lq(-9, -2.5), lq(-9, -2.5), lq(-9, -3), lq(-9, -2.5),
                      lq(-9, -2.5), rng.uniform(170, 305, n), rng.uniform(0, 1, n)], 1)
    den = rng.uniform(0.05, 1.3, n)
    ins = np.stack([den, np.sqrt(1.2 / den), rng.uniform(0.5, 8, n), rng.uniform(0.5, 9, n), rng.uniform(0.3, 2, n),
                    np.where(rng.random(n) < 0.05, 50.0, rng.uniform(200, 1.0e5, n))], 1)
    return cells, ins, rng.normal(size=(n, 8))


@pytest.mark.parametrize("dt,overrides", [(150.0, {}), (225.0, {"do_qa": True, "fast_sat_adj": True}),
                                          (150.0, {"const_vi": True, "rad_snow": False, "rad_rain": False})])
def test_icloud_reverse_matches_dual_numbers(harness, dt, overrides):
    nml = PhysicsConfig(dt_atmos=int(dt), npz=79).mp_namelist()
    nml.update(overrides)
    params = _lib.setup(nml, dt)
    n = 6000
    cells, ins, yb = _cells(n, 17 + int(dt))
    hvr = np.array([0.2, 0.75, 0.45])
    od, oh = np.zeros((n, 15)), np.zeros((n, 15))
    fg, fb = np.zeros((n, 8)), np.zeros((n, 8))

    def p(a):
        return np.ascontiguousarray(a).ctypes.data_as(ctypes.c_void_p)

    harness.icl_check(ctypes.byref(params), ctypes.c_int(n), p(cells), p(ins), p(hvr), p(yb), p(od), p(oh), p(fg),
                      p(fb))
    assert np.array_equal(fg, fb), f"block forward differs in {int(np.sum(np.any(fg != fb, 1)))} cells"
    errs = {}
    for i, name in enumerate(INPUTS):
        errs[name] = float(np.max(np.abs(oh[:, i] - od[:, i])) / max(np.max(np.abs(od[:, i])), 1e-300))
    print(json.dumps({k: f"{v:.1e}" for k, v in errs.items()}))
    assert np.isfinite(oh).all()
    assert all(v <= 1e-12 for v in errs.values()), errs
