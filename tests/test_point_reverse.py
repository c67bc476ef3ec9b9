"""The hand-written reverses of the point-wise processes
(pyshield_amd/csrc/mp_point_adj.h) on the host, against the dual-number
vector-Jacobian products of the same process code (mp_process.h, one dual pass
per input), on random cells that reach every branch:

* icloud (icloud_melt_freeze + icloud_main): warm and cold, below T_MIN, the
  instant-evaporation collapse, p1 below P_MIN, zero and tiny species.  The
  block-wise forward inside the reverse must also reproduce the process code
  bit for bit (every branch of the reverse is the forward's);
* revap_racc (evaporation and accretion of rain);
* the autoconversion, with and without the subgrid width;
* the solid and rain fall speeds (including their clamps) and the
  sedimentation ice melt.

The products must agree to round-off, normalised by each input's largest
cotangent.  The dual-number path itself is checked against torch autograd of
the oracle on the GPU (test_gpu_adjoint.py), and the whole adjoint against it
through these reverses.
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
    out = str(tmp_path_factory.mktemp("pt") / "libpoint_vjp_check.so")
    cmd = [hipcc, "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-x", "hip", "-fPIC", "-shared", os.path.join(HERE, "host", "point_vjp_check.hip"),
           "-I", os.path.join(ROOT, "include"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lib = ctypes.CDLL(out)
    for fn in (lib.icl_check, lib.revap_check, lib.autoconv_check, lib.speed_check, lib.melt_check):
        fn.restype = None
    return lib


def _p(a):
    return np.ascontiguousarray(a).ctypes.data_as(ctypes.c_void_p)


def _norm_errs(oh, od, names):
    return {nm: float(np.max(np.abs(oh[:, i] - od[:, i])) / max(np.max(np.abs(od[:, i])), 1e-300))
            for i, nm in enumerate(names)}


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

    harness.icl_check(ctypes.byref(params), ctypes.c_int(n), _p(cells), _p(ins), _p(hvr), _p(yb), _p(od), _p(oh),
                      _p(fg), _p(fb))
    assert np.array_equal(fg, fb), f"block forward differs in {int(np.sum(np.any(fg != fb, 1)))} cells"
    errs = _norm_errs(oh, od, INPUTS)
    print(json.dumps({k: f"{v:.1e}" for k, v in errs.items()}))
    assert np.isfinite(oh).all()
    assert all(v <= 1e-12 for v in errs.values()), errs


@pytest.mark.parametrize("dt", [150.0, 225.0])
def test_revap_reverse_matches_dual_numbers(harness, dt):
    params = _lib.setup(PhysicsConfig(dt_atmos=int(dt), npz=79).mp_namelist(), dt)
    n = 8000
    cells, ins, yb = _cells(n, 3 + int(dt))
    # warm cells with rain, and a near-saturated share so both evaporation
    # forms (above and inside the subgrid spread) and the accretion occur
    rng = np.random.default_rng(int(dt))
    cells[:, 6] = rng.uniform(230, 305, n)
    cells[:, 2] = 10 ** rng.uniform(-8.5, -2.5, n)
    od, oh = np.zeros((n, 10)), np.zeros((n, 10))
    harness.revap_check(ctypes.byref(params), ctypes.c_int(n), _p(cells), _p(ins), ctypes.c_double(0.15), _p(yb),
                        _p(od), _p(oh))
    errs = _norm_errs(oh, od, INPUTS[:8] + ["den", "denfac"])
    print(json.dumps({k: f"{v:.1e}" for k, v in errs.items()}))
    assert np.isfinite(oh).all()
    assert all(v <= 1e-12 for v in errs.values()), errs


@pytest.mark.parametrize("subgrid,use_ccn", [(0, True), (1, True), (1, False)])
def test_autoconversion_reverse_matches_dual_numbers(harness, subgrid, use_ccn):
    params = _lib.setup(PhysicsConfig(dt_atmos=150, npz=79).mp_namelist(), 150.0)
    # use_ccn = 0 (qc0 divided by the density) is the prognostic-ccn form,
    # which setup rejects; the process code supports it
    params.use_ccn = int(use_ccn)
    n = 8000
    rng = np.random.default_rng(40 + subgrid)
    v = np.stack([rng.uniform(240, 300, n), 10 ** rng.uniform(-7, -2.3, n), 10 ** rng.uniform(-9, -3, n),
                  rng.uniform(0.1, 1.3, n), rng.uniform(5e7, 3e8, n), rng.uniform(1e-3, 1e-1, n),
                  10 ** rng.uniform(-8, -3, n)], 1)
    # half the cells just above the threshold qc, where the sink is dq itself
    qc = params.fac_rc * v[:, 4] / (1.0 if use_ccn else v[:, 3])
    near = rng.random(n) < 0.5
    v[near, 1] = (qc * (1.0 + 10 ** rng.uniform(-7, -1, n)))[near]
    yb = rng.normal(size=(n, 2))
    od, oh = np.zeros((n, 7)), np.zeros((n, 7))
    harness.autoconv_check(ctypes.byref(params), ctypes.c_int(subgrid), ctypes.c_int(n), _p(v), _p(yb), _p(od),
                           _p(oh))
    errs = _norm_errs(oh, od, ["ql", "qr", "den", "ccn", "c_praut", "dl", "fac_rc"])
    print(json.dumps({k: f"{v:.1e}" for k, v in errs.items()}))
    assert np.isfinite(oh).all()
    assert all(v <= 1e-12 for v in errs.values()), errs
    # the threshold's cotangents are exercised
    assert np.abs(od[:, 3]).max() > 0.0 and np.abs(od[:, 6]).max() > 0.0


@pytest.mark.parametrize("overrides", [{}, {"const_vi": True, "const_vs": True, "const_vg": True, "const_vr": True}])
def test_fall_speed_reverses_match_dual_numbers(harness, overrides):
    nml = PhysicsConfig(dt_atmos=150, npz=79).mp_namelist()
    nml.update(overrides)
    params = _lib.setup(nml, 150.0)
    n = 8000
    rng = np.random.default_rng(61)
    # species across the thresholds and up to the speed caps
    v = np.stack([10 ** rng.uniform(-9.5, -1.5, n), 10 ** rng.uniform(-9.5, -2, n), 10 ** rng.uniform(-9.5, -1.5, n),
                  rng.uniform(-90, 30, n), rng.uniform(0.02, 1.3, n), 10 ** rng.uniform(-9.5, -1.5, n)], 1)
    yb = rng.normal(size=(n, 4))
    od, oh = np.zeros((n, 11)), np.zeros((n, 11))
    harness.speed_check(ctypes.byref(params), ctypes.c_int(n), _p(v), _p(yb), _p(od), _p(oh))
    errs = _norm_errs(oh, od, ["qg", "qi", "qs", "tc", "den", "vs_fac", "vg_fac", "vi_fac", "qr", "den_r",
                               "vr_fac"])
    print(json.dumps({k: f"{v:.1e}" for k, v in errs.items()}))
    assert np.isfinite(oh).all()
    assert all(v <= 1e-12 for v in errs.values()), errs


def test_sedimentation_ice_melt_reverse_matches_dual_numbers(harness):
    params = _lib.setup(PhysicsConfig(dt_atmos=150, npz=79).mp_namelist(), 150.0)
    n = 8000
    cells, _, yb = _cells(n, 71)
    cells[:, 6] = np.random.default_rng(72).uniform(265, 290, n)
    od, oh = np.zeros((n, 7)), np.zeros((n, 7))
    harness.melt_check(ctypes.byref(params), ctypes.c_int(n), _p(cells), _p(yb), _p(od), _p(oh))
    errs = _norm_errs(oh, od, INPUTS[:7])
    print(json.dumps({k: f"{v:.1e}" for k, v in errs.items()}))
    assert all(v <= 1e-12 for v in errs.values()), errs
