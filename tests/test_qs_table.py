"""The mixed kernel's FP64-built saturation table (pyshield_amd/csrc/qs_table.h,
written by gen_qs_table.py) against the reference's closed forms wqs2 / iqs2
(microphysics_funcs.py:53-145) evaluated in FP64, on the host build of the
kernel's saturation functions (tests/host/qs_table_check.hip).  The closed
forms themselves are pinned by textbook saturation pressures in
tests/test_oracle.py."""

import ctypes
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TICE, T_SAT_MIN = 273.16, 273.16 - 160.0


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    out = str(tmp_path_factory.mktemp("qs") / "libqs_table_check.so")
    cmd = [hipcc, "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-x", "hip", "-fPIC", "-shared", os.path.join(HERE, "host", "qs_table_check.hip"),
           "-I", os.path.join(ROOT, "include"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    L = ctypes.CDLL(out)
    L.qs_check.restype = None
    return L


def run(L, t, den):
    t = np.ascontiguousarray(t, dtype=np.float64)
    den = np.ascontiguousarray(den, dtype=np.float64)
    out = np.zeros((t.size, 5))
    dq = np.zeros((t.size, 4))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    L.qs_check(ctypes.c_int(t.size), p(t), p(den), p(out), p(dq))
    return out, dq


def test_header_is_the_generators_output():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "pyshield_amd", "csrc", "gen_qs_table.py"), "--check"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_table_matches_the_closed_forms(lib):
    rng = np.random.default_rng(7)
    t = np.concatenate([np.linspace(100.0, 1000.0, 30001), rng.uniform(100.0, 500.0, 20000),
                        [TICE, T_SAT_MIN, TICE + 102.0, 100.0, 1000.0]])
    den = rng.uniform(0.01, 1.4, t.size)
    out, dq = run(lib, t, den)
    rel = lambda a, b: np.abs(a / b - 1.0)  # noqa: E731
    assert rel(out[:, 0], out[:, 1]).max() < 2e-12  # over water, 100-1000 K
    assert rel(dq[:, 0], dq[:, 1]).max() < 2e-12
    assert rel(out[:, 2], out[:, 3]).max() < 5e-13  # iqs: ice below TICE, water above, both clamped
    assert rel(dq[:, 2], dq[:, 3]).max() < 5e-13
    assert np.array_equal(out[:, 4], out[:, 2])     # qs_pick(water=False) is iqs_fast
    ice = (t >= T_SAT_MIN) & (t < TICE)
    assert rel(out[ice, 2], out[ice, 3]).max() < 1e-13


def test_clamps(lib):
    # the reference's iqs clamps below T_SAT_MIN and above TICE + 102: the table
    # reproduces the clamped values; over water the table clamps at 100 K and
    # 1000 K, where below 100 K the ratio is under 1e-15 kg/kg
    t = np.array([50.0, 80.0, 99.0, 110.0, 380.0, 462.0, 999.0, 1200.0])
    den = np.full(t.size, 0.5)
    out, _ = run(lib, t, den)
    assert np.all(np.abs(out[:, 2] / out[:, 3] - 1.0) < 5e-13)
    assert np.all(out[t < 100.0, 0] < 1e-15) and np.all(out[t < 100.0, 1] < 1e-15)
    inside = (t >= 100.0) & (t <= 1000.0)
    assert np.all(np.abs(out[inside, 0] / out[inside, 1] - 1.0) < 2e-12)
