"""The C ABI library loads, exports every entry point include/*.h declares,
and rejects bad arguments before touching the GPU.

The argument checks pass made-up device addresses, so they live in
tests/abi_fake_pointers.py and run in a child process started with every HIP
device hidden; the child refuses to run them if it still sees a device, and
the tests here are then skipped.  No call in this file can reach a kernel."""

import ctypes
import glob
import json
import os
import re
import subprocess
import sys

import pytest

from pyshield_amd import _lib
from pyshield_amd._config import PhysicsConfig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))

sys.path.insert(0, HERE)
import abi_fake_pointers  # noqa: E402


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(shield_\w+)\s*\(", text, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert {"shield_mp_setup", "shield_mp_forward", "shield_fill_gfs_delp"} <= names
    for n in sorted(names):
        assert hasattr(L, n), n
    assert set(_lib.EXPORTS) == names


def test_struct_sizes_match_layout():
    # the ctypes mirrors must have the C sizes: 37 doubles + 20 int32 in the namelist
    assert ctypes.sizeof(_lib.ShieldMpNamelist) == 37 * 8 + 20 * 4
    assert ctypes.sizeof(_lib.ShieldMpShape) == 7 * 8
    assert ctypes.sizeof(_lib.ShieldMpFields) == len(_lib.FIELD_NAMES) * 8
    # 14 output cotangents, 13 input cotangents, the parameter gradient, w's two
    assert ctypes.sizeof(_lib.ShieldMpAdjointFields) == (14 + 13 + 1 + 2) * 8
    # 21 pointers, two int64 strides, three doubles
    assert ctypes.sizeof(_lib.ShieldUpdateFields) == 21 * 8 + 2 * 8 + 3 * 8
    # 31 pointers, two int64 strides, one double
    assert ctypes.sizeof(_lib.ShieldPrologueFields) == 31 * 8 + 2 * 8 + 8
    # 37 pointers, three doubles, four int32
    assert ctypes.sizeof(_lib.ShieldDycoreFields) == 37 * 8 + 3 * 8 + 4 * 4


@pytest.fixture(scope="module")
def fake_pointer_results():
    """Run every check of abi_fake_pointers in one child process that sees no
    HIP device (ROCR/HIP_VISIBLE_DEVICES set to an index that does not exist
    before the child's first HIP call)."""
    env = dict(os.environ, ROCR_VISIBLE_DEVICES="-1", HIP_VISIBLE_DEVICES="-1", CUDA_VISIBLE_DEVICES="-1")
    r = subprocess.run([sys.executable, os.path.join(HERE, "abi_fake_pointers.py")], env=env,
                       capture_output=True, text=True, timeout=300)
    if r.returncode == abi_fake_pointers.NO_DEVICE_EXIT:
        pytest.skip("the child process still saw a HIP device; fake-pointer checks not run")
    assert r.returncode == 0, r.stdout + r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("check", abi_fake_pointers.CHECKS)
def test_rejects_bad_arguments_before_launch(fake_pointer_results, check):
    assert fake_pointer_results[check] == "", fake_pointer_results[check]


def test_update_atmosphere_state_scope():
    from types import SimpleNamespace
    from pyshield_amd.update import UpdateAtmosphereState
    nml = SimpleNamespace(dt_atmos=225, npx=5, npy=5)
    with pytest.raises(ValueError):
        UpdateAtmosphereState(namelist=nml, apply_tendencies=True)   # needs the grid data
    assert UpdateAtmosphereState(namelist=nml)._rdt == pytest.approx(1.0 / 225.0)
    # the doubly periodic form needs no grid fields
    op = UpdateAtmosphereState(namelist=nml, apply_tendencies=True, grid_info=SimpleNamespace(grid_type=4))
    assert op._apply_physics_to_dycore.grid == {}


def test_setup_errors():
    nml = PhysicsConfig().mp_namelist()
    with pytest.raises(_lib.ShieldError, match="timestep"):
        _lib.setup(nml, 0.0)
    with pytest.raises(_lib.ShieldError, match="use_ppm"):
        _lib.setup(dict(nml, use_ppm=1), 150.0)
    with pytest.raises(_lib.ShieldError, match="use_ppm and prog_ccn"):
        _lib.setup(dict(nml, prog_ccn=1), 150.0)


def test_strerror_and_layout():
    L = _lib.lib()
    for code in range(-10, 1):
        assert L.shield_strerror(code)
    if _lib.device_count() == 0:
        # without a device the layout with the most busy lanes wins
        assert _lib.layout(79) == (4, 5)   # 4 columns in 5 slots: 316 of 320 lanes busy
        assert _lib.layout(127) == (1, 2)
        assert _lib.layout(8) == (8, 1)
    with pytest.raises(_lib.ShieldError):
        _lib.layout(1000)
    with pytest.raises(_lib.ShieldError):
        _lib.layout(79, 7)                 # unknown precision mode


