"""The two independent restatements of the reference -- the C oracle
(sequential per column, branches) and the stage-faithful array version
(oracle/mp_numpy.py: vectorised PARALLEL blocks, explicit k loops, masks) --
agree to round-off on every namelist branch the tests use."""

import numpy as np
import pytest

from oracle import mp_numpy
from pyshield_amd._config import PhysicsConfig

from helpers import (PRECIP, TEND, cold_top_state, low_pressure_top_state, norm_err, overdraw_state, run_oracle,
                     small_state)
from test_oracle import NAMELISTS


def _array_run(st, cfg, dt):
    flat = {k: v.reshape(-1, v.shape[-1]) if v.ndim == 3 else v.reshape(-1) for k, v in st.items()}
    return mp_numpy.run(cfg.mp_namelist(), dt, flat)


def _compare(a, b):
    errs = {}
    for n in TEND + ("qaz",):
        errs[n] = norm_err(b[n].reshape(a[n].shape), a[n], floor=1e-30)
    for n in PRECIP:
        # precipitation below 1e-6 mm/day is flux-sum round-off
        errs[n] = norm_err(b[n].reshape(a[n].shape), a[n], floor=1e-6)
    return errs


@pytest.mark.parametrize("overrides", NAMELISTS)
@pytest.mark.parametrize("dt", [150.0, 225.0])
def test_restatements_agree(overrides, dt):
    st = small_state(6, 5, 79, seed=31)
    cfg = PhysicsConfig(dt_atmos=int(dt), **overrides)
    errs = _compare(run_oracle(st, cfg, dt), _array_run(st, cfg, dt))
    for n in TEND + ("qaz",):
        assert errs[n] < 1e-11, (n, errs[n])
    for n in PRECIP:
        assert errs[n] < 1e-7, (n, errs[n])


@pytest.mark.parametrize("make", [cold_top_state, lambda: low_pressure_top_state(True),
                                  lambda: low_pressure_top_state(False), overdraw_state])
def test_restatements_agree_on_special_columns(make):
    st = make()
    cfg = PhysicsConfig()
    errs = _compare(run_oracle(st, cfg, 150.0), _array_run(st, cfg, 150.0))
    assert max(errs[n] for n in TEND + ("qaz",)) < 1e-11, errs
    assert max(errs[n] for n in PRECIP) < 1e-7, errs


def test_restatements_agree_deep_grid():
    from pyshield_amd.synthetic import make_state
    st = make_state(3, 3, 127, 5000, ptop=100.0, qg_median=5e-4)
    cfg = PhysicsConfig(dt_atmos=90)
    errs = _compare(run_oracle(st, cfg, 90.0), _array_run(st, cfg, 90.0))
    assert max(errs[n] for n in TEND) < 1e-11, errs
