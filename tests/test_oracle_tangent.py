"""The oracle's FP64 torch transcription differentiated both ways: forward
mode (the tangent-linear the GPU adjoint's dot-product test uses) against
reverse mode, <J v, w> = <v, J^T w> over the twelve input fields, qaz and the
six process parameters.  Both come from the same code path, so they agree to
round-off; a forward-mode gap (an op without a tangent rule, a parameter
dropped from the dual inputs) breaks the identity."""

import numpy as np

from pyshield_amd._config import PhysicsConfig

from helpers import small_state
from test_gpu_adjoint import STATE, _cotangents, _directions, _oracle_vjp, oracle_jvp_dot


def test_oracle_tangent_linear_matches_reverse_mode():
    dt = 150.0
    st = small_state(2, 2, 79, seed=91)
    qaz = np.random.default_rng(91).uniform(0.0, 1.0, st["pt"].shape)
    cfg = PhysicsConfig(dt_atmos=int(dt), npz=79)
    cot = _cotangents(st, 92)
    v, vq, vp = _directions(st, 93)
    jv_w = oracle_jvp_dot(st, qaz, cfg, dt, v, vq, vp, cot)
    g = _oracle_vjp(st, qaz, cfg, dt, cot)
    v_jtw = (sum(float((g[n] * v[n]).sum()) for n in STATE) + float((g["qaz"] * vq).sum())
             + float((g["params"] * vp).sum()))
    assert abs(jv_w - v_jtw) <= 1e-11 * abs(jv_w), (jv_w, v_jtw)
    # every parameter direction reaches the outputs
    assert np.all(g["params"] != 0.0)
