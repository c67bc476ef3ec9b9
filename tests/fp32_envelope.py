"""How far FP32-accurate rate laws can move the oracle's answer, per cell.

The mixed-precision kernel evaluates the process rate laws in FP32 (about
one FP32 ulp, 1.2e-7 relative, per law) and keeps the saturation mixing
ratios, temperature differences, latent heats and every transfer in FP64.
The reference's column physics has cells whose answer that cannot fix to
1e-5: a sink that removes nearly all of a species amplifies any relative
error of what flows into the cell by (input / residue), and a branch on a
threshold (a fall speed switching to VR_MIN below qr = 1e-8, the
evaporation of all clouds below rh_adj) jumps when its argument moves by one
ulp.  ``envelope`` measures this on the FP64 oracle itself: it reruns the
oracle with every exp, log, sqrt and power outside the saturation functions
multiplied by (1 + sigma N(0, 1)), and returns, per cell, the largest change
of the updated state over the trials.  A cell of the mixed kernel that misses
a bound is accounted for when the oracle's own envelope there is of the same
size; a cell that misses it where the oracle is well conditioned is a defect.

Not collected by pytest (no test_ prefix); only tests/ import it.
"""

import sys

import numpy as np

from oracle import mp_numpy as M

SATURATION = ("wqs1", "wqs2", "iqs1", "iqs2")
# The kernel's FP32 rate laws chain a logarithm, an exp2 and a reciprocal
# (about 1 ulp each) with three to six products (0.5 ulp each): up to about
# 6 ulp, 2-3 ulp typically.  The envelope the cell checks use puts 4 ulp on
# each transcendental of the oracle's rate laws.
KERNEL_SIGMA = 4 * 1.19e-7
UPDATED = ("pt", "qv", "ql", "qr", "qi", "qs", "qg")


class NoisyOps(M.Ops):
    """The oracle's NumPy namespace with a relative error of `sigma` on each
    transcendental result outside the saturation functions."""

    def __init__(self, sigma, rng):
        self.sigma, self.rng = sigma, rng

    def _noise(self, r):
        caller = sys._getframe(2).f_code.co_name
        if caller in SATURATION:
            return r
        return r * (1.0 + self.sigma * self.rng.standard_normal(np.shape(r)))

    def exp(self, x):
        return self._noise(np.exp(x))

    def log(self, x):
        return self._noise(np.log(x))

    def sqrt(self, x):
        return self._noise(np.sqrt(x))

    def power(self, x, y):
        return self._noise(np.power(x, y))


def updated(state, out, dt):
    """The state after the call's forward-Euler update (physics.py:146-194)."""
    return {n: state[n] + out[n + "_dt"] * dt for n in UPDATED}


def envelope(nml, dt, cols, trials=16, sigma=KERNEL_SIGMA, seed=0, qaz=None):
    """Per cell and field of UPDATED, the largest |updated_noisy - updated|
    over `trials` noisy oracle runs of the columns `cols` ([n, nz] fields,
    [n] land and area).  Returns (clean updated state, envelope)."""
    n = cols["pt"].shape[0]
    clean = updated(cols, M.run(nml, dt, cols, qaz=qaz), dt)
    rep = {k: np.concatenate([v] * trials, axis=0) for k, v in cols.items()}
    qrep = None if qaz is None else np.concatenate([qaz] * trials, axis=0)
    out = M.run(nml, dt, rep, qaz=qrep, xp=NoisyOps(sigma, np.random.default_rng(seed)))
    noisy = updated(rep, out, dt)
    env = {}
    for k in UPDATED:
        dev = np.abs(noisy[k] - np.concatenate([clean[k]] * trials, axis=0))
        env[k] = dev.reshape((trials, n) + dev.shape[1:]).max(axis=0)
    return clean, env


T_TOL = 1e-3          # K
Q_RTOL = 1e-5         # relative
Q_FLOOR = 1e-12       # kg/kg
ENV_FACTOR = 2.0


def bound(field, ref):
    """The stated per-cell bound of the mixed-precision call against the FP64
    oracle: 1e-3 K on the temperature, 1e-5 relative with a 1e-12 kg/kg floor
    on the mixing ratios."""
    return np.full_like(ref, T_TOL) if field == "pt" else Q_RTOL * np.abs(ref) + Q_FLOOR


def check_cells(nml, dt, state, tendencies, trials=16, sigma=KERNEL_SIGMA):
    """The mixed call's updated state against the oracle, cell by cell.

    `state` holds the call's inputs ([..., nz] fields, [...] land and area),
    `tendencies` the mixed call's tendencies.  Every cell must be within
    `bound`, or within ENV_FACTOR x the oracle's own envelope under FP32
    rate laws (computed for the columns that hold a cell beyond the bound).
    Returns a report: per field the cells beyond the bound, those the
    envelope accounts for, the largest error / bound and error / envelope,
    and the cells left unexplained."""
    nz = state["pt"].shape[-1]
    cols = {k: (np.asarray(v).reshape(-1, nz) if np.ndim(v) == np.ndim(state["pt"]) else np.asarray(v).reshape(-1))
            for k, v in state.items() if k in M_STATE}
    tend = {k: np.asarray(v).reshape(-1, nz) for k, v in tendencies.items() if k.endswith("_dt")}
    clean = updated(cols, M.run(nml, dt, cols), dt)
    mixed = {n: cols[n] + tend[n + "_dt"] * dt for n in UPDATED}
    err = {n: np.abs(mixed[n] - clean[n]) for n in UPDATED}
    over = {n: err[n] > bound(n, clean[n]) for n in UPDATED}
    bad_cols = np.flatnonzero(np.any([over[n].any(axis=1) for n in UPDATED], axis=0))
    report = {"cells": int(cols["pt"].size), "columns_checked_by_envelope": int(bad_cols.size), "fields": {},
              "unexplained": []}
    env = None
    if bad_cols.size:
        sub = {k: v[bad_cols] for k, v in cols.items()}
        _, env = envelope(nml, dt, sub, trials=trials, sigma=sigma)
    for n in UPDATED:
        r = {"beyond_bound": int(over[n].sum()),
             "worst_over_bound": float(np.max(err[n] / bound(n, clean[n])))}
        if bad_cols.size:
            e, o = err[n][bad_cols], over[n][bad_cols]
            ok = e <= ENV_FACTOR * env[n]
            r["explained"] = int((o & ok).sum())
            r["worst_over_envelope"] = float(np.max(np.where(o, e / np.maximum(env[n], 1e-300), 0.0)))
            for c, k in np.argwhere(o & ~ok):
                report["unexplained"].append((n, int(bad_cols[c]), int(k), float(e[c, k]),
                                              float(bound(n, clean[n])[bad_cols[c], k]), float(env[n][c, k])))
        report["fields"][n] = r
    return report


M_STATE = ("pt", "qv", "ql", "qr", "qi", "qs", "qg", "qa", "ua", "va", "delp", "delz", "w", "land", "area")
