"""Branch signatures of the oracle, to pick points away from every threshold.

``RecordingOps`` is the oracle's NumPy namespace with every branch recorded:
the condition of each ``where`` and the operand each ``minimum``,
``maximum`` and ``abs`` selects (the reference's ``if`` blocks and its
``min``/``max``/``dim`` kinks, microphysics_funcs.py and microphysics.py).
The oracle evaluates every statement over all columns at once, so the n-th
record of two runs on batches of the same width refers to the same statement
and the same columns.  A point x is at least a relative distance d from
every threshold along a direction e when the records at x, x (1 + d e) and
x (1 - d e) agree in its column: finite differences with a smaller step then
stay on one branch.

Not collected by pytest (no test_ prefix); only tests/ import it.
"""

import numpy as np

from oracle import mp_numpy as M


class RecordingOps(M.Ops):
    def __init__(self):
        self.records = []

    def _rec(self, mask, ncol):
        m = np.asarray(mask, dtype=bool)
        if m.ndim == 0:
            return
        # one row per column (the oracle's arrays are [ncol] or [ncol, nz])
        if m.shape[0] == ncol:
            self.records.append(m.reshape(ncol, -1))

    def bind(self, ncol):
        self.ncol = ncol
        return self

    def where(self, c, a, b):
        self._rec(c, self.ncol)
        return np.where(c, a, b)

    def minimum(self, a, b):
        self._rec(np.less(a, b), self.ncol)
        return np.minimum(a, b)

    def maximum(self, a, b):
        self._rec(np.greater(a, b), self.ncol)
        return np.maximum(a, b)

    def abs(self, x):
        self._rec(np.less(x, 0.0), self.ncol)
        return np.abs(x)


def signatures(nml, dt, state):
    """The branch records of one oracle call on `state` ([ncol, nz] fields and
    [ncol] land, area), as one boolean row per column."""
    ncol = state["pt"].shape[0]
    ops = RecordingOps().bind(ncol)
    M.run(nml, dt, state, xp=ops)
    return np.concatenate([r for r in ops.records], axis=1)


def safe_points(nml, dt, cols, candidates, margin):
    """For each candidate (column index, field, level) report whether the
    branch records of its column are unchanged when the field at that level
    is scaled by 1 +- margin.  `cols` holds [ncol, nz] fields and [ncol] land,
    area of the sampled columns."""
    batch = {k: [] for k in cols}
    for c, field, k in candidates:
        for s in (0.0, margin, -margin):
            for name, arr in cols.items():
                row = arr[c].copy()
                if name == field:
                    row[k] *= 1.0 + s
                batch[name].append(row)
    batch = {k: np.stack(v) for k, v in batch.items()}
    sig = signatures(nml, dt, batch)
    sig = sig.reshape(len(candidates), 3, -1)
    return [bool(np.array_equal(s[0], s[1]) and np.array_equal(s[0], s[2])) for s in sig]
