"""Savepoint harness for the ``Microph`` translate test.

The reference validates ``Microphysics`` against serialized Fortran data
through NDSL's translate-test framework:

* ``TranslateMicroph`` (tests/savepoint/translate/translate_microphysics.py)
  names the savepoint variables (``mph_qv1`` .. ``mph_land`` in,
  ``mph_pt_dt`` .. ``mph_vdt`` out), builds a fresh ``Microphysics`` from the
  test's namelist and grid, calls it once with ``timestep = dt_atmos`` and
  returns the ten tendencies;
* ``transform_microphysics_serialized_data`` (translate_physics.py:76-92)
  reshapes the collapsed ``[i*j, 1, k]`` input layout (C order over i, j; k = 0
  the model top, as the dynamical core) to ``[i, j, k]``, and a ``[i*j]`` field
  to ``[i, j]``;
* ``slice_output`` (translate_physics.py:130-181) with TranslateMicroph's final
  ``[:, np.newaxis, ::-1]`` writes the outputs back as ``[i*j, 1, k]`` with
  k = 0 the top;
* the comparison is NDSL's: relative error ``2 |a - b| / (|a| + |b|)`` (0
  where both vanish) below ``max_error``, with values whose reference and
  computed magnitudes are both below a per-variable ``ignore_near_zero_errors``
  threshold accepted (tests/savepoint/translate/overrides/*.yaml hold the
  thresholds per backend).

Here the same test runs on ``pyshield_amd.Microphysics`` (the HIP kernels, FP64
by default).  ``load_savepoint`` reads the savepoint containers: NDSL's
netCDF-4 files (HDF5, read by ``pyshield_amd.hdf5``), classic (version 1/2)
netCDF files and ``.npz`` archives.
"""

import os

import numpy as np

# serialname -> MicrophysicsState field (translate_microphysics.py:18-32)
MICROPH_IN = {
    "mph_qv1": "qvapor",
    "mph_ql1": "qliquid",
    "mph_qr1": "qrain",
    "mph_qi1": "qice",
    "mph_qs1": "qsnow",
    "mph_qg1": "qgraupel",
    "mph_qa1": "qcld",
    "mph_uin": "ua",
    "mph_vin": "va",
    "mph_delp": "delprsi",
    "mph_w": "wmp",
    "mph_dz": "delz",
    "mph_pt": "pt",
    "mph_land": "land",
}
# tendency -> serialname (translate_microphysics.py:34-45)
MICROPH_OUT = {
    "pt_dt": "mph_pt_dt",
    "qv_dt": "mph_qv_dt",
    "ql_dt": "mph_ql_dt",
    "qr_dt": "mph_qr_dt",
    "qi_dt": "mph_qi_dt",
    "qs_dt": "mph_qs_dt",
    "qg_dt": "mph_qg_dt",
    "qa_dt": "mph_qa_dt",
    "udt": "mph_udt",
    "vdt": "mph_vdt",
}

# The reference's Microph thresholds for a GPU backend
# (tests/savepoint/translate/overrides/standard.yaml, gt:gpu entry).
MICROPH_MAX_ERROR = 2.2e-8
MICROPH_NEAR_ZERO = {"mph_ql_dt": 1e-8, "mph_qr_dt": 1e-9, "mph_qg_dt": 1e-18}


def _side(n):
    cn = int(round(np.sqrt(n)))
    if cn * cn != n:
        raise ValueError(f"{n} columns do not form a square tile")
    return cn


def microph_to_ijk(data):
    """[i*j, 1, k] -> [i, j, k] and [i*j] (or [i*j, 1]) -> [i, j], C order over
    (i, j), the k axis unchanged (translate_physics.py:76-92)."""
    a = np.asarray(data)
    if a.ndim == 3:
        cn = _side(a.shape[0])
        return np.reshape(a[:, 0, :], (cn, cn, a.shape[2]))
    if a.ndim in (1, 2) and a.size == a.shape[0]:
        cn = _side(a.shape[0])
        return np.reshape(a.reshape(-1), (cn, cn))
    if a.size == 1:
        return a.reshape(-1)[0]
    raise ValueError(f"unsupported microphysics savepoint shape {a.shape}")


def ijk_to_microph(data):
    """[i, j, k] -> [i*j, 1, k], the layout TranslateMicroph returns."""
    a = np.asarray(data)
    if a.ndim != 3:
        raise ValueError(f"expected [i, j, k], got {a.shape}")
    return np.reshape(a, (a.shape[0] * a.shape[1], a.shape[2]))[:, np.newaxis, :]


def relative_error(computed, reference):
    """NDSL's element-wise relative difference 2 |a - b| / (|a| + |b|), 0 where
    both are 0."""
    a = np.asarray(computed, dtype=np.float64)
    b = np.asarray(reference, dtype=np.float64)
    denom = np.abs(a) + np.abs(b)
    with np.errstate(divide="ignore", invalid="ignore"):
        err = np.where(denom == 0.0, 0.0, 2.0 * np.abs(a - b) / denom)
    return err


def compare(computed, reference, max_error=MICROPH_MAX_ERROR, near_zero=None):
    """Per output variable: (passed, worst relative error, failing count).
    Values below the variable's near-zero threshold in both arrays pass."""
    near_zero = MICROPH_NEAR_ZERO if near_zero is None else near_zero
    report = {}
    for name, ref in reference.items():
        if name not in computed:
            raise KeyError(f"{name} was not computed")
        a, b = np.asarray(computed[name]), np.asarray(ref)
        if a.shape != b.shape:
            raise ValueError(f"{name}: computed {a.shape}, reference {b.shape}")
        err = relative_error(a, b)
        ok = (err < max_error) | (np.isnan(a) & np.isnan(b))
        thr = near_zero.get(name)
        if thr is not None:
            ok |= (np.abs(a) < thr) & (np.abs(b) < thr)
        report[name] = (bool(ok.all()), float(np.nanmax(np.where(ok, 0.0, err))) if err.size else 0.0,
                        int((~ok).sum()))
    return report


_SAVEPOINT_DIMS = (("savepoint", "n_savepoints"), ("rank", "n_ranks"))
# the NAME netCDF-4 gives a dimension that has no coordinate variable
_NC_DIM_ONLY = "This is a netCDF dimension but not a netCDF variable."


def _select(a, dims, savepoint, rank):
    if dims and len(dims) >= 2 and dims[0] in _SAVEPOINT_DIMS[0] and dims[1] in _SAVEPOINT_DIMS[1]:
        return a[savepoint, rank]
    return a


def load_savepoint(path, savepoint=0, rank=0):
    """Variables of one savepoint as numpy arrays.

    ``.npz``: the arrays as stored.  netCDF, netCDF-4 (HDF5) or classic
    format: each variable, with NDSL's leading (savepoint, rank) dimensions
    indexed by ``savepoint`` and ``rank`` when present (a variable whose first
    two dimensions are named ``savepoint`` and ``rank``, or ``n_savepoints``
    / ``n_ranks``); netCDF-4's dimension-only datasets are not variables."""
    ext = os.path.splitext(path)[1].lower()
    if ext == ".npz":
        with np.load(path, allow_pickle=False) as z:
            return {k: np.array(z[k]) for k in z.files}
    from pyshield_amd import hdf5
    if ext in (".nc", ".nc4", ".h5", ".hdf5") and hdf5.is_hdf5(path):
        out = {}
        for name, var in hdf5.read(path).items():
            label = var.attrs.get("NAME")
            if isinstance(label, (bytes, np.bytes_)):
                label = label.decode("utf-8", "replace")
            if var.attrs.get("CLASS") in ("DIMENSION_SCALE", b"DIMENSION_SCALE") and isinstance(label, str) \
                    and label.startswith(_NC_DIM_ONLY):
                continue
            out[name] = np.array(_select(var.data, var.dims, savepoint, rank))
        return out
    if ext in (".nc", ".nc3", ".cdf"):
        from scipy.io import netcdf_file
        out = {}
        with netcdf_file(path, "r", mmap=False) as f:
            for name, var in f.variables.items():
                out[name] = np.array(_select(np.array(var[:]), tuple(var.dimensions), savepoint, rank))
        return out
    raise ValueError(f"unsupported savepoint container {path!r} (netCDF-4/HDF5, classic netCDF or .npz)")


class TranslateMicroph:
    """The reference's TranslateMicroph on the HIP kernels.

    ``namelist``: a ``PhysicsConfig`` (its ``dt_atmos`` is the timestep);
    ``area``: the tile's cell areas ``[i, j]`` from the grid savepoint (the
    reference reads them from the grid data, not from the Microph savepoint);
    ``precision``: "fp64" (the reference's arithmetic) or "mixed"."""

    in_vars = MICROPH_IN
    out_vars = MICROPH_OUT

    def __init__(self, namelist, area, precision="fp64", device="cuda:0"):
        self.namelist = namelist
        self.area = np.asarray(area, dtype=np.float64)
        self.precision = precision
        self.device = device

    def compute(self, inputs):
        """Serialized inputs (``mph_*`` [i*j, 1, k], ``mph_land`` [i*j]) ->
        serialized tendencies ``mph_*_dt`` [i*j, 1, k]."""
        import torch

        from .microphysics import GridData, Microphysics, MicrophysicsState, QuantityFactory, StencilFactory

        missing = [n for n in self.in_vars if n not in inputs]
        if missing:
            raise KeyError(f"savepoint lacks {missing}")
        fields = {self.in_vars[n]: microph_to_ijk(inputs[n]) for n in self.in_vars}
        ni, nj, nz = fields["pt"].shape
        if self.area.shape != (ni, nj):
            raise ValueError(f"area {self.area.shape} does not match the tile ({ni}, {nj})")
        dev = torch.device(self.device)

        def t(x):
            return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device=dev).clone()

        zeros = torch.zeros((ni, nj, nz), dtype=torch.float64, device=dev)
        state = MicrophysicsState(
            pt=t(fields["pt"]), qvapor=t(fields["qvapor"]), qliquid=t(fields["qliquid"]),
            qrain=t(fields["qrain"]), qice=t(fields["qice"]), qsnow=t(fields["qsnow"]),
            qgraupel=t(fields["qgraupel"]), qcld=t(fields["qcld"]), ua=t(fields["ua"]), va=t(fields["va"]),
            delp=zeros.clone(), delz=t(fields["delz"]), omga=zeros.clone(), delprsi=t(fields["delprsi"]),
            wmp=t(fields["wmp"]), dz=zeros.clone(), tendency=zeros, land=t(fields["land"]))
        mp = Microphysics(StencilFactory((ni, nj, nz), device=dev), QuantityFactory((ni, nj, nz), device=dev),
                          GridData(t(self.area)), self.namelist, precision=self.precision)
        mp(state, timestep=float(self.namelist.dt_atmos))
        torch.cuda.synchronize()
        return {ser: ijk_to_microph(getattr(state, name).cpu().numpy()) for name, ser in self.out_vars.items()}
