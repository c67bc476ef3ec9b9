"""TEST INFRASTRUCTURE -- the FP64 oracle of the GFDL microphysics path.

``oracle/gfdl_mp_oracle.c`` is a CPU restatement of pySHiELD's
``Microphysics.__call__`` (pySHiELD/stencils/microphysics.py:1981-2263); this
module loads it through ctypes.  ``oracle/mp_numpy.py`` is a second,
stage-faithful NumPy restatement used to cross-check the C one.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
may import this package, and only as the checker / CPU baseline.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libmp_oracle.so")

NZMAX = 256

_FLOATS = (
    "c_cracw", "c_paut", "c_pgacs", "c_psaci", "ccn_l", "ccn_o",
    "vs_fac", "vg_fac", "vi_fac", "vr_fac",
    "tau_imlt", "tau_i2s", "tau_g2v", "tau_v2g", "tau_l2v",
    "ql_mlt", "qs_mlt", "t_sub", "qi_gen", "qi_lim", "qi0_max",
    "dw_ocean", "dw_land", "rh_inc", "rh_inr",
    "qc_crt", "qi0_crt", "qs0_crt", "rthresh", "mp_time",
    "vg_max", "vi_max", "vr_max", "vs_max",
    "tice", "alin", "clin",
)
_INTS = (
    "hydrostatic", "do_qa", "const_vg", "const_vi", "const_vr", "const_vs", "de_ice",
    "rad_snow", "rad_rain", "do_sedi_heat", "do_sedi_w", "fast_sat_adj",
    "fix_negative", "irain_f", "prog_ccn", "sedi_transport", "use_ppm",
    "z_slope_ice", "z_slope_liq", "reset_qaz",
)


class OracleNml(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in _FLOATS] + [(n, ctypes.c_int32) for n in _INTS]


_DP = ctypes.POINTER(ctypes.c_double)
_FIELD_NAMES = (
    "pt", "qv", "ql", "qr", "qi", "qs", "qg", "qa", "ua", "va", "delp", "delz", "w",
    "land", "area",
    "pt_dt", "qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt", "qa_dt", "udt", "vdt",
    "qaz", "rain", "snow", "ice", "graupel",
)


class OracleFields(ctypes.Structure):
    _fields_ = [(n, _DP) for n in _FIELD_NAMES]


# order of oracle_mp_derived's output vector
DERIVED_NAMES = (
    ["c_air", "c_vap", "d0_vap", "lv00", "cpaut", "so3", "zs", "log_10", "tice0", "t_wfr",
     "fac_rc", "cracs", "csacr", "cgacr", "cgacs"]
    + [f"acco{i}{k}" for i in range(3) for k in range(4)]
    + ["csacw", "csaci", "cgacw", "cgaci", "cracw"]
    + [f"cssub{i}" for i in range(5)] + [f"cgsub{i}" for i in range(5)]
    + [f"crevp{i}" for i in range(5)] + ["cgfr0", "cgfr1"]
    + [f"csmlt{i}" for i in range(5)] + [f"cgmlt{i}" for i in range(5)]
    + ["ces0", "timestep", "mpdt", "rdt", "dts", "dt_rain", "rdts", "dt_evap",
       "fac_i2s", "fac_g2v", "fac_v2g", "fac_imlt", "fac_l2v", "ntimes"]
)

_libs = {}


def build(force=False):
    """Compile the C restatement with gcc (make -C oracle)."""
    if force or not os.path.exists(LIB_PATH) or (
        os.path.getmtime(LIB_PATH) < max(os.path.getmtime(os.path.join(HERE, f))
                                         for f in ("gfdl_mp_oracle.c", "mp_oracle.h"))):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def build_native(outdir):
    """Compile the C restatement for the CPU baseline: -O3 -march=native on
    the host that will run it (make -C oracle native).  Returns its path."""
    subprocess.run(["make", "-s", "-C", HERE, "native", f"NATIVE_OUT={outdir}"], check=True)
    return os.path.join(outdir, "libmp_oracle_native.so")


def lib(path=None):
    """The ctypes library of the checker build, or of another build of the
    same source (``path``, e.g. build_native's)."""
    path = path or LIB_PATH
    if path not in _libs:
        if path == LIB_PATH and not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(path)
        L.oracle_mp_run.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.POINTER(OracleNml),
                                    ctypes.c_double, ctypes.POINTER(OracleFields), ctypes.c_int]
        L.oracle_mp_run.restype = ctypes.c_int
        L.oracle_mp_derived.argtypes = [ctypes.POINTER(OracleNml), ctypes.c_double, _DP, ctypes.c_int]
        L.oracle_mp_derived.restype = ctypes.c_int
        for fn in ("oracle_wqs1", "oracle_iqs1"):
            getattr(L, fn).argtypes = [ctypes.c_double, ctypes.c_double]
            getattr(L, fn).restype = ctypes.c_double
        for fn in ("oracle_wqs2", "oracle_iqs2"):
            getattr(L, fn).argtypes = [ctypes.c_double, ctypes.c_double, _DP]
            getattr(L, fn).restype = ctypes.c_double
        L.oracle_fall_column.argtypes = [ctypes.c_int, ctypes.c_double, _DP, _DP, _DP, _DP, _DP]
        L.oracle_fall_column.restype = ctypes.c_double
        _libs[path] = L
    return _libs[path]


def make_nml(nml: dict) -> OracleNml:
    s = OracleNml()
    for n in _FLOATS:
        setattr(s, n, float(nml[n]))
    for n in _INTS:
        setattr(s, n, int(nml.get(n, 0)))
    return s


def derived(nml: dict, timestep: float) -> dict:
    out = np.zeros(len(DERIVED_NAMES) + 8)
    n = lib().oracle_mp_derived(ctypes.byref(make_nml(nml)), timestep,
                                out.ctypes.data_as(_DP), len(out))
    return dict(zip(DERIVED_NAMES, out[:n]))


def wqs1(t, den):
    return lib().oracle_wqs1(t, den)


def iqs1(t, den):
    return lib().oracle_iqs1(t, den)


def wqs2(t, den):
    d = ctypes.c_double()
    v = lib().oracle_wqs2(t, den, ctypes.byref(d))
    return v, d.value


def iqs2(t, den):
    d = ctypes.c_double()
    v = lib().oracle_iqs2(t, den, ctypes.byref(d))
    return v, d.value


def fall_column(dt, dz1, dp1, vt, q):
    """One implicit upwind fall of a single column (see oracle_fall_column)."""
    nz = len(q)
    arrs = [np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in (dz1, dp1, vt)]
    qq = np.array(q, dtype=np.float64)
    flux = np.zeros(nz)
    s = lib().oracle_fall_column(nz, float(dt), *(a.ctypes.data_as(_DP) for a in arrs),
                                 qq.ctypes.data_as(_DP), flux.ctypes.data_as(_DP))
    return qq, flux, s


TEND_NAMES = ("pt_dt", "qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt", "qa_dt", "udt", "vdt")
PRECIP_NAMES = ("rain", "snow", "ice", "graupel")


def run(nml: dict, timestep: float, state: dict, *, qaz=None, tend=None, nthreads=None, library=None):
    """One Microphysics call on the CPU.

    ``state`` holds float64 arrays [..., nz] for pt qv ql qr qi qs qg qa ua va
    delp delz (w optional) and [...] for land, area.  Returns a dict with the
    ten tendencies, ``qaz``, ``w`` (when given), the updated ``qi``/``qs`` (which
    the reference rewrites when de_ice) and the 2D precipitation in mm/day.
    ``library`` selects another build of the restatement (the CPU baseline's).
    """
    nz = state["pt"].shape[-1]
    ncol = state["pt"].size // nz
    if nthreads is None:
        nthreads = os.cpu_count() or 1

    def arr3(x):
        return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(ncol, nz))

    keep = []
    f = OracleFields()
    out = {}
    for n in ("pt", "qv", "ql", "qr", "qg", "qa", "ua", "va", "delp", "delz"):
        a = arr3(state[n])
        keep.append(a)
        setattr(f, n, a.ctypes.data_as(_DP))
    for n in ("qi", "qs"):
        a = arr3(state[n]).copy()
        out[n] = a
        setattr(f, n, a.ctypes.data_as(_DP))
    if state.get("w") is not None:
        a = arr3(state["w"]).copy()
        out["w"] = a
        setattr(f, "w", a.ctypes.data_as(_DP))
    for n in ("land", "area"):
        a = np.ascontiguousarray(np.asarray(state[n], dtype=np.float64).reshape(ncol))
        keep.append(a)
        setattr(f, n, a.ctypes.data_as(_DP))
    for n in TEND_NAMES:
        a = arr3(tend[n]).copy() if (tend is not None and n in tend) else np.zeros((ncol, nz))
        out[n] = a
        setattr(f, n, a.ctypes.data_as(_DP))
    a = arr3(qaz).copy() if qaz is not None else np.zeros((ncol, nz))
    out["qaz"] = a
    f.qaz = a.ctypes.data_as(_DP)
    for n in PRECIP_NAMES:
        a = np.zeros(ncol)
        out[n] = a
        setattr(f, n, a.ctypes.data_as(_DP))
    rc = lib(library).oracle_mp_run(ncol, nz, ctypes.byref(make_nml(nml)), float(timestep),
                             ctypes.byref(f), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle_mp_run failed with code {rc}")
    shp3 = state["pt"].shape
    shp2 = shp3[:-1]
    res = {}
    for k, v in out.items():
        res[k] = v.reshape(shp2) if k in PRECIP_NAMES else v.reshape(shp3)
    return res
