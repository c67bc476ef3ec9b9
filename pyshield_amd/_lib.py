"""ctypes binding of libshield_mp.so (the C ABI in include/shield_mp.h).

The library is built in-tree by ``pyshield_amd.build``.  Nothing here falls
back to another implementation: if the library is missing or no GPU is
visible, the entry points raise.
"""

import ctypes
import os

from ._config import MP_NAMELIST_FLOATS, MP_NAMELIST_INTS

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libshield_mp.so")

SHIELD_MP_FP64 = 0
SHIELD_MP_MIXED = 1
MODES = {"fp64": SHIELD_MP_FP64, "mixed": SHIELD_MP_MIXED}

# every symbol the header declares
EXPORTS = (
    "shield_mp_setup", "shield_mp_update_timestep", "shield_mp_forward", "shield_mp_forward_update",
    "shield_prepare_microphysics", "shield_update_physics_state_with_tendencies",
    "shield_fill_gfs_delp", "shield_prepare_tendencies_and_update_tracers",
    "shield_strerror", "shield_device_count", "shield_mp_layout", "shield_physics_prologue",
    "shield_mp_adjoint", "shield_mp_adjoint_workspace_bytes", "shield_apply_physics_to_dycore",
    "shield_mp_adjoint_noalloc",
)

_D = ctypes.c_double
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_DP = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p


class ShieldMpNamelist(ctypes.Structure):
    _fields_ = ([(n, _D) for n in MP_NAMELIST_FLOATS]
                + [(n, _I32) for n in MP_NAMELIST_INTS] + [("reset_qaz", _I32)])


class ShieldMpParams(ctypes.Structure):
    _fields_ = [
        ("nml", ShieldMpNamelist),
        *[(n, _D) for n in ("c_air", "c_vap", "d0_vap", "lv00", "cpaut", "so3", "zs", "log_10",
                            "tice0", "t_wfr", "fac_rc", "cracs", "csacr", "cgacr", "cgacs")],
        ("acco", (_D * 4) * 3),
        *[(n, _D) for n in ("csacw", "csaci", "cgacw", "cgaci", "cracw")],
        ("cssub", _D * 5), ("cgsub", _D * 5), ("crevp", _D * 5), ("cgfr", _D * 2),
        ("csmlt", _D * 5), ("cgmlt", _D * 5),
        ("ces0", _D),
        *[(n, _D) for n in ("timestep", "mpdt", "rdt", "dts", "dt_rain", "rdts", "dt_evap",
                            "fac_i2s", "fac_g2v", "fac_v2g", "fac_imlt", "fac_l2v")],
        ("ntimes", _I32), ("use_ccn", _I32), ("do_sedi_w", _I32), ("accumulate", _I32),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            if name in ("nml", "accumulate"):
                continue
            v = getattr(self, name)
            if name == "acco":
                for i in range(3):
                    for k in range(4):
                        out[f"acco{i}{k}"] = v[i][k]
            elif hasattr(v, "__len__"):
                for i, x in enumerate(v):
                    out[f"{name}{i}"] = x
            else:
                out[name] = v
        return out


class ShieldMpShape(ctypes.Structure):
    _fields_ = [(n, _I64) for n in ("ni", "nj", "nk", "si", "sj", "si2", "sj2")]


FIELD_NAMES = (
    "pt", "qv", "ql", "qr", "qi", "qs", "qg", "qa", "ua", "va", "delp", "delz", "w",
    "land", "area",
    "pt_dt", "qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt", "qa_dt", "udt", "vdt",
    "qaz", "rain", "snow", "ice", "graupel",
)


class ShieldMpFields(ctypes.Structure):
    _fields_ = [(n, _VP) for n in FIELD_NAMES]


# shield_mp_adjoint_fields: cotangents of the forward's outputs (read), of its
# inputs (written), and the gradient of the six process parameters
ADJ_OUT_NAMES = ("pt_dt", "qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt", "qa_dt", "udt", "vdt",
                 "rain", "snow", "ice", "graupel")
ADJ_IN_NAMES = ("pt", "qv", "ql", "qr", "qi", "qs", "qg", "qa", "ua", "va", "delp", "delz", "qaz")
AD_PARAM_NAMES = ("fac_rc", "vr_fac", "vs_fac", "vg_fac", "vi_fac", "qi_gen")


class ShieldMpAdjointFields(ctypes.Structure):
    # d_w / g_w: the cotangents of w after and before the call (do_sedi_w)
    _fields_ = [("d_" + n, _VP) for n in ADJ_OUT_NAMES] + [("g_" + n, _VP) for n in ADJ_IN_NAMES] + [
        ("params", _VP), ("d_w", _VP), ("g_w", _VP)]


# physics_updated_* order of shield_update_fields.updated and the FV3 tracer
# order of shield_update_fields.tracers (include/shield_mp.h)
UPDATED_NAMES = ("qvapor", "qliquid", "qrain", "qice", "qsnow", "qgraupel", "qcld", "pt", "ua", "va")
TRACER_NAMES = ("qvapor", "qliquid", "qrain", "qsnow", "qice", "qgraupel")


class ShieldPrologueFields(ctypes.Structure):
    _fields_ = [("delz", _VP), ("delp", _VP), ("q", _VP * 7), ("qsgs_tke", _VP), ("pt", _VP), ("phii", _VP),
                ("prsi", _VP), ("prsik", _VP), ("dm", _VP), ("delprsi", _VP), ("del_gz", _VP), ("phil", _VP),
                ("omga", _VP), ("dz", _VP), ("wmp", _VP), ("tend", _VP * 10), ("si_i", _I64), ("sj_i", _I64),
                ("ptop", _D)]


class ShieldUpdateFields(ctypes.Structure):
    _fields_ = [("updated", _VP * 10), ("u_dt", _VP), ("v_dt", _VP), ("pt_dt", _VP), ("tracers", _VP * 6),
                ("delp", _VP), ("prsi", _VP), ("si_i", _I64), ("sj_i", _I64), ("dt", _D), ("rdt", _D),
                ("q_min", _D)]


DYCORE_PTR_NAMES = ("qvapor", "qliquid", "qrain", "qsnow", "qice", "qgraupel", "pt", "t_dt", "delp", "pe", "peln",
                    "pk", "ps", "ua", "va", "u_srf", "v_srf", "u", "v", "u_dt", "v_dt", "vlon1", "vlon2",
                    "vlon3", "vlat1", "vlat2", "vlat3", "es1_1", "es1_2", "es1_3", "ew2_1", "ew2_2", "ew2_3",
                    "edge_vect_w", "edge_vect_e", "edge_vect_s", "edge_vect_n")


class ShieldDycoreFields(ctypes.Structure):
    _fields_ = [(n, _VP) for n in DYCORE_PTR_NAMES] + [
        ("dt", _D), ("con_cp", _D), ("kappa", _D), ("grid_type", _I32), ("npx", _I32), ("npy", _I32),
        ("pad_", _I32)]


_lib = None


class ShieldError(RuntimeError):
    pass


def lib():
    """Load libshield_mp.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ShieldError(
                f"{LIB_PATH} is missing: build it with `python -m pyshield_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.shield_mp_setup.argtypes = [ctypes.POINTER(ShieldMpNamelist), _D, ctypes.POINTER(ShieldMpParams)]
        L.shield_mp_update_timestep.argtypes = [ctypes.POINTER(ShieldMpParams), _D]
        L.shield_mp_forward.argtypes = [ctypes.POINTER(ShieldMpShape), ctypes.POINTER(ShieldMpParams),
                                        ctypes.POINTER(ShieldMpFields), ctypes.c_int, _VP]
        L.shield_mp_forward_update.argtypes = [ctypes.POINTER(ShieldMpShape), ctypes.POINTER(ShieldMpParams),
                                               ctypes.POINTER(ShieldMpFields), ctypes.POINTER(ShieldUpdateFields),
                                               ctypes.c_int, _VP]
        L.shield_physics_prologue.argtypes = [ctypes.POINTER(ShieldMpShape), ctypes.POINTER(ShieldPrologueFields),
                                              _VP]
        L.shield_apply_physics_to_dycore.argtypes = [ctypes.POINTER(ShieldMpShape),
                                                     ctypes.POINTER(ShieldDycoreFields), _VP]
        L.shield_prepare_microphysics.argtypes = [ctypes.POINTER(ShieldMpShape), _VP, _VP, _VP, _VP, _VP,
                                                  _VP, _VP, _VP * 10, _VP]
        L.shield_update_physics_state_with_tendencies.argtypes = [ctypes.POINTER(ShieldMpShape), _VP * 10,
                                                                  _VP * 10, _VP * 10, _D, _VP]
        L.shield_fill_gfs_delp.argtypes = [ctypes.POINTER(ShieldMpShape), _VP, _VP, _D, _VP]
        L.shield_prepare_tendencies_and_update_tracers.argtypes = (
            [ctypes.POINTER(ShieldMpShape)] + [_VP] * 23 + [_D, _VP])
        L.shield_strerror.argtypes = [ctypes.c_int]
        L.shield_strerror.restype = ctypes.c_char_p
        L.shield_mp_layout.argtypes = [_I64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int)]
        L.shield_mp_adjoint.argtypes = [ctypes.POINTER(ShieldMpShape), ctypes.POINTER(ShieldMpParams),
                                        ctypes.POINTER(ShieldMpFields), ctypes.POINTER(ShieldMpAdjointFields),
                                        _VP, ctypes.c_size_t, _VP]
        L.shield_mp_adjoint_noalloc.argtypes = [ctypes.POINTER(ShieldMpShape), ctypes.POINTER(ShieldMpParams),
                                                ctypes.POINTER(ShieldMpFields), ctypes.POINTER(ShieldMpAdjointFields),
                                                _VP]
        L.shield_mp_adjoint_workspace_bytes.argtypes = [ctypes.POINTER(ShieldMpShape),
                                                        ctypes.POINTER(ShieldMpParams), _I64]
        for fn in EXPORTS:
            if fn != "shield_strerror":
                getattr(L, fn).restype = ctypes.c_int
        L.shield_mp_adjoint_workspace_bytes.restype = ctypes.c_size_t
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        msg = lib().shield_strerror(rc).decode()
        raise ShieldError(f"{what} failed ({rc}): {msg}")


def make_namelist(values: dict) -> ShieldMpNamelist:
    s = ShieldMpNamelist()
    for n in MP_NAMELIST_FLOATS:
        setattr(s, n, float(values[n]))
    for n in MP_NAMELIST_INTS:
        setattr(s, n, int(values[n]))
    s.reset_qaz = int(values.get("reset_qaz", 0))
    return s


def setup(values: dict, timestep: float) -> ShieldMpParams:
    p = ShieldMpParams()
    check(lib().shield_mp_setup(ctypes.byref(make_namelist(values)), float(timestep), ctypes.byref(p)),
          "shield_mp_setup")
    return p


def layout(nk, mode="mixed", ntimes=1):
    """(columns per wavefront, slots) the forward kernel uses for nk levels."""
    P, S = ctypes.c_int(), ctypes.c_int()
    m = MODES[mode] if isinstance(mode, str) else int(mode)
    check(lib().shield_mp_layout(int(nk), m, int(ntimes), ctypes.byref(P), ctypes.byref(S)),
          "shield_mp_layout")
    return P.value, S.value


def device_count():
    return lib().shield_device_count()
