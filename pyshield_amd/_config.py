"""PhysicsConfig: the namelist of the physics suite (pySHiELD/_config.py:16-200).

Same class name, field names, defaults semantics and error behaviour as the
reference: unknown schemes raise ``NotImplementedError`` in ``__post_init__``
(pySHiELD/_config.py:111-118).  The defaults of the microphysics fields are the
NDSL ``NamelistDefaults`` values the reference reads (NDSL is not vendored, so
they are stated here).  ``from_namelist`` accepts any object with the same
attribute names (an NDSL ``Namelist`` or a SimpleNamespace).
"""

import dataclasses
from enum import Enum, unique
from typing import List, Optional, Tuple


DEFAULT_INT = 0
DEFAULT_BOOL = False
DEFAULT_SCHEMES = ["GFS_microphysics"]


@unique
class PHYSICS_PACKAGES(Enum):
    GFS_microphysics = "GFS_microphysics"


def _scheme_names():
    return [p.value for p in PHYSICS_PACKAGES]


# NamelistDefaults for the GFDL microphysics (values of the reference's
# defaults; printed with every benchmark line through PhysicsConfig.mp_namelist)
class NamelistDefaults:
    c_cracw = 0.9
    c_paut = 0.55
    c_pgacs = 2.0e-3
    c_psaci = 0.02
    ccn_l = 300.0
    ccn_o = 100.0
    const_vg = False
    const_vi = False
    const_vr = False
    const_vs = False
    vs_fac = 1.0
    vg_fac = 1.0
    vi_fac = 1.0
    vr_fac = 1.0
    de_ice = False
    layout = (1, 1)
    tau_imlt = 600.0
    tau_i2s = 1000.0
    tau_g2v = 900.0
    tau_v2g = 21600.0
    ql_mlt = 2.0e-3
    qs_mlt = 1.0e-6
    t_sub = 184.0
    qi_gen = 1.82e-6
    qi_lim = 1.0
    qi0_max = 1.0e-4
    rad_snow = True
    rad_rain = True
    dw_ocean = 0.10
    dw_land = 0.20
    tau_l2v = 300.0
    c2l_ord = 4
    do_sedi_heat = True
    do_sedi_w = False
    fast_sat_adj = False
    qc_crt = 5.0e-8
    fix_negative = False
    irain_f = 0
    mp_time = 150.0
    prog_ccn = False
    qi0_crt = 1.0e-4
    qs0_crt = 1.0e-3
    rh_inc = 0.25
    rh_inr = 0.25
    rthresh = 1.0e-5
    sedi_transport = True
    use_ppm = False
    vg_max = 12.0
    vi_max = 0.5
    vr_max = 12.0
    vs_max = 2.0
    z_slope_ice = False
    z_slope_liq = True
    tice = 273.16
    alin = 842.0
    clin = 4.8


# order of the FP64 and int32 members of shield_mp_namelist (include/shield_mp.h)
MP_NAMELIST_FLOATS = (
    "c_cracw", "c_paut", "c_pgacs", "c_psaci", "ccn_l", "ccn_o",
    "vs_fac", "vg_fac", "vi_fac", "vr_fac",
    "tau_imlt", "tau_i2s", "tau_g2v", "tau_v2g", "tau_l2v",
    "ql_mlt", "qs_mlt", "t_sub", "qi_gen", "qi_lim", "qi0_max",
    "dw_ocean", "dw_land", "rh_inc", "rh_inr",
    "qc_crt", "qi0_crt", "qs0_crt", "rthresh", "mp_time",
    "vg_max", "vi_max", "vr_max", "vs_max",
    "tice", "alin", "clin",
)
MP_NAMELIST_INTS = (
    "hydrostatic", "do_qa", "const_vg", "const_vi", "const_vr", "const_vs", "de_ice",
    "rad_snow", "rad_rain", "do_sedi_heat", "do_sedi_w", "fast_sat_adj",
    "fix_negative", "irain_f", "prog_ccn", "sedi_transport", "use_ppm",
    "z_slope_ice", "z_slope_liq",
)


@dataclasses.dataclass
class PhysicsConfig:
    dt_atmos: int = DEFAULT_INT
    hydrostatic: bool = DEFAULT_BOOL
    npx: int = DEFAULT_INT
    npy: int = DEFAULT_INT
    npz: int = DEFAULT_INT
    nwat: int = DEFAULT_INT
    schemes: List = None
    do_qa: bool = DEFAULT_BOOL
    c_cracw: float = NamelistDefaults.c_cracw
    c_paut: float = NamelistDefaults.c_paut
    c_pgacs: float = NamelistDefaults.c_pgacs
    c_psaci: float = NamelistDefaults.c_psaci
    ccn_l: float = NamelistDefaults.ccn_l
    ccn_o: float = NamelistDefaults.ccn_o
    const_vg: bool = NamelistDefaults.const_vg
    const_vi: bool = NamelistDefaults.const_vi
    const_vr: bool = NamelistDefaults.const_vr
    const_vs: bool = NamelistDefaults.const_vs
    vs_fac: float = NamelistDefaults.vs_fac
    vg_fac: float = NamelistDefaults.vg_fac
    vi_fac: float = NamelistDefaults.vi_fac
    vr_fac: float = NamelistDefaults.vr_fac
    de_ice: bool = NamelistDefaults.de_ice
    layout: Tuple[int, int] = NamelistDefaults.layout
    tau_imlt: float = NamelistDefaults.tau_imlt
    tau_i2s: float = NamelistDefaults.tau_i2s
    tau_g2v: float = NamelistDefaults.tau_g2v
    tau_v2g: float = NamelistDefaults.tau_v2g
    ql_mlt: float = NamelistDefaults.ql_mlt
    qs_mlt: float = NamelistDefaults.qs_mlt
    t_sub: float = NamelistDefaults.t_sub
    qi_gen: float = NamelistDefaults.qi_gen
    qi_lim: float = NamelistDefaults.qi_lim
    qi0_max: float = NamelistDefaults.qi0_max
    rad_snow: bool = NamelistDefaults.rad_snow
    rad_rain: bool = NamelistDefaults.rad_rain
    dw_ocean: float = NamelistDefaults.dw_ocean
    dw_land: float = NamelistDefaults.dw_land
    tau_l2v: float = NamelistDefaults.tau_l2v
    c2l_ord: int = NamelistDefaults.c2l_ord
    do_sedi_heat: bool = NamelistDefaults.do_sedi_heat
    do_sedi_w: bool = NamelistDefaults.do_sedi_w
    fast_sat_adj: bool = NamelistDefaults.fast_sat_adj
    qc_crt: float = NamelistDefaults.qc_crt
    fix_negative: bool = NamelistDefaults.fix_negative
    irain_f: int = NamelistDefaults.irain_f
    mp_time: float = NamelistDefaults.mp_time
    prog_ccn: bool = NamelistDefaults.prog_ccn
    qi0_crt: float = NamelistDefaults.qi0_crt
    qs0_crt: float = NamelistDefaults.qs0_crt
    rh_inc: float = NamelistDefaults.rh_inc
    rh_inr: float = NamelistDefaults.rh_inr
    rthresh: float = NamelistDefaults.rthresh
    sedi_transport: bool = NamelistDefaults.sedi_transport
    use_ppm: bool = NamelistDefaults.use_ppm
    vg_max: float = NamelistDefaults.vg_max
    vi_max: float = NamelistDefaults.vi_max
    vr_max: float = NamelistDefaults.vr_max
    vs_max: float = NamelistDefaults.vs_max
    z_slope_ice: bool = NamelistDefaults.z_slope_ice
    z_slope_liq: bool = NamelistDefaults.z_slope_liq
    tice: float = NamelistDefaults.tice
    alin: float = NamelistDefaults.alin
    clin: float = NamelistDefaults.clin
    namelist_override: Optional[str] = None
    # Extension (not in the reference): zero the cloud-fraction accumulator
    # at every call, as GFDL Fortran does.  False keeps pySHiELD's behaviour.
    reset_qaz: bool = False

    def __post_init__(self):
        if self.schemes is None:
            self.schemes = DEFAULT_SCHEMES
        package_schemes = []
        for scheme in self.schemes:
            name = scheme.value if isinstance(scheme, PHYSICS_PACKAGES) else scheme
            if name not in _scheme_names():
                raise NotImplementedError(f"{scheme} physics scheme not implemented")
            package_schemes.append(PHYSICS_PACKAGES[name])
        self.schemes = package_schemes
        if self.namelist_override is not None:
            # The reference reads a Fortran namelist with f90nml, which is not
            # available here; the override path is rejected loudly.
            raise NotImplementedError(
                "namelist_override needs f90nml, which is not installed"
            )

    @classmethod
    def from_namelist(cls, namelist) -> "PhysicsConfig":
        """Build from an object with the reference's namelist attribute names
        (pySHiELD/_config.py:134-200)."""
        names = [f.name for f in dataclasses.fields(cls)]
        kwargs = {}
        for name in names:
            if name in ("schemes", "namelist_override", "reset_qaz"):
                continue
            if hasattr(namelist, name):
                kwargs[name] = getattr(namelist, name)
        return cls(**kwargs)

    def mp_namelist(self) -> dict:
        """The microphysics namelist values, in shield_mp_namelist order."""
        out = {k: float(getattr(self, k)) for k in MP_NAMELIST_FLOATS}
        out.update({k: int(bool(getattr(self, k))) if k != "irain_f" else int(self.irain_f)
                    for k in MP_NAMELIST_INTS})
        out["reset_qaz"] = int(bool(self.reset_qaz))
        return out
