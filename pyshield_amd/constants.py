"""Physical and microphysics constants of the GFDL cloud microphysics path.

Physical constants are the GFS set of ``ndsl.constants`` (the reference imports
them as ``constants.*``; NDSL is not vendored, so the values are stated here and
printed with every benchmark result).  Microphysics constants follow
pySHiELD/functions/microphysics_funcs.py:7-42.
"""

# ndsl.constants (GFS set)
GRAV = 9.80665
RGRAV = 1.0 / GRAV
RDGAS = 287.05
RVGAS = 461.50
HLV = 2.5e6
HLF = 3.3358e5
CP_AIR = 1004.6
TICE = 273.16
C_LIQ = 4185.5
C_ICE = 1972.0
E00 = 611.21
T_MIN = 178.0
KAPPA = RDGAS / CP_AIR
CV_AIR = CP_AIR - RDGAS
CV_VAP = 3.0 * RVGAS
CP_VAP = 4.0 * RVGAS
ZVIR = RVGAS / RDGAS - 1.0
EPS = RDGAS / RVGAS
DC_VAP = CP_VAP - C_LIQ
DC_ICE = C_LIQ - C_ICE
D2ICE = DC_VAP + DC_ICE
LV0 = HLV - DC_VAP * TICE
LI00 = HLF - DC_ICE * TICE
LI2 = LV0 + LI00
T_SAT_MIN = TICE - 160.0
LAT2 = (HLV + HLF) ** 2

# microphysics_funcs.py:7-42
VCONS = 6.6280504
VCONG = 87.2382675
NORMS = 942477796.076938
NORMG = 5026548245.74367
VCONR = 2503.23638966667
NORMR = 25132741228.7183
THR = 1.0e-8
THI = 1.0e-8
THG = 1.0e-8
THS = 1.0e-8
AA = -4.14122e-5
BB = -0.00538922
CC = -0.0516344
DD_FS = 0.00216078
EE = 1.9714
VR_MIN = 1.0e-3
VF_MIN = 1.0e-5
P_MIN = 100.0
DT_FR = 8.0
SFCRHO = 1.2
RHOS = 1.0e2
RHOG = 4.0e2
RHOR = 1.0e3
DZ_MIN_FLIP = 1.0e-2
QCMIN = 1.0e-12
QRMIN = 1.0e-8
QVMIN = 1.0e-20


def as_dict():
    """All constants of this module as a name -> value dict."""
    return {k: v for k, v in globals().items() if k.isupper()}
