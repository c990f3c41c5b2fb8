"""Physical constants and unit factors, with the exact values RH uses.

The numbers (CODATA-1986 era) are those of rh/constant.h:28-73 and
rh/rh.h:36-50; they are kept identical so that results agree with the
reference to round-off.  Derived expressions are evaluated in the same
order as the reference macros expand (e.g. ``LARMOR`` and ``E_ION_HMIN``
have no outer parentheses there).
"""
import math

CLIGHT = 2.99792458E+08
HPLANCK = 6.6260755E-34
KBOLTZMANN = 1.380658E-23
AMU = 1.6605402E-27
M_ELECTRON = 9.1093897E-31
Q_ELECTRON = 1.60217733E-19
EPSILON_0 = 8.854187817E-12
MU_0 = 1.2566370614E-06
RBOHR = 5.29177349E-11
E_RYDBERG = 2.1798741E-18
EV = 1.60217733E-19
THETA0 = 5.03974756E+03
ABARH = 7.42E-41

NM_TO_M = 1.0E-09
CM_TO_M = 1.0E-02
KM_TO_M = 1.0E+03
ERG_TO_JOULE = 1.0E-07
G_TO_KG = 1.0E-03
MICRON_TO_NM = 1.0E+03
MEGABARN_TO_M2 = 1.0E-22

PI = 3.14159265358979
SQRTPI = 1.77245385090551
LG10 = 2.30258509299404568402

# CUBE(CM_TO_M) and SQ(CM_TO_M) as the C macros evaluate them
CUBE_CM_TO_M = CM_TO_M * CM_TO_M * CM_TO_M
SQ_CM_TO_M = CM_TO_M * CM_TO_M

# LARMOR expands to  (Q_ELECTRON / (4.0*PI*M_ELECTRON)) * NM_TO_M
LARMOR_Q = Q_ELECTRON / (4.0 * PI * M_ELECTRON)
# E_ION_HMIN expands to  0.754 * EV
E_ION_HMIN = 0.754 * EV

AIR_TO_VACUUM_LIMIT = 199.9352
VACUUM_TO_AIR_LIMIT = 200.0000

MILLI = 1.0E-03


def pow10(x: float) -> float:
    """RH's POW10(x) = exp(LG10 * x) (rh/rh.h:50), not 10**x."""
    return math.exp(LG10 * x)


def air_to_vacuum(lam_air):
    """rh/vacuumtoair.c:51-66 (wavelengths in nm); scalar or sequence."""
    def one(la):
        if la >= AIR_TO_VACUUM_LIMIT:
            sqwave = (1.0E+07 / la) * (1.0E+07 / la)
            increase = (1.0000834213E+00 + 2.406030E+06 / (1.30E+10 - sqwave)
                        + 1.5997E+04 / (3.89E+09 - sqwave))
            return la * increase
        return la
    try:
        return [one(float(x)) for x in lam_air]
    except TypeError:
        return one(float(lam_air))


def vacuum_to_air(lam_vac):
    """rh/vacuumtoair.c:28-45 (nm); scalar or sequence."""
    def one(la):
        if la >= VACUUM_TO_AIR_LIMIT:
            sqwave = 1.0 / (la * la)
            reduction = 1.0 + 2.735182E-4 + (1.314182 + 2.76249E+4 * sqwave) * sqwave
            return la / reduction
        return la
    try:
        return [one(float(x)) for x in lam_vac]
    except TypeError:
        return one(float(lam_vac))
