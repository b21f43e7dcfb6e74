"""Physical constants and unit conventions of DiskDyn (src/ddyn.cuh:16-38).

Dynamics run in AU, yr and solar masses; every mass is stored as mu = G*M
with G = 4*pi^2 (src/ddyn.cuh:30, src/read_in.cu:76).
"""
MEARTH = 5.97e24            # kg
AU = 1.49598e11             # m
MSUN = 1.98892e30           # kg
RSUN = 695700000.0          # m
YEAR = 3.1556926e7          # s
LIGHT_AU = 63239.6717367    # speed of light in AU/yr
CSPEED = 299792458.0        # m/s
VCONV = 4740.57580894       # 1 AU/yr in m/s
PI = 3.14159265358979323846
PI2 = 6.28318530717958623200
GRAV = 4.0 * PI * PI        # G in AU^3 Msun^-1 yr^-2
PC = 3.08567758149137e16    # m
NSIZE = 2700                # size-bin table cap (src/ddyn.cuh:35)
VERSION = "1.01"            # data-file version (src/ddyn.cu:8)


def kms_to_au_per_yr(v_kms):
    """Stellar-wind speed km/s -> AU/yr, as small_RK4 passes cfg->vsw*year/AU (src/grid.cu:737)."""
    return v_kms * 1000.0 * YEAR / AU
