"""Orbital elements <-> Cartesian state and the barycentre correction.

coe2rv follows src/functions.cu:6-46 (perifocal frame rotated by Omega, omega,
i; mu = (M + m)(1 - beta)); recalc_com follows src/read_in.cu:344-375.  Both are
vectorised with numpy; the torch versions in diskdyn_amd.adjoint reuse the same
algebra for differentiation.
"""
import numpy as np


def coe2rv(M, m, a, e, inc, w, o, nu, beta=0.0):
    """Cartesian (x, y, z, vx, vy, vz) of an orbit around a body of mu = M.

    Argument order and meaning as the reference's coe2rv: M and m are G*mass of
    the central body and the orbiter, a semi-major axis [AU], e eccentricity,
    inc inclination, w argument of periapsis, o longitude of the node, nu true
    anomaly (radians), beta the radiation-pressure ratio that weakens gravity.
    """
    M, m, a, e, inc, w, o, nu, beta = np.broadcast_arrays(
        *[np.asarray(v, dtype=np.float64) for v in (M, m, a, e, inc, w, o, nu, beta)])
    p = a * (1.0 - e * e)
    mu = (M + m) * (1.0 - beta)
    cnu, snu = np.cos(nu), np.sin(nu)
    xe = p * cnu / (1.0 + e * cnu)
    ye = p * snu / (1.0 + e * cnu)
    vxe = -np.sqrt(mu / p) * snu
    vye = np.sqrt(mu / p) * (e + cnu)
    co, so, cw, sw, ci, si = np.cos(o), np.sin(o), np.cos(w), np.sin(w), np.cos(inc), np.sin(inc)
    r11 = co * cw - so * sw * ci
    r12 = -co * sw - so * cw * ci
    r21 = so * cw + co * sw * ci
    r22 = -so * sw + co * cw * ci
    r31 = sw * si
    r32 = cw * si
    return (xe * r11 + ye * r12, xe * r21 + ye * r22, xe * r31 + ye * r32,
            vxe * r11 + vye * r12, vxe * r21 + vye * r22, vxe * r31 + vye * r32)


def rv2coe(M, m, x, y, z, vx, vy, vz, beta=0.0):
    """Elements (a, e, inc) of a Cartesian state (the subset of src/functions.cu:48-154
    that tests and diagnostics use)."""
    mu = (M + m) * (1.0 - beta)
    R = np.sqrt(x * x + y * y + z * z)
    V2 = vx * vx + vy * vy + vz * vz
    hx, hy, hz = y * vz - z * vy, z * vx - x * vz, x * vy - y * vx
    H = np.sqrt(hx * hx + hy * hy + hz * hz)
    rv = x * vx + y * vy + z * vz
    c1 = V2 - mu / R
    ex, ey, ez = (c1 * x - rv * vx) / mu, (c1 * y - rv * vy) / mu, (c1 * z - rv * vz) / mu
    e = np.sqrt(ex * ex + ey * ey + ez * ez)
    a = -mu / 2.0 / (0.5 * V2 - mu / R)
    inc = np.arccos(np.clip(hz / H, -1.0, 1.0))
    return a, e, inc


def recalc_com(x, y, z, vx, vy, vz, mu):
    """Barycentre correction of star + planets, in place (src/read_in.cu:344-375).

    Positions are shifted to the centre of mass; the star (index 0) gets the
    velocity that zeroes the total momentum; planet velocities are unchanged.
    """
    n = len(mu)
    if n == 1:
        return
    M = float(np.sum(mu))
    xc, yc, zc = float(np.dot(mu, x)) / M, float(np.dot(mu, y)) / M, float(np.dot(mu, z)) / M
    x -= xc
    y -= yc
    z -= zc
    vx[0] = -float(np.dot(mu[1:], vx[1:])) / mu[0]
    vy[0] = -float(np.dot(mu[1:], vy[1:])) / mu[0]
    vz[0] = -float(np.dot(mu[1:], vz[1:])) / mu[0]


def kepler_true_anomaly(M, e, iters=60):
    """True anomaly from mean anomaly M (radians) and eccentricity e < 1.

    Newton iteration on Kepler's equation E - e sin E = M, started at
    E = M + e sin M, then nu = 2 atan2(sqrt(1+e) sin(E/2), sqrt(1-e) cos(E/2)).
    """
    M = np.asarray(M, dtype=np.float64)
    e = np.asarray(e, dtype=np.float64)
    E = M + e * np.sin(M)
    for _ in range(iters):
        f = E - e * np.sin(E) - M
        d = 1.0 - e * np.cos(E)
        dE = f / d
        E = E - dE
        if np.all(np.abs(dE) < 1e-15):
            break
    return 2.0 * np.arctan2(np.sqrt(1.0 + e) * np.sin(0.5 * E), np.sqrt(1.0 - e) * np.cos(0.5 * E))


def kepler_propagate_many(mu, r0, v0, t, iters=50):
    """Vectorised two-body propagation of bound orbits: r0, v0 (n, 3) around a
    fixed centre of G*M = mu, advanced by time t; returns (r, v) (n, 3).
    Elements -> eccentric anomaly (Newton on Kepler's equation) -> f and g."""
    r0 = np.asarray(r0, dtype=np.float64)
    v0 = np.asarray(v0, dtype=np.float64)
    R = np.linalg.norm(r0, axis=1)
    V2 = (v0 * v0).sum(1)
    a = 1.0 / (2.0 / R - V2 / mu)
    rv = (r0 * v0).sum(1)
    n = np.sqrt(mu / a ** 3)
    ecosE = 1.0 - R / a
    esinE = rv / np.sqrt(mu * a)
    e = np.hypot(ecosE, esinE)
    E0 = np.arctan2(esinE, ecosE)
    M1 = E0 - esinE + n * t
    E = M1 + e * np.sin(M1)
    for _ in range(iters):
        dE = (E - e * np.sin(E) - M1) / (1.0 - e * np.cos(E))
        E = E - dE
        if np.all(np.abs(dE) < 1e-15):
            break
    dE = E - E0
    f = 1.0 - a / R * (1.0 - np.cos(dE))
    g = t - (dE - np.sin(dE)) / n
    r1 = f[:, None] * r0 + g[:, None] * v0
    R1 = np.linalg.norm(r1, axis=1)
    fdot = -np.sqrt(mu * a) / (R * R1) * np.sin(dE)
    gdot = 1.0 - a / R1 * (1.0 - np.cos(dE))
    return r1, fdot[:, None] * r0 + gdot[:, None] * v0


def kepler_propagate(mu, x, y, z, vx, vy, vz, t):
    """Exact two-body propagation of (r, v) around a fixed centre of G*M = mu by time t.

    Universal-variable-free form for bound orbits: elements -> mean anomaly
    advance -> Cartesian.  Used as the closed form the integrator is checked on.
    """
    r0 = np.array([x, y, z], dtype=np.float64)
    v0 = np.array([vx, vy, vz], dtype=np.float64)
    R = np.linalg.norm(r0)
    V2 = float(v0 @ v0)
    a = 1.0 / (2.0 / R - V2 / mu)
    h = np.cross(r0, v0)
    evec = np.cross(v0, h) / mu - r0 / R
    e = np.linalg.norm(evec)
    n = np.sqrt(mu / a ** 3)
    # eccentric anomaly at t0
    cosE0 = (1.0 - R / a) / e
    sinE0 = float(r0 @ v0) / (e * np.sqrt(mu * a))
    E0 = np.arctan2(sinE0, cosE0)
    M0 = E0 - e * np.sin(E0)
    M1 = M0 + n * t
    E = M1
    for _ in range(100):
        dE = (E - e * np.sin(E) - M1) / (1.0 - e * np.cos(E))
        E -= dE
        if abs(dE) < 1e-16:
            break
    # f and g functions
    dEa = E - E0
    f = 1.0 - a / R * (1.0 - np.cos(dEa))
    g = t - (dEa - np.sin(dEa)) / n
    r1 = f * r0 + g * v0
    R1 = np.linalg.norm(r1)
    fdot = -np.sqrt(mu * a) / (R * R1) * np.sin(dEa)
    gdot = 1.0 - a / R1 * (1.0 - np.cos(dEa))
    v1 = fdot * r0 + gdot * v0
    return r1, v1
