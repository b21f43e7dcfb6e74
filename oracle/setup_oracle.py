"""CPU restatement of the reference's particle set-up for a .param run.

TEST INFRASTRUCTURE ONLY: tests import this module as the checker of the
library's ddyn_read_param planets and ddyn_generate.  Nothing in the library
calls it, and no product path may.

What it restates, from the reference merope82/DiskDyn v1.01:

- ``Ran3``: the random stream of src/ran3.cu:12-48 (Knuth's subtractive
  generator, TAOCP vol. 2, 3.2.2 and 3.6, in Numerical Recipes' parameters),
  written here as a growing list Y with Y[n] = Y[n-55] - Y[n-24] mod 10^9:
  the 55 starting words are Y[-54..0], the first 220 terms after them are the
  warm-up passes, and the deviates are Y[221], Y[222], ... times the double
  nearest 10^-9.
- ``getbod`` (src/functions.cu:235-240) in its mixed precision: numpy's
  longdouble is the x87 80-bit format, C's long double on x86-64.
- ``coe2rv`` (src/functions.cu:6-46) with the reference's products and sums.
- ``genorb_disk`` / ``genorb_blob`` (src/functions.cu:157-232), drawing in the
  order the reference's host compiler evaluates them (left operand of a
  product first).
- ``planets`` (add_star, add_planets, recalc_COM; src/read_in.cu:63-115,
  344-375), ``disk_record`` / ``blob_record`` (add_disks, add_blobs;
  :117-234), ``dust_grid`` (detminmax, confCdisks, confCblobs; :236-342) and
  ``setup`` (the particle part of Grid::setup, src/grid.cu:62-189).

Python's float arithmetic is IEEE double without fused operations and its
math functions call the C library's, so on one machine these values are the
reference's C expressions bit for bit.  The reference is built by nvcc with
-O3 (src/Makefile), whose host compiler turns sin(x) and cos(x) of one
argument into one sincos call; the C library's sincos differs from sin and
cos in the last bit for about one argument in 700, so ``sincos`` here calls
the C library's.  C's round() (half away from zero) is ``c_round``; Python's
round() would round half to even.

What pins it (tests/test_setup_stream.py): the lag relation of the stream
from each output back 55 and 24 outputs, the permutation the initialisation
visits, uniformity, getbod's interval mass against the closed form, coe2rv
against the textbook vis-viva and angular momentum, and a second, slot-ring
formulation of the stream written from the same algorithm.  No output of the
reference's generator ships with it and src/ran3.cu includes the CUDA header,
so the stream itself is parity unpinned against the reference's own numbers.
"""
import ctypes
import ctypes.util
import math

import numpy as np

LD = np.longdouble

PI = 3.14159265358979323846
PI2 = 6.28318530717958623200
GRAV = 4.0 * PI * PI
MSUN = 1.98892e30
MEARTH = 5.97e24
MGRID = 1.1

_M = 1000000000
_SEED_WORD = 161803398


_LIBM = ctypes.CDLL(ctypes.util.find_library("m"))
_LIBM.sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
_LIBM.sincos.restype = None


def sincos(x):
    """(sin x, cos x) from the C library's sincos."""
    s, c = ctypes.c_double(), ctypes.c_double()
    _LIBM.sincos(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def c_round(x):
    """C's round(): half away from zero."""
    a = abs(x)
    f = math.floor(a)
    r = f + (1.0 if a - f >= 0.5 else 0.0)
    return math.copysign(r, x)


def _c_rem(a, m):
    """C's % on integers (truncating toward zero)."""
    r = abs(a) % m
    return -r if a < 0 else r


class Ran3:
    """The reference's stream as the lagged list Y (see the module text)."""

    def __init__(self, seed=-2):
        mag = -seed if seed < 0 else seed
        words = [0] * 56
        prev, cur = _c_rem(_SEED_WORD - mag, _M), 1
        words[55] = prev
        for n in range(1, 55):
            words[(21 * n) % 55] = cur
            nxt = prev - cur
            if nxt < 0:
                nxt += _M
            prev, cur = cur, nxt
        self.y = words[1:56]          # Y[-54] .. Y[0]
        for _ in range(220):
            self._extend()
        self.count = 0

    def _extend(self):
        v = self.y[-55] - self.y[-24]
        if v < 0:
            v += _M
        self.y.append(v)
        if len(self.y) > 4096:
            del self.y[:-55]
        return v

    def word(self):
        self.count += 1
        return self._extend()

    def __call__(self):
        return self.word() * (1.0 / _M)


def getbod(mmax, C, ome):
    """(mu of the body, the next mmax) for the ladder's current mmax (a numpy
    longdouble)."""
    mmin = np.power(np.power(mmax, LD(ome)) - LD(ome / C), LD(1.0 / ome))
    m = LD(C) * (np.power(mmax, LD(1.0 + ome)) - np.power(mmin, LD(1.0 + ome))) / LD(1.0 + ome) / LD(MSUN) * LD(GRAV)
    return float(m), mmin


def coe2rv(M, m, a, e, i, w, o, nu, beta):
    p = a * (1.0 - e * e)
    mu = (M + m) * (1.0 - beta)
    sqrt = math.sqrt
    snu, cnu = sincos(nu)
    so, co = sincos(o)
    sw, cw = sincos(w)
    si, ci = sincos(i)
    xe = p * cnu / (1.0 + e * cnu)
    ye = p * snu / (1.0 + e * cnu)
    ze = 0.0
    vxe = -sqrt(mu / p) * snu
    vye = sqrt(mu / p) * (e + cnu)
    vze = 0.0
    x = xe * (co * cw - so * sw * ci) + ye * (-co * sw - so * cw * ci) + ze * (so * si)
    y = xe * (so * cw + co * sw * ci) + ye * (-so * sw + co * cw * ci) + ze * (-co * si)
    z = xe * sw * si + ye * cw * si + ze * ci
    vx = vxe * (co * cw - so * sw * ci) + vye * (-co * sw - so * cw * ci) + vze * (so * si)
    vy = vxe * (so * cw + co * sw * ci) + vye * (-so * sw + co * cw * ci) + vze * (-co * si)
    vz = vxe * sw * si + vye * cw * si + vze * ci
    return x, y, z, vx, vy, vz


def _normal(r):
    u1 = r()
    u2 = r()
    return math.sqrt(-2.0 * math.log(u1)) * math.cos(PI2 * u2)


def genorb_disk(d, Mstar, m, r, beta):
    while True:
        e = _normal(r) * d["sigmae"]
        if e < 0:
            e = -e
        if not e >= 1.0:
            break
    amin = d["Rin"] / (1.0 - e)
    amax = d["Rout"] / (1.0 + e)
    g = d["gammas"]
    a = math.pow(r() * (math.pow(amax, 2.0 + g) - math.pow(amin, 2.0 + g)) + math.pow(amin, 2.0 + g), 1.0 / (2.0 + g))
    i = _normal(r) * d["dhdr"]
    if i < 0:
        i = -i
    nu = PI2 * r()
    w = PI2 * r()
    o = PI2 * r()
    if beta >= 0.5:
        beta = 0.0
    return coe2rv(Mstar, m, a, e, i, w, o, nu, beta)


def genorb_blob(b, Mstar, m, r, beta):
    while True:
        a = b["a"] + _normal(r) * b["sigma"]
        if not a <= 0.0:
            break
    nu = b["nu"] + _normal(r) * b["sigma"] / a
    if beta >= 0.5:
        beta = 0.0
    x, y, z, vx, vy, vz = coe2rv(Mstar, m, a, b["ecc"], b["inc"], b["w"], b["o"], nu, beta)
    if b["collcloud"] == 0:
        voff = b["vkepp"] * 0.01 * math.sqrt(vx * vx + vy * vy + vz * vz)
        costheta = -1.0 + 2.0 * r()
        sintheta = math.sqrt(1.0 - costheta * costheta)
        gamma = PI2 * r()
        singamma, cosgamma = sincos(gamma)
        vx += voff * sintheta * cosgamma
        vy += voff * sintheta * singamma
        vz += voff * costheta
    return x, y, z, vx, vy, vz


def planets(mstar, records):
    """Star and planets (mu = G M, positions, velocities) from the star's mass
    [Msun] and the planets' records (i, a, e, nu, Omega, omega [deg], m [Msun])."""
    P = [[0.0] * 6 + [mstar * GRAV]]
    for rec in records:
        i, a, e, nu, o, w, m = rec
        i *= PI_DEG
        nu *= PI_DEG
        o *= PI_DEG
        w *= PI_DEG
        m *= GRAV
        P.append(list(coe2rv(P[0][6], m, a, e, i, w, o, nu, 0.0)) + [m])
    if len(P) > 1:
        X = Y = Z = M = 0.0
        for p in P:
            X += p[6] * p[0]
            Y += p[6] * p[1]
            Z += p[6] * p[2]
            M += p[6]
        X /= M
        Y /= M
        Z /= M
        for p in P:
            p[0] -= X
            p[1] -= Y
            p[2] -= Z
        vx = vy = vz = 0.0
        for p in P[1:]:
            vx -= p[6] * p[3]
            vy -= p[6] * p[4]
            vz -= p[6] * p[5]
        P[0][3] = vx / P[0][6]
        P[0][4] = vy / P[0][6]
        P[0][5] = vz / P[0][6]
    return P


PI_DEG = math.pi / 180.0


def disk_record(rec, rho):
    """add_disks: Rin Rout dh/dr sigma_e amax[km] amin[um] eta mtot[M_earth] gamma Ngrav Nsmall."""
    Rin, Rout, dhdr, sigmae, amax, amin, etaa, mtot, gammas, Ngrav, Nsmall = rec
    return dict(Rin=Rin, Rout=Rout, dhdr=math.atan(dhdr), sigmae=sigmae, amax=amax * 1000.0, amin=amin / 1.0e6,
                eta=(etaa + 2.0) / 3.0, Mtot=mtot * MEARTH,
                mmax=4.0 * math.pow(amax * 1000.0, 3.0) * PI * rho / 3.0,
                mmin=4.0 * math.pow(amin / 1.0e6, 3.0) * PI * rho / 3.0,
                gammas=gammas, Ngrav=int(Ngrav), Nsmall=int(Nsmall))


def blob_record(rec, rho):
    """add_blobs: a e i w Omega nu [deg] amax[km] amin[um] eta mtot Ngrav Nsmall kind value."""
    a, ecc, inc, w, o, nu, amax, amin, etaa, mtot, Ngrav, Nsmall, kind, val = rec
    cloud = kind in ("cloud", "Cloud")
    return dict(a=a, ecc=ecc, inc=inc * math.pi / 180.0, w=w * math.pi / 180.0, o=o * math.pi / 180.0,
                nu=nu * math.pi / 180.0, sigma=val if cloud else 0.0, amax=amax * 1000.0, amin=amin / 1.0e6,
                eta=(etaa + 2.0) / 3.0, Mtot=mtot * MEARTH,
                mmax=4.0 * math.pow(amax * 1000.0, 3.0) * PI * rho / 3.0,
                mmin=4.0 * math.pow(amin / 1.0e6, 3.0) * PI * rho / 3.0,
                Ngrav=int(Ngrav), Nsmall=int(Nsmall), collcloud=1 if cloud else 0, vkepp=0.0 if cloud else val)


def dust_grid(disks, blobs, rho, dmin, dmax):
    """detminmax, then confCdisks and confCblobs; returns (mmin, f)."""
    kappa = 4.0 * PI * rho / 3.0
    for d in disks + blobs:
        if d["amin"] < dmin:
            d["amin"] = dmin
            d["mmin"] = kappa * math.pow(d["amin"], 3.0)
        d["dmax"] = dmax if d["amax"] > dmax else d["amax"]
        d["dmmax"] = kappa * math.pow(d["dmax"], 3.0)
    if disks:
        amin, amax = disks[0]["amin"], disks[0]["dmax"]
    elif blobs:
        amin, amax = blobs[0]["amin"], blobs[0]["dmax"]
    else:
        amin, amax = dmin, dmax
    for d in disks[1:] + blobs:
        if d["amin"] < amin:
            amin = d["amin"]
        if d["dmax"] > amax:
            amax = d["dmax"]
    mmin = kappa * math.pow(amin, 3.0)
    mmax = kappa * math.pow(amax, 3.0)
    n_grid = int(math.log(mmax / mmin) / math.log(MGRID) + 1.0)
    f = math.log(mmax / mmin) / (float(n_grid) - 1.0)
    for d in disks + blobs:
        C = d["Mtot"]
        if d["eta"] == 1:
            C /= (d["mmax"] - d["mmin"])
        elif d["eta"] == 2:
            C /= math.log(d["mmax"] / d["mmin"])
        else:
            C /= (math.pow(d["mmax"], 2.0 - d["eta"]) - math.pow(d["mmin"], 2.0 - d["eta"])) / (2.0 - d["eta"])
        d["C"] = C
    return mmin, f


def setup(P, disks, blobs, rho, dmin, dmax, beta, seed=-2):
    """Grid::setup's particles: arrays x, y, z, vx, vy, vz, N, mu (float64) and
    id (int32) in the reference's order.  P from planets(); disks and blobs
    from disk_record / blob_record after dust_grid(); beta the float32 bin
    table."""
    mmin_g, f = dust_grid(disks, blobs, rho, dmin, dmax)
    r = Ran3(seed)
    rows = [tuple(p[:6]) + (1.0, p[6], -1) for p in P]
    Mstar = P[0][6]
    for kind, objs, orb in (("disk", disks, genorb_disk), ("blob", blobs, genorb_blob)):
        for d in objs:
            mmax = LD(d["mmax"])
            for _ in range(d["Ngrav"]):
                m, mmax = getbod(mmax, d["C"], 1.0 - d["eta"])
                rows.append(orb(d, Mstar, m, r, 0.0) + (1.0, m, -1))
    for kind, objs, orb in (("disk", disks, genorb_disk), ("blob", blobs, genorb_blob)):
        for d in objs:
            mx = mmin_g * math.exp(c_round(math.log(d["dmmax"] / mmin_g) / f) * f)
            mn = mmin_g * math.exp(c_round(math.log(d["mmin"] / mmin_g) / f) * f)
            ngrid = int(c_round(math.log(mx / mn) / math.log(MGRID) + 1))
            npg = int(c_round(float(d["Nsmall"]) / float(ngrid)))
            for l in range(ngrid):
                m = mn * math.exp(float(l) * f)
                size = math.pow(m * 3.0 / 4.0 / PI / rho, 1.0 / 3.0)
                bid = int(c_round(math.log(m / mmin_g) / f))
                if size < dmin:
                    bid = -1
                if size - dmax > 1e-9:
                    bid = -1
                lo = mn * math.exp((float(l) - 0.5) * f)
                hi = mn * math.exp((float(l) + 0.5) * f)
                m *= GRAV / MSUN
                eta = d["eta"]
                if eta == 0:
                    N = d["C"] * (hi - lo)
                elif eta == 1:
                    N = d["C"] * math.log(hi / lo)
                else:
                    N = d["C"] * (math.pow(hi, 1.0 - eta) - math.pow(lo, 1.0 - eta)) / (1.0 - eta)
                b = 0.0 if bid < 0 else float(np.float32(beta[bid]))
                for _ in range(npg):
                    rows.append(orb(d, Mstar, m, r, b) + (N / npg, m, bid))
    cols = list(zip(*rows))
    out = {k: np.array(cols[j], dtype=np.float64) for j, k in enumerate(("x", "y", "z", "vx", "vy", "vz", "N", "mu"))}
    out["id"] = np.array(cols[8], dtype=np.int32)
    return out
