"""Seeded synthetic disks: the five benchmark systems of this path.

Random numbers come from numpy's Philox4x64-10 counter-based generator
(Salmon et al., SC'11) keyed by the seed, so the same inputs are produced on
any host.  Orbits are sampled in elements and converted with coe2rv around the
star (src/functions.cu:6-46, as genorb_disk does at src/functions.cu:157-191);
mean anomalies are uniform and converted to true anomalies with Kepler's
equation.  Star and planets are shifted to the barycentre with recalc_COM
(src/read_in.cu:344-375) before any swarm body or grain is added, as
Grid::setup does (src/grid.cu:106-189).  Every generator takes size overrides
so tests can run the same system at reduced sizes.
"""
import math

import numpy as np

from .orbits import coe2rv, kepler_true_anomaly, recalc_com
from .system import FP32_FORCE, FP64, Params, Workload, make_state
from .imaging import OpticsTables, View
from .units import GRAV, LIGHT_AU, PC, kms_to_au_per_yr

DEG = math.pi / 180.0
VSW_DEFAULT_KMS = 450.0     # stellar-wind speed of the reference's example inputs [km/s]


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def _angles(rng, n):
    two_pi = 2.0 * math.pi
    return rng.uniform(0.0, two_pi, n), rng.uniform(0.0, two_pi, n), rng.uniform(0.0, two_pi, n)


def _rayleigh(rng, sigma, n, cap=None):
    v = rng.rayleigh(sigma, n)
    if cap is not None:
        bad = v >= cap
        while np.any(bad):
            v[bad] = rng.rayleigh(sigma, int(bad.sum()))
            bad = v >= cap
    return v


def _place_orbits(st, lo, hi, mu_star, m, a, e, inc, w, o, M, beta=0.0):
    nu = kepler_true_anomaly(M, e)
    x, y, z, vx, vy, vz = coe2rv(mu_star, m, a, e, inc, w, o, nu, beta)
    st.x[lo:hi], st.y[lo:hi], st.z[lo:hi] = x, y, z
    st.vx[lo:hi], st.vy[lo:hi], st.vz[lo:hi] = vx, vy, vz
    st.mu[lo:hi] = m


def _star_and_planets(st, mstar, planets, rng):
    """planets: list of (m [Msun], a, e, inc, w, o, M) with None for random angles."""
    st.mu[0] = mstar * GRAV
    for k, (m, a, e, inc, w, o, M) in enumerate(planets, start=1):
        w = rng.uniform(0, 2 * math.pi) if w is None else w
        o = rng.uniform(0, 2 * math.pi) if o is None else o
        M = rng.uniform(0, 2 * math.pi) if M is None else M
        nu = float(kepler_true_anomaly(np.array([M]), np.array([e]))[0])
        r = coe2rv(st.mu[0], m * GRAV, a, e, inc, w, o, nu, 0.0)
        st.x[k], st.y[k], st.z[k], st.vx[k], st.vy[k], st.vz[k] = [float(v) for v in r]
        st.mu[k] = m * GRAV
    npl = len(planets) + 1
    sl = slice(0, npl)
    x, y, z, vx, vy, vz = (a[sl].copy() for a in (st.x, st.y, st.z, st.vx, st.vy, st.vz))
    recalc_com(x, y, z, vx, vy, vz, st.mu[sl])
    st.x[sl], st.y[sl], st.z[sl], st.vx[sl], st.vy[sl], st.vz[sl] = x, y, z, vx, vy, vz
    return npl


def config1(n_dust=4096, seed=1001, nsteps=1000, dt=0.01):
    """Plain FP64, small planetary system: 1 Msun star, planets 9.5e-4 Msun at
    5.2 AU and 2.86e-4 Msun at 9.5 AU (e = 0.05, i = 0.5 deg), and massless
    test particles a~U(30,50), e~U(0,0.1), i~U(0,5 deg); beta = 0, no drag."""
    rng = _rng(seed)
    st = make_state(3, n_dust)
    npl = _star_and_planets(st, 1.0, [(9.5e-4, 5.2, 0.05, 0.5 * DEG, None, None, None),
                                      (2.86e-4, 9.5, 0.05, 0.5 * DEG, None, None, None)], rng)
    g = slice(npl, npl + n_dust)
    a = rng.uniform(30.0, 50.0, n_dust)
    e = rng.uniform(0.0, 0.1, n_dust)
    inc = rng.uniform(0.0, 5.0 * DEG, n_dust)
    w, o, M = _angles(rng, n_dust)
    _place_orbits(st, g.start, g.stop, st.mu[0], 0.0, a, e, inc, w, o, M)
    st.mu[g] = 0.0
    prm = Params(n_planets=npl, precision=FP64)
    return Workload("cfg1_planetary", st, prm, dt, nsteps, seed,
                    {"desc": "star + 2 planets + massless test particles, plain FP64"})


def config1_jacobi(n_dust=256, seed=1011, nsteps=1000, dt=0.01, mp=9.5e-4, ap=5.2):
    """Circular restricted three-body variant: one planet on an exactly circular
    barycentric orbit, test particles as in config1.  The Jacobi integral
    E - n Lz of each test particle is conserved by the exact dynamics."""
    rng = _rng(seed)
    st = make_state(2, n_dust)
    mus, mup = GRAV, mp * GRAV
    mtot = mus + mup
    n = math.sqrt(mtot / ap ** 3)
    phase = rng.uniform(0, 2 * math.pi)
    c, s = math.cos(phase), math.sin(phase)
    rs, rp = -mup / mtot * ap, mus / mtot * ap
    st.mu[0], st.mu[1] = mus, mup
    for k, r in ((0, rs), (1, rp)):
        st.x[k], st.y[k] = r * c, r * s
        st.vx[k], st.vy[k] = -n * r * s, n * r * c
    a = rng.uniform(30.0, 50.0, n_dust)
    e = rng.uniform(0.0, 0.1, n_dust)
    inc = rng.uniform(0.0, 5.0 * DEG, n_dust)
    w, o, M = _angles(rng, n_dust)
    _place_orbits(st, 2, 2 + n_dust, mus, 0.0, a, e, inc, w, o, M)
    st.mu[2:] = 0.0
    prm = Params(n_planets=2, precision=FP64)
    return Workload("cfg1J_jacobi", st, prm, dt, nsteps, seed,
                    {"mean_motion": n, "desc": "circular restricted three-body test"})


def jacobi_constants(st, mean_motion):
    """Jacobi integral E - n Lz of every test particle of a config1_jacobi
    state (star at row 0, planet at row 1, barycentre at rest at the origin)."""
    g = slice(2, None)
    v2 = st.vx[g] ** 2 + st.vy[g] ** 2 + st.vz[g] ** 2
    r0 = np.sqrt((st.x[g] - st.x[0]) ** 2 + (st.y[g] - st.y[0]) ** 2 + (st.z[g] - st.z[0]) ** 2)
    r1 = np.sqrt((st.x[g] - st.x[1]) ** 2 + (st.y[g] - st.y[1]) ** 2 + (st.z[g] - st.z[1]) ** 2)
    energy = 0.5 * v2 - st.mu[0] / r0 - st.mu[1] / r1
    return energy - mean_motion * (st.x[g] * st.vy[g] - st.y[g] * st.vx[g])


def config2(n_swarm=65536, seed=1002, nsteps=200, dt=0.02, precision=FP32_FORCE, eps=1e-3):
    """Self-gravitating swarm: 1 Msun star and n_swarm bodies with masses
    log-uniform in [1e-10, 1e-7] Msun, a~U(20,40), e~Rayleigh(0.02),
    i~Rayleigh(0.01 rad), Plummer eps = 1e-3 AU on swarm-swarm pairs."""
    rng = _rng(seed)
    st = make_state(1 + n_swarm, 0)
    st.mu[0] = GRAV
    m = np.exp(rng.uniform(math.log(1e-10), math.log(1e-7), n_swarm)) * GRAV
    a = rng.uniform(20.0, 40.0, n_swarm)
    e = _rayleigh(rng, 0.02, n_swarm, cap=0.9)
    inc = _rayleigh(rng, 0.01, n_swarm, cap=math.pi / 2)
    w, o, M = _angles(rng, n_swarm)
    _place_orbits(st, 1, 1 + n_swarm, st.mu[0], m, a, e, inc, w, o, M)
    prm = Params(n_planets=1, precision=precision, eps=eps, fp64_radius=10.0 * eps)
    return Workload("cfg2_swarm", st, prm, dt, nsteps, seed,
                    {"desc": "self-gravitating swarm"})


def grain_beta(s_um, rho=2.5, qpr=1.0, lum=0.1, mstar=0.5):
    """beta = 0.57 Q_pr (L/Lsun)/(M/Msun) / (rho[g cm^-3] s[um]) (src/functions.cu:375)."""
    return 0.57 * qpr * (lum / mstar) / (rho * s_um)


def config3(n_dust=262144, n_swarm=1024, seed=1003, nsteps=500, dt=0.05, precision=FP32_FORCE,
            eps=1e-3, vsw_kms=VSW_DEFAULT_KMS):
    """Dusty debris ring: 0.5 Msun, 0.1 Lsun star with SW/PR drag ratio 0.35; one
    3e-5 Msun planet at 20 AU (e = 0.1); n_swarm planetesimals of 1e-9 Msun;
    n_dust grains with n(s) ~ s^-3.5 on [1, 1000] um, rho = 2.5 g/cm^3,
    beta = 0.57/(rho s) (L/M), launched at a~U(30,40), e~U(0,0.05), i~U(0,3 deg).
    Radiation pressure, PR drag and wind drag on."""
    rng = _rng(seed)
    mstar = 0.5
    st = make_state(2 + n_swarm, n_dust)
    npl = _star_and_planets(st, mstar, [(3e-5, 20.0, 0.1, 0.0, None, None, None)], rng)
    # planetesimals: grain orbit distribution (a~U(30,40), e~U(0,0.05), i~U(0,3 deg))
    lo, hi = npl, npl + n_swarm
    a = rng.uniform(30.0, 40.0, n_swarm)
    e = rng.uniform(0.0, 0.05, n_swarm)
    inc = rng.uniform(0.0, 3.0 * DEG, n_swarm)
    w, o, M = _angles(rng, n_swarm)
    _place_orbits(st, lo, hi, st.mu[0], 1e-9 * GRAV, a, e, inc, w, o, M)
    # grains: inverse CDF of n(s) ~ s^-3.5 on [1, 1000] um
    u = rng.uniform(0.0, 1.0, n_dust)
    smin, smax, q = 1.0, 1000.0, -2.5
    s_um = (smin ** q + u * (smax ** q - smin ** q)) ** (1.0 / q)
    beta = grain_beta(s_um, lum=0.1, mstar=mstar)
    vsw = kms_to_au_per_yr(vsw_kms)
    betasw = 0.35 * beta * vsw / LIGHT_AU
    mgrain = (4.0 / 3.0) * math.pi * (s_um * 1e-4) ** 3 * 2.5 * 1e-3 / 1.98892e30 * GRAV
    a = rng.uniform(30.0, 40.0, n_dust)
    e = rng.uniform(0.0, 0.05, n_dust)
    inc = rng.uniform(0.0, 3.0 * DEG, n_dust)
    w, o, M = _angles(rng, n_dust)
    _place_orbits(st, hi, hi + n_dust, st.mu[0], mgrain, a, e, inc, w, o, M, beta)
    st.beta_p = beta
    st.betasw_p = betasw
    prm = Params(n_planets=npl, SWQ=1, precision=precision, eps=eps, Rmin=0.1, vsw=vsw,
                 fp64_radius=10.0 * eps)
    return Workload("cfg3_debris_ring", st, prm, dt, nsteps, seed,
                    {"s_um": s_um, "desc": "dusty debris ring with radiation, PR and wind drag"})


def config4(n_dust=32768, seed=1004, nsteps=400, dt=0.05, planet_m=1e-4, planet_a=25.0,
            planet_e=0.15):
    """Adjoint system: 1 Msun star, one planet (m, a, e; i = Omega = omega = nu = 0)
    and n_dust ring particles a~U(30,45), e~U(0,0.05), i~U(0,2 deg),
    beta~U(0,0.3) per particle; radiation pressure and PR drag, no wind."""
    rng = _rng(seed)
    st = make_state(2, n_dust)
    npl = _star_and_planets(st, 1.0, [(planet_m, planet_a, planet_e, 0.0, 0.0, 0.0, 0.0)], rng)
    a = rng.uniform(30.0, 45.0, n_dust)
    e = rng.uniform(0.0, 0.05, n_dust)
    inc = rng.uniform(0.0, 2.0 * DEG, n_dust)
    w, o, M = _angles(rng, n_dust)
    beta = rng.uniform(0.0, 0.3, n_dust)
    _place_orbits(st, npl, npl + n_dust, st.mu[0], 0.0, a, e, inc, w, o, M, beta)
    st.mu[npl:] = 0.0
    st.beta_p = beta
    st.betasw_p = np.zeros(n_dust)
    prm = Params(n_planets=npl, precision=FP64)
    return Workload("cfg4_adjoint", st, prm, dt, nsteps, seed,
                    {"planet": (planet_m, planet_a, planet_e), "ckpt_every": 20,
                     "desc": "adjoint ring: radiation + PR drag"})


def config5(n_swarm=1048576, seed=1005, nsteps=20, dt=0.02, precision=FP32_FORCE, eps=5e-4):
    """Large self-gravitating disk: 1 Msun star and n_swarm bodies, dN/dm ~ m^-1.8
    on [1e-12, 1e-8] Msun, a~U(10,100) (Sigma ~ 1/r), e, i ~ Rayleigh(0.03),
    eps = 5e-4 AU."""
    rng = _rng(seed)
    st = make_state(1 + n_swarm, 0)
    st.mu[0] = GRAV
    u = rng.uniform(0.0, 1.0, n_swarm)
    p = -0.8   # integral exponent of m^-1.8
    mlo, mhi = 1e-12, 1e-8
    m = (mlo ** p + u * (mhi ** p - mlo ** p)) ** (1.0 / p) * GRAV
    a = rng.uniform(10.0, 100.0, n_swarm)
    e = _rayleigh(rng, 0.03, n_swarm, cap=0.9)
    inc = _rayleigh(rng, 0.03, n_swarm, cap=math.pi / 2)
    w, o, M = _angles(rng, n_swarm)
    _place_orbits(st, 1, 1 + n_swarm, st.mu[0], m, a, e, inc, w, o, M)
    prm = Params(n_planets=1, precision=precision, eps=eps, fp64_radius=10.0 * eps)
    return Workload("cfg5_large_disk", st, prm, dt, nsteps, seed,
                    {"desc": "large self-gravitating disk"})


CONFIGS = {1: config1, 2: config2, 3: config3, 4: config4, 5: config5}


def ring_imaging(w, n_bins=64, n_wav=48, n_theta=41, n_r=100, teff=3600.0, dist_pc=10.0, xsize=256, ysize=256,
                 fov=100.0, incl_deg=30.0, pa_deg=20.0, rot_deg=10.0, Rmin=0.1, Rmax=200.0):
    """Optical tables and an observer for imaging the debris ring of config3
    (which must carry its grain sizes, w.meta["s_um"]): the grains get size-bin
    ids on n_bins log bins over [1, 1000] um; smooth synthetic optics with
    x = 2 pi s / lambda: Qabs = x / (1 + x), Qsca = 0.5 x^2 / (1 + x^2), a
    forward-peaked phase function exp(a mu) with a = 2 x / (1 + x); a
    blackbody star of teff at dist_pc (radius from L = 0.1 Lsun); grain
    temperatures 278.3 K (L/Lsun)^(1/4) (R/AU)^(-1/2) (1 + 1/x_10um)^(1/5) on
    n_r radii from Rmin to Rmax.  Sets the grains' ids in w.state; returns the
    tables and the view."""
    st = w.state
    s_um = np.asarray(w.meta["s_um"], dtype=np.float64)
    smin, smax = 1.0, 1000.0
    b = np.floor(np.log(s_um / smin) / np.log(smax / smin) * n_bins).astype(np.int64)
    st.id[st.n_grav:] = np.clip(b, 0, n_bins - 1).astype(np.int32)
    size_um = smin * (smax / smin) ** ((np.arange(n_bins) + 0.5) / n_bins)
    wav = np.geomspace(0.3e-6, 1.0e-3, n_wav)
    x = 2.0 * math.pi * size_um[:, None] * 1e-6 / wav[None, :]
    qabs = x / (1.0 + x)
    qsca = 0.5 * x * x / (1.0 + x * x)
    mu = np.linspace(-1.0, 1.0, n_theta)
    a = 2.0 * x / (1.0 + x)
    qpf = np.exp(a[:, None, :] * mu[None, :, None])
    h, c, k = 6.626068e-34, 299792458.0, 1.3806503e-23
    lum = 0.1 * 3.827e26
    sigma = 2.0 * math.pi ** 5 * k ** 4 / (15.0 * h ** 3 * c ** 2)
    rstar = math.sqrt(lum / (4.0 * math.pi * sigma * teff ** 4))
    dist = dist_pc * PC
    influx = math.pi * 2 * h * c * c / (wav ** 5 * (np.exp(h * c / (wav * k * teff)) - 1.0)) * (rstar / dist) ** 2
    R = Rmin * (Rmax / Rmin) ** (np.arange(n_r) / (n_r - 1.0))
    x10 = 2.0 * math.pi * size_um / 10.0
    Td = 278.3 * 0.1 ** 0.25 * R[None, :] ** -0.5 * (1.0 + 1.0 / x10[:, None]) ** 0.2
    t = OpticsTables(w=wav, in_=influx, size=size_um * 1e-6, Qabs=qabs, Qsca=qsca, Qpfunc=qpf, Td=Td, Rmin=Rmin,
                     Rmax=Rmax)
    v = View(dist=dist, rot=rot_deg * DEG, incl=incl_deg * DEG, PA=pa_deg * DEG, fov=fov, xsize=xsize, ysize=ysize)
    return t, v
