"""CPU tests of a .param run's set-up (libddyn_host.so, ddyn_setup.cpp): the
optical data and spectrum readers, the dust size grid, the grain coefficient
tables with the sublimation radius, and the particle generation of disks and
blobs.  What pins them: closed forms (Stefan-Boltzmann luminosity, the
blackbody equilibrium distance, beta = 0.57 Q_pr (L/M)/(s rho), exact log-log
interpolation of power laws, the mass ladder's total mass, the bin weights'
telescoping sum) and the distributions the mathematics fixes (Kolmogorov-
Smirnov tests of the orbital-element laws).  The particles themselves, bit
for bit against the oracle's restatement of the reference's set-up, are in
tests/test_setup_stream.py.
"""
import math

import numpy as np
import pytest
from scipy import stats

from diskdyn_amd import hostio, orbits
from diskdyn_amd.units import AU, GRAV, MSUN, RSUN
from oracle import setup_oracle as S
from tests import opticslib as OL

MEARTH = 5.97e24
WAV_UM = np.geomspace(0.05, 2000.0, 800)
SIZES_UM = np.geomspace(0.1, 2000.0, 14)


@pytest.fixture
def files(tmp_path):
    spec = tmp_path / "star.dat"
    OL.write_spectrum(spec, 5800.0)
    comp = tmp_path / "bb.dat"
    OL.gray_table(comp, SIZES_UM, WAV_UM, qabs=1.0, qsca=0.0, n_theta=21)
    return tmp_path, comp, spec


def _setup(tmp_path, comp, spec, disks, blobs="", physics="yes yes no"):
    p = OL.setup_param(tmp_path / "run.param", comp, spec, disks, blobs, physics)
    cfg = hostio.read_param(p)
    opt = hostio.Optics(cfg.comp.decode(), cfg.specfile.decode(), cfg)
    hostio.set_dust_grid(cfg, opt)
    return cfg, opt


DISK = "30.0 50.0 0.05 0.05 100.0 1.0 3.67 0.1 -1.0 0 {nsmall}"


def test_optics_reader_shapes_and_errors(files):
    tmp, comp, spec = files
    cfg, opt = _setup(tmp, comp, spec, DISK.format(nsmall=0))
    info = opt.info()
    assert info["n_wav"] == len(WAV_UM) and info["n_theta"] == 21 and info["n_sizes"] == len(SIZES_UM)
    assert abs(info["dmin"] - 0.1e-6) < 1e-20 and abs(info["dmax"] - 2000e-6) < 1e-16
    with pytest.raises(hostio.HostIOError) as e:
        hostio.Optics(str(tmp / "missing.dat"), str(spec), cfg)
    assert e.value.code == 39
    bad = tmp / "bad.dat"
    bad.write_text(comp.read_text().replace("# 2000\n", "# 2000\n1 2\n", 1)[:-200])
    with pytest.raises(hostio.HostIOError) as e:
        hostio.Optics(str(bad), str(spec), cfg)
    assert e.value.code == 42


def test_dust_grid_clips_to_optical_range(files):
    """detminmax: the disk's 1 um .. 100 km range clipped to the optical 0.1 um
    .. 2 mm for grains; the grid spans it in mass steps of about 1.1."""
    tmp, comp, spec = files
    cfg, _ = _setup(tmp, comp, spec, DISK.format(nsmall=0))
    rho = 2500.0
    kappa = 4 * math.pi * rho / 3
    d = cfg.disks[0]
    assert abs(d.amin - 1e-6) < 1e-21 and abs(d.dmax - 2000e-6) < 1e-18
    assert abs(d.dmmax / (kappa * 2000e-6 ** 3) - 1) < 1e-14
    assert abs(cfg.amin - 1e-6) < 1e-21 and abs(cfg.amax - 2000e-6) < 1e-18
    ratio = (2000.0 / 1.0) ** 3
    assert cfg.n_grid == int(math.log(ratio) / math.log(1.1) + 1.0)
    assert abs(cfg.f - math.log(ratio) / (cfg.n_grid - 1)) < 1e-15
    # the normalisation: Mtot = C int m^(1 - eta) dm from the grains' mmin to the disk's mmax
    eta = (3.67 + 2.0) / 3.0
    C = d.C
    M = C * (d.mmax ** (2 - eta) - d.mmin ** (2 - eta)) / (2 - eta)
    assert abs(M / (0.1 * MEARTH) - 1) < 1e-12


def test_luminosity_beta_and_sublimation_radius_of_blackbody_grains(files):
    """Qabs = 1, Qsca = 0: L = 4 pi R^2 sigma T^4 (Stefan-Boltzmann), Q_pr = 1 so
    beta = 0.57 (L/Lsun)/(M/Msun) / (s[um] rho[g/cm^3]), and the sublimation
    radius is the blackbody equilibrium distance 0.5 R (T / 2000 K)^2."""
    tmp, comp, spec = files
    cfg, opt = _setup(tmp, comp, spec, DISK.format(nsmall=0))
    t = hostio.grain_tables(cfg, opt)
    L_sb = 4 * math.pi * RSUN ** 2 * OL.SIGMA_SB * 5800.0 ** 4 / OL.LSUN_W
    assert abs(t["lum"] / L_sb - 1) < 2e-4
    beta_cf = 0.57 * t["lum"] / 1.0 / (t["size"] * 1e6) / 2.5
    assert np.allclose(t["beta"], beta_cf, rtol=2e-7, atol=0)
    rmin_cf = 0.5 * RSUN * (5800.0 / 2000.0) ** 2 / AU
    assert abs(t["Rmin"] / rmin_cf - 1) < 2e-4
    # sizes: the grid amin exp(i f_a) up to amax
    assert abs(t["size"][0] - cfg.amin) < 1e-21 and abs(t["size"][-1] / cfg.amax - 1) < 1e-13
    assert np.allclose(np.diff(np.log(t["size"])), math.log(cfg.amax / cfg.amin) / (cfg.n_grid - 1), rtol=1e-9)


def test_scattering_asymmetry_and_wind_drag(tmp_path):
    """Gray grains with Qabs = 0.6, Qsca = 0.8 and phase function 1 + 0.9 mu:
    Q_pr = Qabs + Qsca (1 - g) with g the trapezoid rule's 0.3 (1 + h^2/2);
    beta_sw = 3 Mdot Cd v_sw / (32 pi G M rho s) with Cd = 2, in SI units
    (G M from the gravitational constant, an independent route to the
    reference's 4 pi^2 AU^3 / yr^2 to about 1e-4); q = kqe s 1e9."""
    spec = tmp_path / "star.dat"
    OL.write_spectrum(spec, 5800.0)
    comp = tmp_path / "gray.dat"
    OL.gray_table(comp, SIZES_UM, WAV_UM, qabs=0.6, qsca=0.8, g_lin=0.9, n_theta=21)
    cfg, opt = _setup(tmp_path, comp, spec, DISK.format(nsmall=0))
    t = hostio.grain_tables(cfg, opt)
    g = OL.trapezoid_g(0.9, 21)
    qpr = 0.6 + 0.8 * (1 - g)
    beta_cf = 0.57 * qpr * t["lum"] / (t["size"] * 1e6) / 2.5
    assert np.allclose(t["beta"], beta_cf, rtol=3e-7, atol=0)
    mdot = 2e-14 * MSUN / 3.1556926e7           # kg/s
    gm = 6.67384e-11 * MSUN
    bsw = 3 * mdot * 2.0 * 450e3 / (32 * math.pi * gm * 2500.0 * t["size"])
    assert np.allclose(t["betasw"], bsw, rtol=5e-4)
    assert np.allclose(t["q"], 27.5 / 3.1556926e7 * t["size"] * 1e9, rtol=1e-6)
    # wind drag off: beta_sw = 0; radiation pressure off: beta = 0
    cfg2, opt2 = _setup(tmp_path, comp, spec, DISK.format(nsmall=0), physics="no no no")
    t2 = hostio.grain_tables(cfg2, opt2)
    assert not t2["beta"].any() and not t2["betasw"].any()


def test_size_interpolation_is_exact_for_power_laws(tmp_path):
    """Qabs ~ s^1.5 on the table's sizes: the log-log interpolation to the grid
    gives s^1.5 at every grid size, seen through beta of Qsca = 0 grains
    (beta ~ Q_pr / s)."""
    spec = tmp_path / "star.dat"
    OL.write_spectrum(spec, 5800.0)
    comp = tmp_path / "pow.dat"
    OL.gray_table(comp, SIZES_UM, WAV_UM, qabs=0.01, qsca=0.0, qabs_pow=1.5, n_theta=21)
    cfg, opt = _setup(tmp_path, comp, spec, DISK.format(nsmall=0))
    t = hostio.grain_tables(cfg, opt)
    q = 0.01 * (t["size"] * 1e6 / SIZES_UM[0]) ** 1.5
    beta_cf = 0.57 * q * t["lum"] / (t["size"] * 1e6) / 2.5
    assert np.allclose(t["beta"], beta_cf, rtol=5e-7, atol=0)


def _elements(st, lo, hi, mu_star, beta=None):
    """Orbital elements (a, e, i, omega, Omega, nu) of rows [lo, hi) in the
    frame the particles are generated in: Kepler orbits of mu_star + m about
    the origin, the barycentre of the star and planets (the reference places
    them there too, after recalc_COM, src/functions.cu:188); the textbook
    state-to-elements map (angular momentum, node and eccentricity vectors)."""
    r = st.pos()[lo:hi]
    v = st.vel()[lo:hi]
    b = np.zeros(hi - lo) if beta is None else beta
    mu = (mu_star + st.mu[lo:hi]) * (1 - b)
    R = np.linalg.norm(r, axis=1)
    h = np.cross(r, v)
    H = np.linalg.norm(h, axis=1)
    ev = (np.sum(v * v, 1) - mu / R)[:, None] * r / mu[:, None] - (np.sum(r * v, 1) / mu)[:, None] * v
    e = np.linalg.norm(ev, axis=1)
    a = 1.0 / (2.0 / R - np.sum(v * v, 1) / mu)
    inc = np.arccos(np.clip(h[:, 2] / H, -1, 1))
    n = np.stack([-h[:, 1], h[:, 0], np.zeros(len(R))], axis=1)
    N = np.linalg.norm(n, axis=1)
    o = np.arctan2(n[:, 1], n[:, 0])
    w = np.arccos(np.clip(np.sum(n * ev, 1) / (N * e), -1, 1))
    w = np.where(ev[:, 2] < 0, 2 * math.pi - w, w)
    nu = np.arccos(np.clip(np.sum(ev * r, 1) / (e * R), -1, 1))
    nu = np.where(np.sum(r * v, 1) < 0, 2 * math.pi - nu, nu)
    return a, e, inc, w, o, nu


def _gen(tmp_path, disks, blobs=""):
    spec = tmp_path / "star.dat"
    OL.write_spectrum(spec, 5800.0)
    comp = tmp_path / "bb.dat"
    OL.gray_table(comp, SIZES_UM, WAV_UM, n_theta=21)
    cfg, opt = _setup(tmp_path, comp, spec, disks, blobs)
    t = hostio.grain_tables(cfg, opt)
    st = hostio.generate(cfg, opt, t["beta"])
    return cfg, opt, t, st


def test_disk_bodies_follow_the_mass_ladder_and_orbit_laws(tmp_path):
    """30,000 gravitating bodies of one disk (Rin 20, Rout 60 AU, dh/dr 0.05,
    sigma_e 0.1, Sigma ~ r^-1): masses down the ladder of dN/dm = C m^-eta
    with one body per interval (total mass in closed form), e ~ |N(0, 0.1)|,
    i ~ |N(0, atan 0.05)|, a^(2 + gamma) uniform between Rin/(1-e) and
    Rout/(1+e), angles uniform.  The star (the reference's generation
    coordinates) sits at the origin, the planets after recalc_COM."""
    n = 30000
    cfg, opt, t, st = _gen(tmp_path, f"20.0 60.0 0.05 0.1 100.0 1.0 3.67 1000.0 -1.0 {n} 0")
    npl = cfg.n_planets
    assert st.n_grav == npl + n and st.n_dust == 0
    m = st.mu[npl:] * MSUN / GRAV
    assert np.all(np.diff(m) < 0)
    d = cfg.disks[0]
    eta = d.eta
    ome = 1 - eta
    mN = (d.mmax ** ome - n * ome / d.C) ** (1 / ome)
    Mtot = d.C * (d.mmax ** (2 - eta) - mN ** (2 - eta)) / (2 - eta)
    # The ladder in the reference's precision (ome / C and 1 / ome rounded to
    # double, the powers in long double) ends 5e-11 off the closed form's mN;
    # the 30,000 bodies span only a narrow mass range below mmax, so the total
    # C (mmax^(2-eta) - mN^(2-eta)) / (2-eta) cancels and that becomes 1.1e-7
    # (6.9e-11 with every step in long double).
    assert abs(m.sum() / Mtot - 1) < 1e-6
    # and the masses are the reference's getbod ladder, bit for bit
    ladder, top = [], S.LD(d.mmax)
    for _ in range(n):
        mu, top = S.getbod(top, d.C, 1.0 - d.eta)
        ladder.append(mu)
    assert np.array_equal(st.mu[npl:], np.array(ladder))
    a, e, inc, w, o, nu = _elements(st, npl, npl + n, st.mu[0])
    assert stats.kstest(e / 0.1, stats.halfnorm.cdf).pvalue > 1e-3
    assert stats.kstest(inc / math.atan(0.05), stats.halfnorm.cdf).pvalue > 1e-3
    p = 2.0 + d.gammas
    amin, amax = 20.0 / (1 - e), 60.0 / (1 + e)
    u = (a ** p - amin ** p) / (amax ** p - amin ** p)
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    for ang in (w, o, nu):
        assert stats.kstest(np.mod(ang, 2 * math.pi) / (2 * math.pi), "uniform").pvalue > 1e-3
    assert np.array_equal(st.id[npl:], np.full(n, -1, dtype=np.int32))


def test_disk_grains_bins_weights_and_radiation_pressure(tmp_path):
    """Grains of one disk: Npg tracers per mass bin of ratio exp(f), ids of the
    bins, N per tracer the bin's share of C int m^-eta dm (the weights of all
    tracers telescope to the integral over the whole range), orbits launched
    under gravity reduced by (1 - beta) of the bin (vis-viva with mu (1 - beta))."""
    cfg, opt, t, st = _gen(tmp_path, "30.0 40.0 0.05 0.05 100.0 1.0 3.67 0.1 -1.0 0 20000")
    ng, nd = st.n_grav, st.n_dust
    d = cfg.disks[0]
    mx = cfg.mmin * math.exp(round(math.log(d.dmmax / cfg.mmin) / cfg.f) * cfg.f)
    mn = cfg.mmin * math.exp(round(math.log(d.mmin / cfg.mmin) / cfg.f) * cfg.f)
    ngrid = round(math.log(mx / mn) / math.log(1.1) + 1)
    npg = round(20000 / ngrid)
    assert nd == ngrid * npg
    ids = st.id[ng:]
    assert np.array_equal(np.unique(ids), np.arange(ids.min(), ids.min() + ngrid))
    assert np.all(np.bincount(ids - ids.min()) == npg)
    m_kg = st.mu[ng:] * MSUN / GRAV
    assert np.allclose(m_kg, cfg.mmin * np.exp(ids * cfg.f), rtol=1e-12)
    lo, hi = mn * math.exp(-0.5 * cfg.f), mn * math.exp((ngrid - 0.5) * cfg.f)
    eta = d.eta
    total = d.C * (hi ** (1 - eta) - lo ** (1 - eta)) / (1 - eta)
    assert abs(st.N[ng:].sum() / total - 1) < 1e-10
    beta = np.asarray(t["beta"], float)[ids]
    beta = np.where(beta >= 0.5, 0.0, beta)
    r = np.linalg.norm(st.pos()[ng:], axis=1)
    v2 = np.sum(st.vel()[ng:] ** 2, axis=1)
    a, e, *_ = _elements(st, ng, ng + nd, st.mu[0], beta)
    vis = (st.mu[0] + st.mu[ng:]) * (1 - beta) * (2 / r - 1 / a)
    assert np.allclose(v2, vis, rtol=1e-10)
    # a between Rin/(1-e) and Rout/(1+e): the orbit stays inside [Rin, Rout] while the
    # interval is not empty (e < (Rout - Rin)/(Rout + Rin)); beyond, the reference's
    # formula samples the inverted interval, and so does this one
    lo_a, hi_a = 30.0 / (1 - e), 40.0 / (1 + e)
    assert np.all((a >= np.minimum(lo_a, hi_a) * (1 - 1e-12)) & (a <= np.maximum(lo_a, hi_a) * (1 + 1e-12)))
    ok = e < 10.0 / 70.0
    assert np.all((a[ok] * (1 - e[ok]) >= 30.0 - 1e-9) & (a[ok] * (1 + e[ok]) <= 40.0 + 1e-9))


def test_blob_cloud_and_collision_kick(tmp_path):
    """Blobs: a ~ N(a_blob, sigma) (> 0) with the blob's e, i, Omega, omega and
    nu ~ N(nu_blob, sigma / a); a collision blob adds a kick of vkepp % of the
    orbital speed in an isotropic direction (|dv| exact, cos(theta) uniform)."""
    blobs = ("40.0 0.1 5.0 10.0 20.0 30.0 100.0 1.0 3.67 0.1 4000 0 cloud 2.0\n"
             "40.0 0.1 5.0 10.0 20.0 30.0 100.0 1.0 3.67 0.1 4000 0 collision 3.0")
    cfg, opt, t, st = _gen(tmp_path, "", blobs)
    npl = cfg.n_planets
    for b, lo in ((cfg.blobs[0], npl), (cfg.blobs[1], npl + 4000)):
        rows = slice(lo, lo + 4000)
        r = st.pos()[rows]
        if b.collcloud == 1:
            a, e, inc, w, o, nu = _elements(st, lo, lo + 4000, st.mu[0])
            assert np.allclose(e, 0.1, rtol=1e-9) and np.allclose(inc, b.inc, rtol=1e-9)
            assert stats.kstest((a - 40.0) / 2.0, "norm").pvalue > 1e-3
            dnu = np.angle(np.exp(1j * (nu - b.nu)))
            assert stats.kstest(dnu * a / 2.0, "norm").pvalue > 1e-3
        else:
            # the unkicked orbit through the same position: its true anomaly from the
            # position in the orbital plane, its a from r = a (1 - e^2) / (1 + e cos nu)
            x0 = np.array(orbits.coe2rv(GRAV, 0.0, 1.0, 0.0, b.inc, b.w, b.o, 0.0, 0.0)[:3]).ravel()
            y0 = np.array(orbits.coe2rv(GRAV, 0.0, 1.0, 0.0, b.inc, b.w, b.o, math.pi / 2, 0.0)[:3]).ravel()
            nu = np.arctan2(r @ y0, r @ x0)
            rn = np.linalg.norm(r, axis=1)
            a = rn * (1 + 0.1 * np.cos(nu)) / (1 - 0.01)
            m = st.mu[rows]
            v0 = np.stack(orbits.coe2rv(st.mu[0], m, a, 0.1, b.inc, b.w, b.o, nu, 0.0)[3:], axis=1)
            dv = st.vel()[rows] - v0
            sp = np.linalg.norm(v0, axis=1)
            assert np.allclose(np.linalg.norm(dv, axis=1), 0.03 * sp, rtol=1e-7)
            ct = dv[:, 2] / np.linalg.norm(dv, axis=1)
            assert stats.kstest((ct + 1) / 2, "uniform").pvalue > 1e-3


def test_generation_is_seeded_and_counts_match(tmp_path):
    cfg, opt, t, st = _gen(tmp_path, "30.0 40.0 0.05 0.05 100.0 1.0 3.67 0.1 -1.0 100 3000")
    st2 = hostio.generate(cfg, opt, t["beta"])
    assert np.array_equal(st.x, st2.x) and np.array_equal(st.vz, st2.vz)
    cfg.seed = 12345
    st3 = hostio.generate(cfg, opt, t["beta"])
    assert not np.array_equal(st.x[cfg.n_planets:], st3.x[cfg.n_planets:])
    assert np.array_equal(st.x[:cfg.n_planets], st3.x[:cfg.n_planets])


def test_generation_refuses_laws_it_cannot_draw(tmp_path):
    """A blob whose semi-major axes can never be positive (a < 0, no spread)
    and a disk whose eccentricities are almost never below 1 end the
    generation with error 54 instead of looping."""
    for disks, blobs in (("", "-5.0 0.1 5.0 10.0 20.0 30.0 100.0 1.0 3.67 0.1 10 0 cloud 0.0"),
                         ("30.0 40.0 0.05 1.0e12 100.0 1.0 3.67 0.1 -1.0 10 0", "")):
        spec, comp = tmp_path / "star.dat", tmp_path / "bb.dat"
        OL.write_spectrum(spec, 5800.0)
        OL.gray_table(comp, SIZES_UM, WAV_UM, n_theta=21)
        cfg, opt = _setup(tmp_path, comp, spec, disks, blobs)
        t = hostio.grain_tables(cfg, opt)
        with pytest.raises(hostio.HostIOError) as e:
            hostio.generate(cfg, opt, t["beta"])
        assert e.value.code == 54


def test_restart_ids_from_masses(tmp_path):
    """Grid::read_in recomputes each body's bin from its mass; beyond the grid
    (gravitating bodies) and massless bodies get -1."""
    cfg, opt, t, st = _gen(tmp_path, "30.0 40.0 0.05 0.05 100.0 1.0 3.67 0.1 -1.0 50 3000")
    ids = hostio.restart_ids(cfg, st)
    assert np.array_equal(ids[st.n_grav:], st.id[st.n_grav:])
    assert np.all(ids[:st.n_grav] == -1)
