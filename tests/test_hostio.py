"""CPU tests of the host envelope (libddyn_host.so): .param reader, det_step,
data-file writer and reader.  The .param text below is written for these
tests; the reference's example files are parsed too when present."""
import glob
import math
import os

import numpy as np
import pytest

from diskdyn_amd import hostio, orbits, synth
from diskdyn_amd.units import AU, GRAV, MSUN, RSUN, YEAR

PARAM = """# test system
1
0
no
5.0 10.0 10.0
20.0
200
# star: M sptype Teff R B P mloss vsw
1.5 A0V 9500 1.7 1.0 2.5 2e-14 450
10.0 100.0 30.0 60.0 0.0 128 128
0.5 1.0 "10.0"
Library
yes yes no
begin_planets
1.0 5.0 0.1 30.0 40.0 50.0 1e-3
2.0 12.0 0.05 200.0 10.0 80.0 3e-4
end_planets
2.5 optical/astrosil.dat 1.0
begin_disks
30.0 50.0 0.05 0.05 1.0 1.0 -3.5 0.1 -1.0 0 0
end_disks
begin_blobs
40.0 0.1 5.0 10.0 20.0 30.0 1.0 1.0 -3.5 0.01 0 0 cloud 0.5
end_blobs
teststub
"""


@pytest.fixture
def param_file(tmp_path):
    p = tmp_path / "sys.param"
    p.write_text(PARAM)
    return p


def test_read_param_records_and_units(param_file):
    cfg = hostio.read_param(param_file)
    assert cfg.verb == 1 and cfg.gl == 0 and cfg.n_step == 200
    assert (cfg.t_write, cfg.t_fits, cfg.t_sed, cfg.t_end) == (5.0, 10.0, 10.0, 20.0)
    assert cfg.n_planets == 3 and cfg.n_disks == 1 and cfg.n_blobs == 1
    assert cfg.pm[0] == 1.5 * GRAV and cfg.pm[1] == 1e-3 * GRAV
    assert abs(cfg.rstar - 1.7 * RSUN / AU) < 1e-18
    assert abs(cfg.Bstar - 1.0e-4 * YEAR * YEAR / MSUN) < 1e-30
    assert abs(cfg.Wz - 365.2422 / 2.5) < 1e-12
    assert cfg.vsw == 450e3 and cfg.SWQ == 1 and cfg.FLQ == 0 and cfg.radQ == 1
    assert cfg.teff == 9500.0 and cfg.specfile == b"Kurucz/Kurucz_9500.dat"
    assert cfg.rho == 2500.0 and cfg.comp == b"optical/astrosil.dat"
    assert cfg.nwavs == 3 and abs(cfg.wavs[2] - 10e-6) < 1e-18
    assert cfg.name_stub == b"teststub"
    assert abs(cfg.disks[0].dhdr - math.atan(0.05)) < 1e-15 and cfg.disks[0].eta == (-3.5 + 2) / 3
    b = cfg.blobs[0]
    assert b.collcloud == 1 and b.sigma == 0.5 and abs(b.inc - 5 * math.pi / 180) < 1e-15
    # planets: coe2rv with (i, a, e, nu, Omega, omega, m) then the barycentre correction
    n = cfg.n_planets
    mu = np.array(cfg.pm[:n])
    P = np.array([[cfg.pvx[j], cfg.pvy[j], cfg.pvz[j]] for j in range(n)])
    X = np.array([[cfg.px[j], cfg.py[j], cfg.pz[j]] for j in range(n)])
    assert np.abs((mu[:, None] * P).sum(0)).max() < 1e-15
    assert np.abs((mu[:, None] * X).sum(0)).max() < 1e-13
    rel = X[1] - X[0]
    a, e, inc = orbits.rv2coe(mu[0], mu[1], *rel, *(P[1] - 0.0 * P[0]))
    assert abs(inc - math.pi / 180) < 1e-9 and abs(e - 0.1) < 1e-9 and abs(a - 5.0) < 1e-9
    # the planet's velocity is its coe2rv velocity (recalc_COM leaves planets' velocities)
    r = orbits.coe2rv(mu[0], mu[1], 5.0, 0.1, math.pi / 180, 50 * math.pi / 180, 40 * math.pi / 180,
                      30 * math.pi / 180)
    assert np.allclose(P[1], [float(r[3]), float(r[4]), float(r[5])], rtol=1e-14)


@pytest.mark.parametrize("bad,code", [
    ("1\n0\nmaybe\n", 8), ("-1\n", 4), ("1\n0\nno\n0 1 1\n", 9),
    ("1\n0\nno\n1 1 1\n10\n100\n1.0 G2V 5800 1.0 1.0 25\n", 13),
    ("1\n0\nno\n1 1 1\n10\n100\n1.0 G2V 5800 1.0 1.0 25 1e-14 450\n10 100 0 0 0 8 8\n0.5\nLibrary\nyes no perhaps\n", 31),
    ("1\n0\nno\n1 1 1\n10\n100\n", 17),
])
def test_read_param_errors(tmp_path, bad, code):
    p = tmp_path / "bad.param"
    p.write_text(bad)
    with pytest.raises(hostio.HostIOError) as e:
        hostio.read_param(p)
    assert e.value.code == code


def test_missing_file():
    with pytest.raises(hostio.HostIOError) as e:
        hostio.read_param("/nonexistent/file.param")
    assert e.value.code == 3


def test_det_step_rule(param_file):
    cfg = hostio.read_param(param_file)
    st = hostio.state_from_config(cfg)
    dt = hostio.det_step(cfg, st)
    cands = []
    for i in range(1, st.n_grav):
        r3 = ((st.x[i] - st.x[0]) ** 2 + (st.y[i] - st.y[0]) ** 2 + (st.z[i] - st.z[0]) ** 2) ** 1.5
        cands.append(2 * math.pi * math.sqrt(r3 / (st.mu[0] + st.mu[i])) / cfg.n_step)
    cands.append(2 * math.pi * math.sqrt(30.0 / st.mu[0]) / cfg.n_step)            # disk Rin term
    cands.append(2 * math.pi * math.sqrt((40.0 * 0.9 - 0.5) / st.mu[0]) / cfg.n_step)   # blob term
    assert abs(dt - min(cands)) <= 1e-15 * dt


def test_datafile_round_trip(param_file, tmp_path):
    cfg = hostio.read_param(param_file)
    w = synth.config3(n_dust=64, n_swarm=5)
    st = w.state
    st.N[st.n_grav + 3] = 0.0
    st.id[st.n_grav:] = np.arange(st.n_dust, dtype=np.int32) % 7
    cfg.n_planets = w.params.n_planets
    run = hostio.RunInfo(time=12.5, dt=0.05, tlastd=10.0, tlastf=10.0, tlasts=10.0, Rmax=38.0,
                         n_mass=st.n_grav - cfg.n_planets, n_dust=st.n_dust, n_all=st.n_all)
    path = tmp_path / "out.dat"
    hostio.write_datafile(path, cfg, st, run)
    lines = path.read_text().splitlines()
    assert lines[1] == "#Calculated_with_DiskDyn_Version 1.01"
    assert lines[106].startswith("#Data:") and lines[107] == "#Planets:"
    assert lines[108 + cfg.n_planets] == "#Bodies:"
    assert any(l.startswith("## ") for l in lines[110:])        # grains beyond Rmax are marked
    cfg2, run2, st2 = hostio.read_datafile(path)
    assert run2.n_all == st.n_all and run2.n_dust == st.n_dust and run2.time == 12.5
    assert cfg2.n_planets == cfg.n_planets and cfg2.t_end == cfg.t_end and cfg2.SWQ == 1
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        a, b = getattr(st2, k), getattr(st, k)
        assert np.allclose(a, b, rtol=6e-9, atol=1e-300)            # %.8e
    assert np.allclose(st2.mu[st.n_grav:], st.mu[st.n_grav:], rtol=6e-9)
    assert np.array_equal(st2.N, st.N) and np.array_equal(st2.id[st.n_grav:], st.id[st.n_grav:])
    assert np.allclose(st2.beta_p, st.beta_p, rtol=6e-9)


def test_version_mismatch(param_file, tmp_path):
    cfg = hostio.read_param(param_file)
    st = hostio.state_from_config(cfg)
    run = hostio.RunInfo(n_all=st.n_all, n_dust=0)
    path = tmp_path / "v.dat"
    hostio.write_datafile(path, cfg, st, run)
    txt = path.read_text().replace("Version 1.01", "Version 0.99")
    path.write_text(txt)
    with pytest.raises(hostio.HostIOError) as e:
        hostio.read_datafile(path)
    assert e.value.code == 55


REF_TESTS = "/root/reference/tests"


@pytest.mark.skipif(not os.path.isdir(REF_TESTS), reason="reference example inputs not present")
def test_reference_example_params_parse():
    files = sorted(glob.glob(os.path.join(REF_TESTS, "*.param")))
    assert files
    for f in files:
        cfg = hostio.read_param(f)
        assert cfg.n_planets >= 1 and cfg.t_end > 0 and cfg.n_step > 0
        n = cfg.n_planets
        mu = np.array(cfg.pm[:n])
        P = np.array([[cfg.pvx[j], cfg.pvy[j], cfg.pvz[j]] for j in range(n)])
        assert np.abs((mu[:, None] * P).sum(0)).max() <= 1e-12 * max(1.0, np.abs(mu[:, None] * P).max())


def test_restart_with_wind_drag_and_grains_is_refused(tmp_path):
    """A data file carries beta per grain but no beta_sw or charge tables, so
    ddyn -m refuses a restart with grains and wind drag or the Lorentz force
    on, instead of running them at zero (the refusal comes before any GPU
    call)."""
    import subprocess

    from diskdyn_amd import hostio
    from diskdyn_amd.system import make_state
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    ddyn = os.path.join(root, "diskdyn_amd", "ddyn")
    p = tmp_path / "sys.param"
    p.write_text(PARAM)
    cfg = hostio.read_param(p)
    for swq, flq in ((1, 0), (0, 1)):
        cfg.SWQ, cfg.FLQ = swq, flq
        st = make_state(cfg.n_planets, 3)
        for j in range(cfg.n_planets):
            st.x[j], st.y[j], st.vy[j], st.mu[j] = cfg.px[j], cfg.py[j], cfg.pvy[j], cfg.pm[j]
        st.x[cfg.n_planets:] = [30.0, 31.0, 32.0]
        st.vy[cfg.n_planets:] = 1.1
        st.beta_p = np.full(3, 0.1)
        run = hostio.RunInfo()
        run.n_mass, run.n_dust, run.n_all, run.Rmax = 0, 3, st.n_all, 1e300
        f = tmp_path / f"grains_{swq}{flq}.dat"
        hostio.write_datafile(f, cfg, st, run)
        r = subprocess.run([ddyn, "-m", str(f), "--quiet"], cwd=tmp_path, capture_output=True, text=True,
                           timeout=60)
        print(r.returncode, r.stderr)
        assert r.returncode == 2 and "beta_sw and charge tables" in r.stderr


def test_datafile_writer_matches_reference_layout_byte_for_byte(tmp_path):
    """The library's writer against tests/golden/datafile_tiny.dat, the text of
    the reference's fprintf sequence (src/grid.cu:441-673) for a tiny system
    with a disk, a blob, a massive body and grains beyond the outer radius,
    inside the star and without a size bin (restated in
    tests/golden/make_datafile_golden.py)."""
    from diskdyn_amd.system import DiskState
    from tests.golden import make_datafile_golden as G
    c, run_d, s, beta_bins = G.tiny_system()
    cfg = hostio.Config()
    for k in ("verb", "gpuid", "gl", "t_gl", "t_write", "t_sed", "t_fits", "t_end", "n_step", "teff", "rstar",
              "Bstar", "Wz", "mloss", "vsw", "radQ", "SWQ", "FLQ", "rho", "kqe", "dist", "fov", "PA", "inc",
              "longitude", "xsize", "ysize", "amin", "amax", "mmin", "mmax", "n_grid", "f", "n_wav", "n_theta",
              "n_planets", "n_disks", "n_blobs", "seed"):
        setattr(cfg, k, c[k])
    for k in ("sptype", "name_stub", "specfile", "comp"):
        setattr(cfg, k, c[k].encode())
    cfg.nwavs = len(c["wavs"])
    for i, v in enumerate(c["wavs"]):
        cfg.wavs[i] = v
    for k, v in c["disk"].items():
        setattr(cfg.disks[0], k, v)
    for k, v in c["blob"].items():
        setattr(cfg.blobs[0], k, v)
    for k, arr in (("x", "px"), ("y", "py"), ("z", "pz"), ("vx", "pvx"), ("vy", "pvy"), ("vz", "pvz"), ("m", "pm")):
        for i, v in enumerate(c["planets"][k]):
            getattr(cfg, arr)[i] = v
    run = hostio.RunInfo()
    for k, v in run_d.items():
        setattr(run, k, v)
    n = len(s["x"])
    run.n_all = n
    st = DiskState(n - run.n_dust, run.n_dust, *[np.ascontiguousarray(s[k]) for k in ("x", "y", "z", "vx", "vy", "vz",
                                                                                      "mu", "N")], s["id"])
    out = tmp_path / "tiny.dat"
    hostio.write_datafile(out, cfg, st, run, beta_bins=beta_bins)
    got = out.read_bytes()
    want = open(G.FIXTURE, "rb").read()
    assert want == G.expected_text(c, run_d, s, beta_bins).encode()
    if got != want:
        for i, (a, b) in enumerate(zip(got.decode().splitlines(), want.decode().splitlines())):
            assert a == b, f"line {i + 1}: {a!r} != {b!r}"
    assert got == want
